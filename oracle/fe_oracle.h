/*
 * fe_oracle.h -- CPU restatement of FixedEffects.jl v3.1.0's residualization
 * path.  TEST INFRASTRUCTURE ONLY: this library is the checker that the HIP
 * library (libfehip) is compared against, and the CPU baseline that bench.py
 * times beside it.  Nothing in the product links or calls it.
 *
 * Every function restates one reference function; the citation names the
 * file:line it follows (paths relative to the reference repository root).
 * Conventions follow the reference: refs are 1-based int64 (Julia Vector{Int}),
 * a NULL interaction means UnitWeights, a NULL weight vector means UnitWeights.
 *
 * Two precisions are built from one source (fe_oracle_impl.inc):
 *   _d : T = Float64 (double_precision = true, the default for :cpu)
 *   _f : T = Float32 (double_precision = false)
 * Interactions and weights are always passed as double, as the reference
 * keeps them in the user's element type and promotes inside each kernel.
 */
#ifndef FE_ORACLE_H
#define FE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* bucketize_refs  (ext/MetalExt.jl:40-64): stable counting sort of
 * observations by level.  perm (N) and offsets (n+1) are 1-based, exactly as
 * the reference builds them. */
void oracle_bucketize(const int64_t *refs, int64_t N, int64_t n,
                      int64_t *perm, int64_t *offsets);

typedef struct oracle_solver_d oracle_solver_d;
typedef struct oracle_solver_f oracle_solver_f;

/* LSMR outcome (src/utils/lsmr.jl:218-221) plus the stop code. */
typedef struct {
    int64_t iters;   /* ConvergenceHistory.mvps (iteration count) */
    int32_t istop;   /* 0..7 as in src/utils/lsmr.jl:208-215 */
    int32_t converged;
    double normr, normAr, normA, condA, normx;
} oracle_lsmr_info;

#define ORACLE_DECLARE(S, T)                                                   \
    /* scale!  (src/SolverCPU.jl:78-89) */                                     \
    void oracle_scale_##S(T *scale, int64_t n, const int64_t *refs,            \
                          const double *z, const double *w, int64_t N);        \
    /* cache!  (src/SolverCPU.jl:91-95), threaded over 100k-row chunks */      \
    void oracle_cache_##S(T *cache, const int64_t *refs, const double *z,      \
                          const double *w, const T *scale, int64_t N,          \
                          int nthreads);                                       \
    /* gather!  (src/SolverCPU.jl:21-27): fecoef[refs[i]] += a*y[i]*cache[i] */\
    void oracle_gather_##S(T *fecoef, const int64_t *refs, T alpha,            \
                           const T *y, const T *cache, int64_t N);             \
    /* scatter!  (src/SolverCPU.jl:29-34): y[i] += a*fecoef[refs[i]]*cache[i] */\
    void oracle_scatter_##S(T *y, T alpha, const T *fecoef,                    \
                            const int64_t *refs, const T *cache, int64_t N,    \
                            int nthreads);                                     \
    /* AbstractFixedEffectSolver{T}(fes, w, Val{:cpu})  (src/SolverCPU.jl:55-64) \
     * inputs are copied; refs[g] 1-based in 1..n[g] */                        \
    oracle_solver_##S *oracle_create_##S(int64_t N, int32_t G,                 \
                                         const int64_t *const *refs,           \
                                         const int64_t *n,                     \
                                         const double *const *interaction,     \
                                         const double *weights, int nthreads); \
    void oracle_destroy_##S(oracle_solver_##S *s);                             \
    /* update_weights!  (src/SolverCPU.jl:67-76); w NULL = UnitWeights */      \
    void oracle_update_weights_##S(oracle_solver_##S *s, const double *w);     \
    /* read-outs of the solver state for the checks */                         \
    void oracle_get_scale_##S(const oracle_solver_##S *s, int32_t g, T *out);  \
    void oracle_get_cache_##S(const oracle_solver_##S *s, int32_t g, T *out);  \
    /* mul!(y, A, x, a, b)  (src/AbstractFixedEffectLinearMap.jl:37-44);       \
     * x is the p-length concatenation of the G coefficient blocks */          \
    void oracle_mul_A_##S(oracle_solver_##S *s, T *y, const T *x, T a, T b);   \
    /* mul!(x, A', y, a, b)  (src/AbstractFixedEffectLinearMap.jl:26-35) */    \
    void oracle_mul_At_##S(oracle_solver_##S *s, T *x, const T *y, T a, T b);  \
    /* lsmr!(x, A, b, v, h, hbar; atol, btol, conlim, maxiter)                 \
     * (src/utils/lsmr.jl:46-222); x (p) in/out, b (N) overwritten */          \
    void oracle_lsmr_##S(oracle_solver_##S *s, T *x, T *b, double atol,        \
                         double btol, double conlim, int64_t maxiter,          \
                         oracle_lsmr_info *info);                              \
    /* inner solve_residuals!(r, feM; tol, maxiter)                            \
     * (src/AbstractFixedEffectSolver.jl:49-68); y (N, double) in/out */       \
    void oracle_solve_residuals_##S(oracle_solver_##S *s, double *y,           \
                                    double tol, int64_t maxiter,               \
                                    int64_t *iter, int32_t *converged);        \
    /* inner solve_coefficients!(r, feM; tol, maxiter)                         \
     * (src/AbstractFixedEffectSolver.jl:147-160): writes the unscaled,       \
     * not yet normalized coefficients (p) and the LSMR iteration count.      \
     * The returned iteration count is div(mvps, 2) as the reference returns. */\
    void oracle_solve_coefficients_##S(oracle_solver_##S *s, const double *y,  \
                                       double tol, int64_t maxiter,            \
                                       double *coef, int64_t *iter,            \
                                       int32_t *converged);

ORACLE_DECLARE(d, double)
ORACLE_DECLARE(f, float)

/* normalize!  (src/FixedEffect.jl:106-111) with components (:68-98) and
 * rescale! (:113-135).  coef is the p-length concatenation of the G blocks
 * (double, as collected by solve_coefficients!), normalized in place.
 * unit_interaction[g] != 0 marks the FEs whose interaction is UnitWeights. */
void oracle_normalize(double *coef, int64_t N, int32_t G,
                      const int64_t *const *refs, const int64_t *n,
                      const int32_t *unit_interaction);

/* components  (src/FixedEffect.jl:68-98): label[g][k] = 0-based component id
 * of level k (1-based k stored at index k-1), -1 for levels never observed.
 * Returns the number of components. */
int64_t oracle_components(int64_t N, int32_t G, const int64_t *const *refs,
                          const int64_t *n, int64_t *const *label);

/* full  (src/FixedEffect.jl:137-140): out[g*N + i] = coef_g[refs_g[i]] */
void oracle_full(const double *coef, int64_t N, int32_t G,
                 const int64_t *const *refs, const int64_t *n, double *out);

#ifdef __cplusplus
}
#endif
#endif
