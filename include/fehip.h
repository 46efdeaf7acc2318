/*
 * fehip.h -- C-ABI of libfehip.so, the MI355X (gfx950) fixed-effect
 * residualization solver.
 *
 * The library replaces, for one new backend (`method = :rocm`), the part of
 * FixedEffects.jl v3.1.0 that a backend plugs in behind
 * `AbstractFixedEffectSolver{T}(fes, w, Val{method})`:
 *   - FixedEffectLinearMap mul! and its adjoint (src/AbstractFixedEffectLinearMap.jl:26-44,
 *     kernels src/SolverCPU.jl:21-34),
 *   - the scale / cache construction (src/SolverCPU.jl:67-95),
 *   - the LSMR loop (src/utils/lsmr.jl:46-222),
 *   - the inner solve_residuals! / solve_coefficients! methods
 *     (src/AbstractFixedEffectSolver.jl:49-68, :74-97, :147-160),
 *   - normalize! / components / full (src/FixedEffect.jl:57-140) and the
 *     grouping of raw integer keys (src/FixedEffect.jl:18-21), as optional
 *     device steps.
 * The outer keyword methods stay with the caller.
 *
 * Conventions
 *   - refs are 1-based, as FixedEffect.refs (src/FixedEffect.jl:7-21); the
 *     create call validates 1 <= ref <= n (the reference leaves an out-of-range
 *     ref undefined under @inbounds).
 *   - interaction[g] == NULL means UnitWeights (no interaction); weights ==
 *     NULL means UnitWeights.
 *   - every host buffer is copied; the caller keeps ownership.
 *   - calls block until results are in the caller's buffers.  A handle is not
 *     thread-safe; each handle owns its own HIP stream.
 *   - every entry point returns 0 on success or a negative FEHIP_E* code;
 *     fehip_last_error() then holds a message (thread-local).
 */
#ifndef FEHIP_H
#define FEHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fehip_solver fehip_t;
typedef struct fehip_mlp fehip_mlp_t;  /* the warm-start network */

/* status codes */
#define FEHIP_OK 0
#define FEHIP_EDIM -1  /* inconsistent lengths / sizes */
#define FEHIP_EREF -2  /* a ref outside 1..n */
#define FEHIP_EHIP -3  /* HIP runtime error */
#define FEHIP_EOOM -4  /* device allocation failed */
#define FEHIP_EARG -5  /* invalid argument */

/* arithmetic of the solve */
#define FEHIP_FP64 0   /* Float64 everywhere: double_precision = true */
#define FEHIP_FP32 1   /* Float32 everywhere, LSMR scalars included:
                          double_precision = false (src/utils/lsmr.jl:59) */
#define FEHIP_MIXED 2  /* Float32 streams with segment sums beyond Float32
                          (FP64 sums of exact products, compensated pairs),
                          Float64 LSMR scalars, Float64 refinement sweeps */

/* Builds the solver: uploads refs, builds the per-FE CSR permutations on the
 * device, computes scales and caches.
 * Replaces AbstractFixedEffectSolver{T}(fes, w, Val{:cpu})
 * (src/SolverCPU.jl:55-64) and FixedEffectLinearMapCPU{T}(fes)
 * (src/SolverCPU.jl:13-17).
 *   refs[g]        : N int64 refs of FE g (Julia Vector{Int}), 1-based
 *   nlevels[g]     : fe.n
 *   interaction[g] : N doubles or NULL (UnitWeights)
 *   weights        : N doubles or NULL (UnitWeights) */
int fehip_create(fehip_t **out, int64_t N, int32_t G,
                 const int64_t *const *refs, const int64_t *nlevels,
                 const double *const *interaction, const double *weights,
                 int32_t precision);

/* Same as fehip_create with int32 refs (1-based). */
int fehip_create_i32(fehip_t **out, int64_t N, int32_t G,
                     const int32_t *const *refs, const int64_t *nlevels,
                     const double *const *interaction, const double *weights,
                     int32_t precision);

/* update_weights!(feM, w)  (src/SolverCPU.jl:67-76): recomputes scales and
 * caches; w NULL = UnitWeights. */
int fehip_update_weights(fehip_t *h, const double *weights);

/* Residualizes K columns in one batched solve.
 * Replaces the inner solve_residuals!(r, feM; tol, maxiter)
 * (src/AbstractFixedEffectSolver.jl:49-68) and, for K > 1, its multi-column
 * loop (:74-97): each column gets its own LSMR scalars and stopping test.
 *   cols[c]       : host column c (N doubles), overwritten with residuals
 *   iters[c]      : lsmr iterations + 1, or 1 when G == 1 (:57-61)
 *   converged[c]  : istop not in (3, 6, 7) (src/utils/lsmr.jl:218)
 * maxiter is the outer maxiter; LSMR runs with maxiter - 1 (:59). */
int fehip_solve_residuals(fehip_t *h, double *const *cols, int32_t K,
                          double tol, int64_t maxiter, int64_t *iters,
                          uint8_t *converged);

/* Same as fehip_solve_residuals with the K columns already in device memory:
 * column c at dY + c * ldY (device pointer, doubles), overwritten in place.
 * This is the entry point that keeps inputs resident in HBM. */
int fehip_solve_residuals_device(fehip_t *h, double *dY, int32_t K,
                                 int64_t ldY, double tol, int64_t maxiter,
                                 int64_t *iters, uint8_t *converged);

/* Many panels in one call (multi-panel form of the multi-column
 * solve_residuals!, src/AbstractFixedEffectSolver.jl:74-97, over several
 * fixed-effect structures): panel i residualizes K[i] <= 32 device-resident
 * columns dY[i] (column c at dY[i] + c*ldY[i]) with its own handle hs[i].
 * All panels' LSMR iterations are in flight together, each on its handle's
 * stream; the host advances whichever panel's stop flag has come back.
 * Results are those of fehip_solve_residuals_device on each handle alone.
 * iters[i] / converged[i] (K[i] entries each) may be NULL.  Handles must be
 * distinct. */
int fehip_solve_residuals_panels(fehip_t *const *hs, int32_t npanels, double *const *dY, const int32_t *K,
                                 const int64_t *ldY, double tol, int64_t maxiter, int64_t *const *iters,
                                 uint8_t *const *converged);

/* fehip_solve_residuals_panels on host columns: cols[i][c] is column c (N_i
 * doubles, overwritten with its residuals) of panel i; the library stages
 * each panel's columns in device memory of its own, one H2D and one D2H per
 * column.  Same multi-column solve_residuals! over several fixed-effect
 * structures (src/AbstractFixedEffectSolver.jl:74-97); K[i] <= 32. */
int fehip_solve_residuals_panels_host(fehip_t *const *hs, int32_t npanels, double *const *const *cols,
                                      const int32_t *K, double tol, int64_t maxiter, int64_t *const *iters,
                                      uint8_t *const *converged);

/* Replaces the inner solve_coefficients!(r, feM; tol, maxiter)
 * (src/AbstractFixedEffectSolver.jl:147-160) up to the collect at :158:
 * LSMR from x0 = 0 with maxiter, then x .*= scale.  coef[g] receives the
 * nlevels[g] unnormalized coefficients of FE g; the caller applies
 * normalize! and full (src/FixedEffect.jl:106-140), or calls
 * fehip_solve_coefficients_full, which does both on the device.
 *   iters     : div(lsmr iterations, 2) (:159)
 *   converged : istop not in (3, 6, 7) */
int fehip_solve_coefficients(fehip_t *h, const double *y, double tol,
                             int64_t maxiter, double *const *coef,
                             int64_t *iters, uint8_t *converged);

/* The whole inner solve_coefficients!(r, feM) (src/AbstractFixedEffectSolver.jl
 * :147-160) on the device: the solve above, then normalize! (src/FixedEffect.jl
 * :106-135: per connected component of the FEs without interaction, every FE
 * but the first demeaned, the means added to the first) and full (:137-140).
 *   coef : optional, G host arrays of nlevels[g] doubles (normalized)
 *   obs  : optional, G host arrays of N doubles, obs[g][i] = coef[g][refs[g][i]]
 *          (the reference's return value); NULL entries are skipped
 * Components are computed once per handle (they depend on the refs only). */
int fehip_solve_coefficients_full(fehip_t *h, const double *y, double tol,
                                  int64_t maxiter, double *const *coef,
                                  double *const *obs, int64_t *iters,
                                  uint8_t *converged);

/* LSMR algorithm of the residual solves (no reference counterpart).
 * FEHIP_LSMR (default): the reference's LSMR on the whole map A = [A_1 ... A_G]
 * (src/utils/lsmr.jl), iteration counts comparable with the reference.
 * FEHIP_LSMR_PROJECTED: LSMR on the map with the FE of most levels (s)
 * eliminated (Frisch-Waugh-Lovell): (I - A_s A_s') A_o over the other FEs'
 * coefficients, right-hand side (I - A_s A_s') b, where A_s' and A_s are the
 * exact within-level demeaning (the columns of A_s are orthonormal).  The
 * least-squares residual is the same; LSMR needs fewer iterations (a smaller
 * condition number) and the same stopping rule applies to the projected
 * problem.  Applies to fehip_solve_residuals*, panels, the _x0/_learned
 * forms and, in FEHIP_FP64 / FEHIP_FP32, to the coefficient solves: there
 * the eliminated FE's coefficients are the least-squares fit of the rest,
 * A_s'(b - A_o x_o), and iters is div(projected LSMR iterations, 2). */
#define FEHIP_LSMR 0
#define FEHIP_LSMR_PROJECTED 1
int fehip_set_algorithm(fehip_t *h, int32_t algorithm);

/* FixedEffect(keys...) grouping (src/FixedEffect.jl:18-21, which calls
 * GroupedArray(keys..., sort = nothing)): dense 1-based refs for the distinct
 * combinations of ncols int64 key columns (keys[c]: N host values), numbered
 * in ascending lexicographic key order (GroupedArray's numbering for
 * sort = nothing is unspecified; residuals do not depend on it).
 *   refs    : N host int64, out
 *   ngroups : number of levels (the FixedEffect's n), out
 * Needs no handle; N < 2^31. */
int fehip_group(int64_t N, int32_t ncols, const int64_t *const *keys,
                int64_t *refs, int64_t *ngroups);

/* No counterpart in the reference: fehip_group for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same numbering as fehip_group, on
 * device-resident data and without any allocation: its scratch is a fixed
 * device array of the library (one per GPU, held until the process ends).
 *   dkeys[c] : N device int64 keys of column c (host array of device pointers)
 *   drefs    : N device int64, out: 1-based refs
 *   ngroups  : host int64, out
 * N <= 2^20 (FEHIP_EDIM beyond).  Runs on the null stream and returns when
 * drefs is written.  Calls on one GPU share the scratch: calls that overlap
 * on one GPU (other threads) race. */
int fehip_group_device_noalloc(int64_t N, int32_t ncols, const int64_t *const *dkeys, int64_t *drefs,
                               int64_t *ngroups);

/* normalize!(fecoefs, fes) (src/FixedEffect.jl:106-111) of G host arrays in
 * place, for the handle's fixed effects. */
int fehip_normalize(fehip_t *h, double *const *coef);

/* Initial iterate (no reference counterpart as an argument; the reference's
 * lsmr! takes it through its x argument, src/utils/lsmr.jl:42-43, and
 * solve_residuals! always passes x0 = A'b, src/AbstractFixedEffectSolver.jl:56).
 * Same as fehip_solve_residuals with LSMR started from the unscaled
 * coefficients coef0[c] (sum(nlevels) doubles, FE blocks concatenated) for
 * column c; coef0 == NULL or coef0[c] == NULL keeps the reference start. */
int fehip_solve_residuals_x0(fehip_t *h, double *const *cols, int32_t K,
                             double tol, int64_t maxiter,
                             const double *const *coef0, int64_t *iters,
                             uint8_t *converged);

/* fehip_solve_coefficients with LSMR started from coef0 (unscaled,
 * sum(nlevels) doubles) instead of 0; coef0 == NULL is the reference start. */
int fehip_solve_coefficients_x0(fehip_t *h, const double *y, double tol,
                                int64_t maxiter, const double *coef0,
                                double *const *coef, int64_t *iters,
                                uint8_t *converged);

/* Warm-start features of one column y (N doubles), computed on the device:
 * per level (sum(nlevels) rows x 16 floats) log1p(count), weighted group mean
 * / sigma, log1p(degree in the FE graph), one alternating-projection pass /
 * sigma, interaction flag, log(sum w z^2)/2, one-hot FE index, empty flag.
 * coef_ap (optional) receives the alternating-projection coefficients
 * (unscaled), sigma (optional) the rms of sqrt(w) .* y. */
int fehip_features(fehip_t *h, const double *y, float *feats, double *coef_ap,
                   double *sigma);

/* Residualizes K columns with LSMR started from the network's prediction:
 * coefficients = alternating-projection pass + net(features) * sigma.
 * Otherwise as fehip_solve_residuals (host columns) /
 * fehip_solve_residuals_device (device columns). */
int fehip_solve_residuals_learned(fehip_t *h, double *const *cols, int32_t K,
                                  double tol, int64_t maxiter,
                                  fehip_mlp_t *net, int64_t *iters,
                                  uint8_t *converged);
int fehip_solve_residuals_learned_device(fehip_t *h, double *dY, int32_t K,
                                         int64_t ldY, double tol,
                                         int64_t maxiter, fehip_mlp_t *net,
                                         int64_t *iters, uint8_t *converged);

/* The warm-start network (no reference counterpart): 16 -> 128 -> 128 -> 128
 * -> 1 perceptron with ReLU, bf16 MFMA products with FP32 accumulation on the
 * GPU, FP32 master weights, Adam.  Parameters are one flat FP32 vector:
 * W0 [16x128], b0 [128], W1 [128x128], b1, W2 [128x128], b2, W3 [128x16]
 * (column 0 live), b3 [16]. */
int fehip_mlp_create(fehip_mlp_t **out, uint64_t seed);
int fehip_mlp_nparams(const fehip_mlp_t *m, int64_t *n);
int fehip_mlp_get_params(fehip_mlp_t *m, float *out);
int fehip_mlp_set_params(fehip_mlp_t *m, const float *in);
/* Gradient of 0.5 * mean (net(X) - t)^2 over n rows (X: n x 16 floats), in
 * the parameter layout; loss (optional) receives the loss.  No update. */
int fehip_mlp_gradient(fehip_mlp_t *m, const float *X, const float *t,
                       int64_t n, float *grad, double *loss);
/* Adam training over `epochs` shuffled passes of mini-batches. */
int fehip_mlp_fit(fehip_mlp_t *m, const float *X, const float *t, int64_t n,
                  int32_t epochs, int32_t batch, float lr, uint64_t seed,
                  double *loss_per_epoch);
int fehip_mlp_predict(fehip_mlp_t *m, const float *X, int64_t n, float *y);
void fehip_mlp_destroy(fehip_mlp_t *m);

/* Test hook: the CSR bucketing of FE g over the caller's observation order,
 * as bucketize_refs builds it (ext/MetalExt.jl:40-64) but 0-based:
 * perm (N) lists observations grouped by level, ascending within a level;
 * offsets (nlevels+1) delimit the levels.  Built on the device. */
int fehip_get_csr(fehip_t *h, int32_t g, int32_t *perm, int32_t *offsets);

/* Test hook: scale vector of FE g (nlevels[g] doubles), as scale!
 * computes it (src/SolverCPU.jl:78-89), widened to double. */
int fehip_get_scale(fehip_t *h, int32_t g, double *scale);

/* Test hook: cache vector of FE g in the caller's observation order
 * (N doubles), as cache! computes it (src/SolverCPU.jl:91-95). */
int fehip_get_cache(fehip_t *h, int32_t g, double *cache);

/* Test hook: one application of the linear map on K = 1 column, in the
 * caller's observation order and with the handle's precision.
 *   fehip_mul_A : y <- b*y + a*A*x   (src/AbstractFixedEffectLinearMap.jl:37-44)
 *   fehip_mul_At: x <- b*x + a*A'*y  (src/AbstractFixedEffectLinearMap.jl:26-35)
 * x is the concatenation of the G coefficient blocks (sum(nlevels) doubles). */
int fehip_mul_A(fehip_t *h, double *y, const double *x, double a, double b);
int fehip_mul_At(fehip_t *h, double *x, const double *y, double a, double b);

/* Test hook: LSMR scalars of column `col` of the last solve on this handle
 * (the ConvergenceHistory and estimates of src/utils/lsmr.jl:170-192, 218-220):
 * out[0] iterations, out[1] istop, out[2] normr, out[3] normAr, out[4] normA,
 * out[5] condA, out[6] normx, out[7] alpha, out[8] beta. */
int fehip_lsmr_info(fehip_t *h, int32_t col, double *out);

/* Test hook (no reference counterpart): the launch plan of K columns in one
 * storage precision (fp32_storage 0: the FP64 streams of FEHIP_FP64 and of
 * FEHIP_MIXED's FP64 phase; 1: the FP32 streams of FEHIP_FP32 and of
 * FEHIP_MIXED's FP32 phase), built if not yet there.
 *   out[0] 1 if the few-level FEs are summed inside the fused pass
 *   out[1] positions per fused-pass tile      out[2] fused-pass tiles
 *   out[3] tiles of the D'u passes (all FEs)  out[4] fix-up items
 *   out[5] few-level accumulator slots        out[6] pieces of the piece form (0: CSR form)
 *   out[7] workgroups of the CSR D'u launch   out[8] columns per lane of the fused pass
 * out holds 9 elements. */
int fehip_plan_info(fehip_t *h, int32_t K, int32_t fp32_storage, int64_t *out);

/* Measurement hook (no reference counterpart): average device time in ms of
 * each launch of one LSMR iteration on K columns.  `reps` iterations of a
 * real LSMR start on filler data (b = 1, x = 0, zero tolerances) run eagerly
 * in their order, with a HIP event between consecutive launches on the
 * handle's stream, in the launch configuration the solve uses -- so each
 * kernel meets the cache state the iteration leaves it:
 *   out_ms[0] one pass: u = A v - alpha u (scatter!, src/SolverCPU.jl:29-34)
 *             with the raw D'u sums of the FE with most levels and of the
 *             few-level FEs (gather!, src/SolverCPU.jl:21-27)
 *   out_ms[1] beta = ||u|| finalize
 *   out_ms[2] D'u of the other many-level FEs (CSR-ordered gathers); by
 *             default one launch whose first blocks also do the fix-up of
 *             the first FE's levels and of the few-level FEs (v = A'u/beta -
 *             beta v)
 *   out_ms[3] D'u of few-level FEs when they do not fit the pass (else 0)
 *   out_ms[4] fix-up after the CSR pass: segments longer than a CSR tile,
 *             levels without observations (about 0 when there are none);
 *             with FEHIP_SEG_FIX=0, the whole fix-up
 *   out_ms[5] alpha = ||v|| and plane-rotation finalize
 *   out_ms[6] LSMR vector update (src/utils/lsmr.jl:138-142, :192)
 *   out_ms[7] two events recorded back to back with no launch between
 *             them: the timing's own overhead, included in every slot above
 *             (the ||x|| and stopping-rule step runs inside the next
 *             iteration's beta finalize, out_ms[1]) */
int fehip_time_kernels(fehip_t *h, int32_t K, int32_t reps, double *out_ms);

/* Schedule of the LSMR launches (no reference counterpart; the reference's
 * CPU loops have no schedule): tiles of the fused pass over the FE with most
 * levels and of the CSR D'u launch (one wave per tile; 0 = sized by
 * occupancy, or FEHIP_TILES_S / FEHIP_TILES_C), few-level FEs accumulated in
 * the fused pass (1), in their own launch (0) or by default (-1), and LSMR
 * iterations per captured graph (0 = keep).  Drops the handle's plans; the
 * next solve rebuilds them.  Results do not depend on the schedule beyond
 * the order of floating-point sums. */
int fehip_set_schedule(fehip_t *h, int64_t tiles_s, int64_t tiles_c, int32_t fuse_dense, int32_t poll_iters);

/* Autotuner (no reference counterpart): times the launches of one LSMR
 * iteration on K columns of this handle's data (`reps` launches each) under
 * candidate schedules around the occupancy-sized default and keeps the
 * fastest on the handle.  out[6] = {tiles_s, tiles_c, fuse_dense,
 * poll_iters, best ms per iteration, default ms per iteration}. */
int fehip_autotune(fehip_t *h, int32_t K, int32_t reps, double *out);

/* Dimensions of a handle. */
int fehip_dims(const fehip_t *h, int64_t *N, int32_t *G, int64_t *p);

/* Message for the last failed call on this thread ("" if none). */
const char *fehip_last_error(void);

/* Frees the handle and its device memory (NULL is a no-op). */
void fehip_destroy(fehip_t *h);

#ifdef __cplusplus
}
#endif
#endif
