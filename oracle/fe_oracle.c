/*
 * fe_oracle.c -- CPU restatement of FixedEffects.jl v3.1.0's residualization
 * path.  TEST INFRASTRUCTURE ONLY: the checker for libfehip and the CPU
 * baseline timed beside it.  Not part of the product; never linked by it.
 *
 * Built by oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).  Contraction
 * is off because Julia does not fuse a*b+c outside @fastmath/muladd.
 */
#include "fe_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* GroupedArrays' @spawn_for_chunks 100_000 (src/SolverCPU.jl:31,92) */
#define OR_CHUNK 100000

/* bucketize_refs  (ext/MetalExt.jl:40-64): counts (:42-45), offsets =
 * vcat(1, 1 .+ cumsum(counts)) (:47-53), then a stable placement in ascending
 * i through a per-level cursor (:56-62).  All indices 1-based. */
void oracle_bucketize(const int64_t *refs, int64_t N, int64_t n,
                      int64_t *perm, int64_t *offsets)
{
    int64_t *counts = (int64_t *)calloc(n, sizeof(int64_t));
    for (int64_t i = 0; i < N; i++) counts[refs[i] - 1] += 1;
    offsets[0] = 1;
    for (int64_t k = 0; k < n; k++) offsets[k + 1] = offsets[k] + counts[k];
    int64_t *next = counts; /* reuse: next = offsets[1:n] */
    for (int64_t k = 0; k < n; k++) next[k] = offsets[k];
    for (int64_t i = 0; i < N; i++) {
        int64_t r = refs[i] - 1;
        int64_t p = next[r];
        perm[p - 1] = i + 1;
        next[r] = p + 1;
    }
    free(counts);
}

#define OR_S d
#define OR_T double
#define OR_BIG 1e100
#include "fe_oracle_impl.inc"
#undef OR_S
#undef OR_T
#undef OR_BIG

/* convert(Float32, 1e100) is Inf32 (src/utils/lsmr.jl:92) */
#define OR_S f
#define OR_T float
#define OR_BIG INFINITY
#include "fe_oracle_impl.inc"
#undef OR_S
#undef OR_T
#undef OR_BIG

/* components  (src/FixedEffect.jl:68-98): breadth-first walk over the
 * observation/level graph.  refsrev (:57-63) is the level -> observations
 * inverted index, built here as a CSR.  The reference's Set iteration order
 * does not change which observations and levels end up together, so the
 * labels are the same partition; ids are assigned in order of each
 * component's first observation, as the reference pushes components. */
int64_t oracle_components(int64_t N, int32_t G, const int64_t *const *refs,
                          const int64_t *n, int64_t *const *label)
{
    int64_t **roff = (int64_t **)malloc(sizeof(int64_t *) * G);
    int64_t **robs = (int64_t **)malloc(sizeof(int64_t *) * G);
    for (int32_t g = 0; g < G; g++) {
        roff[g] = (int64_t *)malloc(sizeof(int64_t) * (n[g] + 1));
        robs[g] = (int64_t *)malloc(sizeof(int64_t) * N);
        oracle_bucketize(refs[g], N, n[g], robs[g], roff[g]);
        for (int64_t k = 0; k < n[g]; k++) label[g][k] = -1;
    }
    unsigned char *visited = (unsigned char *)calloc(N, 1);
    int64_t *queue = (int64_t *)malloc(sizeof(int64_t) * N);
    int64_t ncomp = 0;
    for (int64_t i0 = 0; i0 < N; i0++) {
        if (visited[i0]) continue;
        int64_t head = 0, tail = 0;
        queue[tail++] = i0;
        visited[i0] = 1;
        while (head < tail) {
            int64_t i = queue[head++];
            for (int32_t g = 0; g < G; g++) {
                int64_t k = refs[g][i] - 1;
                if (label[g][k] >= 0) continue;          /* ref in component */
                label[g][k] = ncomp;                     /* push!(component) */
                for (int64_t q = roff[g][k] - 1; q < roff[g][k + 1] - 1; q++) {
                    int64_t j = robs[g][q] - 1;           /* union!(tovisit) */
                    if (!visited[j]) {
                        visited[j] = 1;
                        queue[tail++] = j;
                    }
                }
            }
        }
        ncomp++;
    }
    for (int32_t g = 0; g < G; g++) { free(roff[g]); free(robs[g]); }
    free(roff); free(robs); free(visited); free(queue);
    return ncomp;
}

/* rescale!  (src/FixedEffect.jl:113-135) applied through normalize!
 * (:106-111) to the FEs whose interaction is UnitWeights, when there are at
 * least two of them.  Per component: FE j = last..2 is demeaned over the
 * component's levels and the means are added to FE 1.  Level sums run in
 * ascending level order (the reference iterates a Set). */
void oracle_normalize(double *coef, int64_t N, int32_t G,
                      const int64_t *const *refs, const int64_t *n,
                      const int32_t *unit_interaction)
{
    int32_t nsel = 0;
    int32_t *sel = (int32_t *)malloc(sizeof(int32_t) * G);
    int64_t *off = (int64_t *)malloc(sizeof(int64_t) * (G + 1));
    off[0] = 0;
    for (int32_t g = 0; g < G; g++) {
        off[g + 1] = off[g] + n[g];
        if (unit_interaction[g]) sel[nsel++] = g;
    }
    if (nsel >= 2) {
        const int64_t **srefs = (const int64_t **)malloc(sizeof(int64_t *) * nsel);
        int64_t *sn = (int64_t *)malloc(sizeof(int64_t) * nsel);
        int64_t **label = (int64_t **)malloc(sizeof(int64_t *) * nsel);
        for (int32_t j = 0; j < nsel; j++) {
            srefs[j] = refs[sel[j]];
            sn[j] = n[sel[j]];
            label[j] = (int64_t *)malloc(sizeof(int64_t) * sn[j]);
        }
        int64_t nc = oracle_components(N, nsel, srefs, sn, label);
        double *m = (double *)calloc(nc, sizeof(double));
        double *sum = (double *)malloc(sizeof(double) * nc);
        int64_t *cnt = (int64_t *)malloc(sizeof(int64_t) * nc);
        for (int32_t j = nsel - 1; j >= 1; j--) {
            double *c = coef + off[sel[j]];
            for (int64_t q = 0; q < nc; q++) { sum[q] = 0.0; cnt[q] = 0; }
            for (int64_t k = 0; k < sn[j]; k++)
                if (label[j][k] >= 0) { sum[label[j][k]] += c[k]; cnt[label[j][k]]++; }
            for (int64_t q = 0; q < nc; q++) sum[q] = sum[q] / (double)cnt[q];
            for (int64_t k = 0; k < sn[j]; k++)
                if (label[j][k] >= 0) c[k] -= sum[label[j][k]];
            for (int64_t q = 0; q < nc; q++) m[q] += sum[q];
        }
        double *c1 = coef + off[sel[0]];
        for (int64_t k = 0; k < sn[0]; k++)
            if (label[0][k] >= 0) c1[k] += m[label[0][k]];
        for (int32_t j = 0; j < nsel; j++) free(label[j]);
        free(label); free(srefs); free(sn); free(m); free(sum); free(cnt);
    }
    free(sel);
    free(off);
}

/* full  (src/FixedEffect.jl:137-140) */
void oracle_full(const double *coef, int64_t N, int32_t G,
                 const int64_t *const *refs, const int64_t *n, double *out)
{
    int64_t off = 0;
    for (int32_t g = 0; g < G; g++) {
        for (int64_t i = 0; i < N; i++) out[g * N + i] = coef[off + refs[g][i] - 1];
        off += n[g];
    }
}
