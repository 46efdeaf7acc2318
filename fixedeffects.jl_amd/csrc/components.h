// components.h -- normalize! and full on the device (src/FixedEffect.jl:57-140).
//
// The coefficients of a fixed-effects least-squares problem are determined
// only up to one constant per FE and connected component of the
// observation/level graph; the reference pins them by demeaning every
// FE but the first within each component and adding the means to the first
// (rescale!, :113-135).  The components come from a graph walk over
// observations (:68-98).  Here:
//
//   * nodes are the levels of the participating FEs in the coefficient index
//     space (FE g's level l is node off[g] + l); every observation links the
//     level of the first participating FE with the level of each other one;
//   * k_cc_hook unites the two ends of every link with a lock-free union-find
//     (a root is hooked under a smaller root with compare-and-swap, path
//     halving on the way up), so parent[v] <= v always and each tree's root is
//     the smallest node of its component -- a label independent of timing;
//   * roots are numbered densely in node order, nodes are stably sorted by
//     component number (unobserved levels last), and one workgroup per
//     component forms the per-FE sums in a fixed order (deterministic
//     results), then applies the shifts;
//   * k_full expands the coefficients to observations (full, :137-140).
#pragma once

#include "fehip_common.h"

namespace fehip {

__device__ __forceinline__ int32_t cc_load(const int32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void cc_store(int32_t *p, int32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x.  Every value ever held by parent[y] is <= y, and strictly less
// once y stops being a root, so the walk strictly decreases and ends even on
// a stale read.  Path halving stores a grandparent (same tree, smaller id)
// into a node that is no longer a root; roots are only written by the CAS.
__device__ __forceinline__ int32_t cc_find(int32_t *parent, int32_t x) {
    int32_t cur = x;
    int32_t nxt = cc_load(parent + cur);
    while (nxt != cur) {
        const int32_t nn = cc_load(parent + nxt);
        if (nn != nxt) cc_store(parent + cur, nn);
        cur = nxt;
        nxt = nn;
    }
    return cur;
}

// Unite the trees of a and b: the larger root is hooked under the smaller.
// A failed CAS means that root gained a parent; continuing from the new roots
// strictly lowers the larger of the two, so the loop ends.
__device__ __forceinline__ void cc_unite(int32_t *parent, int32_t a, int32_t b) {
    int32_t ra = cc_find(parent, a), rb = cc_find(parent, b);
    while (ra != rb) {
        if (ra < rb) {
            const int32_t t = ra;
            ra = rb;
            rb = t;
        }
        const int32_t old = atomicCAS(parent + ra, ra, rb);
        if (old == ra) break;
        ra = cc_find(parent, old);
        rb = cc_find(parent, rb);
    }
}

struct CcParams {
    const int32_t *refs[kMaxG];  // participating FEs, internal order, 0-based
    int64_t node0[kMaxG];        // node of level 0 of each participating FE
    int32_t nu;                  // participating FEs (>= 2)
};

__global__ void k_cc_init(int32_t *__restrict__ parent, uint8_t *__restrict__ seen, int64_t p) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < p) {
        parent[v] = (int32_t)v;
        seen[v] = 0;
    }
}

// One observation per thread: mark its levels observed and link them.
__global__ __launch_bounds__(kBlock) void k_cc_hook(CcParams prm, int64_t N, int32_t *parent,
                                                    uint8_t *__restrict__ seen) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int32_t a = (int32_t)(prm.node0[0] + prm.refs[0][i]);
    seen[a] = 1;
    for (int j = 1; j < prm.nu; j++) {
        const int32_t b = (int32_t)(prm.node0[j] + prm.refs[j][i]);
        seen[b] = 1;
        cc_unite(parent, a, b);
    }
}

// flag[v] = 1 for the root of every observed component (after k_cc_hook).
__global__ void k_cc_roots(const int32_t *__restrict__ parent, const uint8_t *__restrict__ seen, int64_t p,
                           uint32_t *__restrict__ flag) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < p) flag[v] = (seen[v] && parent[v] == (int32_t)v) ? 1u : 0u;
}

// key[v] = component number of node v (roots numbered in node order by the
// exclusive scan rid), or the component count for an unobserved level.
__global__ void k_cc_label(int32_t *parent, const uint8_t *__restrict__ seen, const uint32_t *__restrict__ rid,
                           int64_t p, int32_t *__restrict__ key) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= p) return;
    key[v] = seen[v] ? (int32_t)rid[cc_find(parent, (int32_t)v)] : (int32_t)rid[p];
}

// Component c occupies positions [start[c], start[c + 1]) of the sorted node
// list.  Component numbers are dense, so every number up to the largest key
// starts somewhere; start[last key + 1] = p closes the list.
__global__ void k_cc_starts(const int32_t *__restrict__ skeys, int64_t p, int32_t *__restrict__ start) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j > p) return;
    if (j == p) {
        start[skeys[p - 1] + 1] = (int32_t)p;
        return;
    }
    if (j == 0 || skeys[j] != skeys[j - 1]) start[skeys[j]] = (int32_t)j;
}

struct RescaleParams {
    int64_t off[kMaxG + 1];  // coefficient offsets of all FEs (node space)
    int32_t role[kMaxG];     // -1: not participating, 0: first, j >= 1: j-th other
    int32_t G;
    int32_t nu;
};

__device__ __forceinline__ int cc_fe_of(const RescaleParams &prm, int64_t v) {
    int g = 0;
    while (g + 1 < prm.G && v >= prm.off[g + 1]) g++;
    return g;
}

// rescale! (:113-135), one workgroup per component (grid-stride over the
// component numbers 0..*ncomp-1).  Per participating FE: sum and count of its
// levels in the component, lane-strided then reduced in a fixed tree; the
// first thread forms the means from the last FE down to the second (:117-128)
// and their total m; then every level of FE j >= 2 loses its mean and every
// level of the first gains m (:130-133).
__global__ __launch_bounds__(kBlock) void k_rescale(RescaleParams prm, const int32_t *__restrict__ nodes,
                                                    const int32_t *__restrict__ start,
                                                    const uint32_t *__restrict__ ncomp, double *__restrict__ coef) {
    __shared__ double s_sum[kWavesPerBlock][kMaxG];
    __shared__ int32_t s_cnt[kWavesPerBlock][kMaxG];
    __shared__ double s_shift[kMaxG];
    const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
    const int64_t C = (int64_t)ncomp[0];
    for (int64_t c = blockIdx.x; c < C; c += gridDim.x) {
        const int32_t lo = start[c], hi = start[c + 1];
        double sum[kMaxG];
        int32_t cnt[kMaxG];
#pragma unroll
        for (int g = 0; g < kMaxG; g++) {
            sum[g] = 0.0;
            cnt[g] = 0;
        }
        for (int32_t e = lo + threadIdx.x; e < hi; e += kBlock) {
            const int32_t v = nodes[e];
            const int g = cc_fe_of(prm, v);
            const double x = coef[v];
#pragma unroll
            for (int q = 0; q < kMaxG; q++)
                if (q == g) {
                    sum[q] += x;
                    cnt[q] += 1;
                }
        }
#pragma unroll
        for (int q = 0; q < kMaxG; q++) {
            if (q >= prm.G) continue;
            double s = sum[q];
            int32_t n = cnt[q];
            for (int d = kWave / 2; d >= 1; d >>= 1) {
                s += __shfl_xor(s, d);
                n += __shfl_xor(n, d);
            }
            if (lane == 0) {
                s_sum[wv][q] = s;
                s_cnt[wv][q] = n;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double tot[kMaxG];
            int32_t num[kMaxG];
            for (int q = 0; q < prm.G; q++) {
                tot[q] = 0.0;
                num[q] = 0;
                for (int w = 0; w < kWavesPerBlock; w++) {
                    tot[q] += s_sum[w][q];
                    num[q] += s_cnt[w][q];
                }
            }
            double m = 0.0;
            int first = -1;
            for (int j = prm.nu - 1; j >= 1; j--) {
                for (int q = 0; q < prm.G; q++) {
                    if (prm.role[q] != j) continue;
                    const double mj = num[q] > 0 ? tot[q] / (double)num[q] : 0.0;
                    s_shift[q] = -mj;
                    m += mj;
                }
            }
            for (int q = 0; q < prm.G; q++) {
                if (prm.role[q] == 0) first = q;
                if (prm.role[q] < 0) s_shift[q] = 0.0;
            }
            if (first >= 0) s_shift[first] = m;
        }
        __syncthreads();
        for (int32_t e = lo + threadIdx.x; e < hi; e += kBlock) {
            const int32_t v = nodes[e];
            const int g = cc_fe_of(prm, v);
            coef[v] += s_shift[g];
        }
        __syncthreads();
    }
}

// full (:137-140): out[caller's observation] = coef[level], for one FE.
__global__ void k_full(const double *__restrict__ coef, const int32_t *__restrict__ ref_int,
                       const int32_t *__restrict__ perm_s, int64_t N, double *__restrict__ out) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < N) out[perm_s[r]] = coef[ref_int[r]];
}

}  // namespace fehip
