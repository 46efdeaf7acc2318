// fehip_common.h -- shared types and wave helpers for libfehip (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fehip {

constexpr int kWave = 64;
constexpr int kBlock = 256;               // 4 waves per workgroup
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kMaxG = 16;                 // fixed effects per handle
constexpr int kMaxK = 32;                 // columns per batched launch
constexpr int kFinBlock = 1024;           // threads of a finalize launch

// What a reduction launch and the finalize launch after it do with the
// column norms the reduction forms.
enum Mode : int {
    MODE_NONE = 0,  // plain mul!: no norm, no LSMR bookkeeping
    MODE_INIT = 1,  // LSMR set-up (src/utils/lsmr.jl:63-98)
    MODE_ITER = 2,  // LSMR iteration (src/utils/lsmr.jl:102-215)
};

// Per-column LSMR state, resident in device memory for the whole solve.
// Tr is the reference's scalar type Tr (src/utils/lsmr.jl:59-60): Float64 for
// FEHIP_FP64 and FEHIP_MIXED, Float32 for FEHIP_FP32.
template <typename Tr>
struct LsmrState {
    Tr alpha, beta;
    Tr zetabar, alphabar, rho, rhobar, cbar, sbar;      // :70-75
    Tr betadd, betad, rhodold, tautildeold, thetatilde;  // :81-85
    Tr zeta, d;                                          // :86-87
    Tr normA2, maxrbar, minrbar;                         // :90-92
    Tr normb, normr, normAr, normA, condA, normx;        // :89, :95-98
    Tr inv_beta;   // stored u is raw: u = u_raw * inv_beta (:65, :107)
    Tr inv_alpha;  // v is normalized by the update kernel (:68, :110)
    Tr c_hbar, c_x, c_h;  // coefficients of the pending h/hbar/x update (:138-142)
    Tr ctol;
    double atol, btol;
    int64_t iter, maxiter;
    int32_t istop, done, vskip, active;
    int32_t pending;  // this iteration's stopping test is due (the next beta finalize runs it)
};

// Coefficients the LSMR bookkeeping hands to the next matvec launch, per
// column: target <- b[c]*target + a[c]*(A x or A'y); on[c] == 0 skips column c.
struct ColOps {
    double a[kMaxK];
    double b[kMaxK];
    int32_t on[kMaxK];
};

__device__ __forceinline__ uint64_t lanes_le(int l) {
    return l >= 63 ? ~0ull : ((2ull << l) - 1ull);
}
__device__ __forceinline__ uint64_t lanes_lt(int l) {
    return l <= 0 ? 0ull : (~0ull >> (64 - l));
}

// Sum over the lanes of one column c (lanes c, c+K, c+2K, ...) with a fixed
// butterfly order; every lane of the column receives the total.
template <int K, typename A>
__device__ __forceinline__ A column_allreduce(A v) {
#pragma unroll
    for (int d = K; d < kWave; d <<= 1) v += __shfl_xor(v, d, kWave);
    return v;
}

// Block-wide per-column sum of a double (lane layout lane = psub*K + c).
// Result valid in threads 0..K-1 (thread c holds column c).  Uses `red`
// (kWavesPerBlock*K doubles of LDS).
template <int K>
__device__ __forceinline__ double block_column_sum(double v, double *red) {
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    v = column_allreduce<K>(v);
    if (lane < K) red[wave * K + lane] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x < K) {
#pragma unroll
        for (int w = 0; w < kWavesPerBlock; w++) s += red[w * K + threadIdx.x];
    }
    __syncthreads();
    return s;
}

// Deterministic sum of per-column partials laid out [part*K + c] (two spans
// of parts).  All NT threads of the block take part; thread c < K returns
// column c's total.  Fixed order: thread t owns column t%K and parts
// r = t/K, r + R, r + 2R, ... (R = NT/K) accumulated in four interleaved
// chains, then a fixed tree over the R threads of the column.
template <int K, int NT>
__device__ double block_sum_partials(const double *p0, int64_t n0, const double *p1, int64_t n1,
                                     double *red /* NT doubles */) {
    constexpr int R = NT / K;
    const int t = threadIdx.x, c = t % K, r = t / K;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (int span = 0; span < 2; span++) {
        const double *p = span ? p1 : p0;
        const int64_t n = span ? n1 : n0;
        int64_t j = r;
        for (; j + 3 * R < n; j += 4 * R) {
            s0 += p[j * K + c];
            s1 += p[(j + R) * K + c];
            s2 += p[(j + 2 * R) * K + c];
            s3 += p[(j + 3 * R) * K + c];
        }
        for (; j < n; j += R) s0 += p[j * K + c];
    }
    red[t] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    for (int w = R / 2; w >= 1; w >>= 1) {
        if (r < w) red[t] += red[t + w * K];
        __syncthreads();
    }
    double out = (t < K) ? red[t] : 0.0;
    __syncthreads();
    return out;
}

// Two such sums in one pass, A over pa (na parts) and B over pb (nb parts),
// each in block_sum_partials's order (so each total is bitwise that
// function's with one span): their loads are in flight together.
template <int K, int NT>
__device__ void block_sum_partials_pair(const double *pa, int64_t na, const double *pb, int64_t nb,
                                        double *red /* 2 * NT doubles */, double &oa, double &ob) {
    constexpr int R = NT / K;
    const int t = threadIdx.x, c = t % K, r = t / K;
    double a[4] = {0.0, 0.0, 0.0, 0.0}, b[4] = {0.0, 0.0, 0.0, 0.0};
    int64_t ja = r, jb = r;
    while (ja + 3 * R < na || jb + 3 * R < nb) {
        if (ja + 3 * R < na) {
#pragma unroll
            for (int k = 0; k < 4; k++) a[k] += pa[(ja + k * R) * K + c];
            ja += 4 * R;
        }
        if (jb + 3 * R < nb) {
#pragma unroll
            for (int k = 0; k < 4; k++) b[k] += pb[(jb + k * R) * K + c];
            jb += 4 * R;
        }
    }
    for (; ja < na; ja += R) a[0] += pa[ja * K + c];
    for (; jb < nb; jb += R) b[0] += pb[jb * K + c];
    red[t] = (a[0] + a[1]) + (a[2] + a[3]);
    red[NT + t] = (b[0] + b[1]) + (b[2] + b[3]);
    __syncthreads();
    for (int w = R / 2; w >= 1; w >>= 1) {
        if (r < w) {
            red[t] += red[t + w * K];
            red[NT + t] += red[NT + t + w * K];
        }
        __syncthreads();
    }
    oa = (t < K) ? red[t] : 0.0;
    ob = (t < K) ? red[NT + t] : 0.0;
    __syncthreads();
}

}  // namespace fehip
