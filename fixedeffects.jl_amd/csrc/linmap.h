// linmap.h -- the fixed-effect linear map A (N x p) and its adjoint on gfx950,
// batched over K columns, plus the fused LSMR vector update.
//
// Layout in HBM (internal observation order = observations sorted stably by
// the level of the FE with most levels, "FE s"):
//   u, r      : N x K, observation-major (u[i*K + c]); one ref/cache load
//               serves all K columns.
//   x,v,h,hbar: p x K, level-major (x[(off_g + k)*K + c]).
//   refs_g    : int32, 0-based, internal order.   cache_g: T, internal order.
//   CSR of FE g != s: perm (CSR position -> internal observation) and the
//               cache in CSR order; FE s needs neither (its levels are
//               contiguous runs of the internal order).
//
// Wave lane layout for K columns: lane = psub*K + c, P = 64/K positions per
// wave step.  Reductions are formed in a fixed order (no float atomics), so
// results are bitwise reproducible run to run.
#pragma once

#include <type_traits>

#include "fehip_common.h"

namespace fehip {

// ---------------------------------------------------------------------------
// Accumulator types.  FP64 and FP32 modes accumulate in the storage type, as
// the reference's gather!/scatter! do (src/SolverCPU.jl:21-34).  The mixed
// mode accumulates FP32 products with a compensated (two-sum) pair.
// ---------------------------------------------------------------------------
struct F32x2 {
    float hi, lo;
    __device__ F32x2() : hi(0.f), lo(0.f) {}
    __device__ F32x2(float x) : hi(x), lo(0.f) {}
    __device__ F32x2(float h, float l) : hi(h), lo(l) {}
    __device__ explicit operator double() const { return (double)hi + (double)lo; }
    __device__ explicit operator float() const { return hi + lo; }
};
// Knuth two-sum of the high parts, low parts summed, then renormalized.
__device__ __forceinline__ F32x2 operator+(F32x2 a, F32x2 b) {
#pragma clang fp contract(off)
    const float s = a.hi + b.hi;
    const float bp = s - a.hi;
    const float e = (a.hi - (s - bp)) + (b.hi - bp);
    const float lo = (a.lo + b.lo) + e;
    const float h = s + lo;
    return F32x2(h, lo - (h - s));
}
__device__ __forceinline__ F32x2 &operator+=(F32x2 &a, F32x2 b) { return a = a + b; }
__device__ __forceinline__ F32x2 __shfl_up(F32x2 v, unsigned d, int w) {
    return F32x2(::__shfl_up(v.hi, d, w), ::__shfl_up(v.lo, d, w));
}
__device__ __forceinline__ F32x2 __shfl_xor(F32x2 v, int m, int w) {
    return F32x2(::__shfl_xor(v.hi, m, w), ::__shfl_xor(v.lo, m, w));
}
__device__ __forceinline__ F32x2 __shfl(F32x2 v, int l, int w) {
    return F32x2(::__shfl(v.hi, l, w), ::__shfl(v.lo, l, w));
}
using ::__shfl;
using ::__shfl_up;
using ::__shfl_xor;

template <typename Acc, typename T>
__device__ __forceinline__ Acc mul_acc(T a, T b) {
    return Acc(a * b);
}
// exact product of two floats as a pair (fma gives the rounding error)
template <>
__device__ __forceinline__ F32x2 mul_acc<F32x2, float>(float a, float b) {
    const float p = a * b;
    return F32x2(p, __builtin_fmaf(a, b, -p));
}
// exact product of two floats in a double (24 + 24 significant bits)
template <>
__device__ __forceinline__ double mul_acc<double, float>(float a, float b) {
    return (double)a * (double)b;
}

template <typename Acc>
__device__ __forceinline__ double acc_to_double(Acc a) {
    return (double)a;
}
template <typename T, typename Acc>
__device__ __forceinline__ T acc_to_T(Acc a) {
    return (T)a;
}

// Cross-lane moves through DPP (no LDS round trip): CTRL is the DPP control
// (row_shr:n = 0x110+n, row_bcast:15 = 0x142, row_bcast:31 = 0x143), RM the
// row mask; lanes without a source read 0.
template <int CTRL, int RM>
__device__ __forceinline__ int dpp_i32(int x) {
    return __builtin_amdgcn_update_dpp(0, x, CTRL, RM, 0xF, true);
}
template <int CTRL, int RM>
__device__ __forceinline__ double dpp_move(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)dpp_i32<CTRL, RM>((int)(unsigned)u);
    const unsigned hi = (unsigned)dpp_i32<CTRL, RM>((int)(unsigned)(u >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
template <int CTRL, int RM>
__device__ __forceinline__ float dpp_move(float v) {
    return __builtin_bit_cast(float, dpp_i32<CTRL, RM>(__builtin_bit_cast(int, v)));
}
template <int CTRL, int RM>
__device__ __forceinline__ F32x2 dpp_move(F32x2 v) {
    return F32x2(dpp_move<CTRL, RM>(v.hi), dpp_move<CTRL, RM>(v.lo));
}
// Swap of the wave's 32-lane halves (gfx950 v_permlane32_swap, a VALU op):
// every lane gets the value of lane (l + 32) mod 64.
__device__ __forceinline__ int half_swap_i32(int x) {
    const auto r = __builtin_amdgcn_permlane32_swap(x, x, false, false);
    return (threadIdx.x & 32) ? (int)r[0] : (int)r[1];
}
__device__ __forceinline__ double half_swap(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)half_swap_i32((int)(unsigned)u);
    const unsigned hi = (unsigned)half_swap_i32((int)(unsigned)(u >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ float half_swap(float v) {
    return __builtin_bit_cast(float, half_swap_i32(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ F32x2 half_swap(F32x2 v) { return F32x2(half_swap(v.hi), half_swap(v.lo)); }

// value of lane l, uniform across the wave
__device__ __forceinline__ double lane_value(double v, int l) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ float lane_value(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ __forceinline__ F32x2 lane_value(F32x2 v, int l) { return F32x2(lane_value(v.hi, l), lane_value(v.lo, l)); }

// Load of element i of set-up data (head words and their prefixes) at a
// wave-uniform index, through the scalar unit: the constant address space
// tells the compiler that no kernel of the iteration writes it.  As a vector
// load it took a vector-memory slot per wave step and, since vector loads
// complete in order, its move to a scalar register waited for every load the
// software pipeline had in flight.
template <typename V>
__device__ __forceinline__ V ld_const(const V *base, int64_t i) {
    return ((const __attribute__((address_space(4))) V *)base)[i];
}

// Segmented inclusive scan over the positions of one wave step.  rs is the
// psub of the latest segment head at or before this lane (0 if none); lanes
// only add values of their own run.  K = 1: scan within rows of 16 lanes
// (DPP row shifts), then the rows' totals are carried across with
// row_bcast:15 and row_bcast:31; a lane adds a moved value only if no head
// lies between its source and itself.  K > 1: shuffles by multiples of K.
template <int K, typename A>
__device__ __forceinline__ A seg_scan(A v, int psub, int rs) {
    if constexpr (K == 1) {
        const int r16 = psub & 15;
        A up = dpp_move<0x111, 0xF>(v);
        if (r16 >= 1 && psub - 1 >= rs) v += up;
        up = dpp_move<0x112, 0xF>(v);
        if (r16 >= 2 && psub - 2 >= rs) v += up;
        up = dpp_move<0x114, 0xF>(v);
        if (r16 >= 4 && psub - 4 >= rs) v += up;
        up = dpp_move<0x118, 0xF>(v);
        if (r16 >= 8 && psub - 8 >= rs) v += up;
        up = dpp_move<0x142, 0xA>(v);  // rows 1 and 3 read lanes 15 and 47
        if ((psub & 16) && rs < (psub & ~15)) v += up;
        up = dpp_move<0x143, 0xC>(v);  // rows 2 and 3 read lane 31
        if (psub >= 32 && rs < 32) v += up;
        return v;
    } else if constexpr (K == 32) {
        // two positions per step: the upper half adds the lower half's value
        // unless its own position starts a segment
        const A up = half_swap(v);
        if (psub == 1 && rs == 0) v += up;
        return v;
    } else {
        constexpr int P = kWave / K;
#pragma unroll
        for (int d = 1; d < P; d <<= 1) {
            A up = __shfl_up(v, (unsigned)(d * K), kWave);
            if (psub - d >= rs) v += up;
        }
        return v;
    }
}

// ---------------------------------------------------------------------------
// LSMR scalar bookkeeping (src/utils/lsmr.jl), run by one thread per column
// of the single-workgroup finalize launch that follows each reduction
// (k_fin_db, k_fin_dtx).  Contraction is disabled so the scalar
// recurrences round like the reference's.
// ---------------------------------------------------------------------------
template <typename Tr>
__device__ __forceinline__ Tr big_minrbar() { return (Tr)1e100; }
template <>
__device__ __forceinline__ float big_minrbar<float>() { return __builtin_inff(); }  // Float32(1e100) = Inf32

// after u = A v - alpha u (or u = b - A x at set-up): beta = ||u||  (:64-65, :105-107)
template <typename Tr>
__device__ void fin_after_db(LsmrState<Tr> &s, double usq, int mode, int c, ColOps *nx) {
#pragma clang fp contract(off)
    if (s.done) {
        nx->on[c] = 0;
        return;
    }
    const Tr beta = (Tr)sqrt(usq);
    s.beta = beta;
    if (mode == MODE_INIT) {
        s.normb = beta;
        s.inv_beta = beta > (Tr)0 ? (Tr)1 / beta : (Tr)1;
        nx->a[c] = (double)s.inv_beta;  // v = A'u (:66)
        nx->b[c] = 0.0;
        nx->on[c] = 1;
    } else {
        s.iter += 1;
        if (beta > (Tr)0) {
            s.inv_beta = (Tr)1 / beta;
            s.vskip = 0;
            nx->a[c] = (double)s.inv_beta;  // v = A'u - beta v (:108)
            nx->b[c] = (double)(-beta);
            nx->on[c] = 1;
        } else {
            s.inv_beta = (Tr)1;
            s.vskip = 1;  // the reference skips the v update (:106-111)
            nx->on[c] = 0;
        }
    }
}

// after v = A'u - beta v: alpha = ||v|| and the plane rotations (:67-98 at
// set-up; :109-182 per iteration).  Leaves the h/hbar/x coefficients.
template <typename Tr>
__device__ void fin_after_dtx(LsmrState<Tr> &s, double vsq, int mode, int c, ColOps *nx_db) {
#pragma clang fp contract(off)
    if (s.done) {
        nx_db->on[c] = 0;
        return;
    }
    if (mode == MODE_INIT) {
        const Tr alpha = (Tr)sqrt(vsq);
        s.alpha = alpha;
        s.inv_alpha = alpha > (Tr)0 ? (Tr)1 / alpha : (Tr)1;
        s.zetabar = alpha * s.beta;
        s.alphabar = alpha;
        s.rho = 1; s.rhobar = 1; s.cbar = 1; s.sbar = 0;
        s.betadd = s.beta; s.betad = 0; s.rhodold = 1; s.tautildeold = 0; s.thetatilde = 0;
        s.zeta = 0; s.d = 0;
        s.normA = -1; s.condA = -1; s.normx = -1;
        s.normA2 = alpha * alpha;
        s.maxrbar = 0;
        s.minrbar = big_minrbar<Tr>();
        s.normb = s.beta;
        s.istop = 0;
        s.normr = s.beta;
        s.normAr = alpha * s.beta;
        s.iter = 0;
        s.vskip = 0;
        if (s.normAr == (Tr)0 || s.maxiter <= 0) s.done = 1;  // :101-102
        nx_db->a[c] = 1.0;
        nx_db->b[c] = (double)(-(alpha * s.inv_beta));
        nx_db->on[c] = s.done ? 0 : 1;
        return;
    }
    if (!s.vskip) {
        const Tr alpha = (Tr)sqrt(vsq);
        s.alpha = alpha;
        s.inv_alpha = alpha > (Tr)0 ? (Tr)1 / alpha : (Tr)1;
    } else {
        s.inv_alpha = (Tr)1;
    }
    const Tr alpha = s.alpha, beta = s.beta, lambda = (Tr)0;
    // Qhat_{k,2k+1} (:114-116)
    const Tr alphahat = (Tr)sqrt((double)(s.alphabar * s.alphabar + lambda * lambda));
    const Tr chat = s.alphabar / alphahat;
    const Tr shat = lambda / alphahat;
    // Q_i (:119-124)
    const Tr rhoold = s.rho;
    s.rho = (Tr)sqrt((double)(alphahat * alphahat + beta * beta));
    const Tr cc = alphahat / s.rho;
    const Tr ss = beta / s.rho;
    const Tr thetanew = ss * alpha;
    s.alphabar = cc * alpha;
    // Qbar_i (:127-135)
    const Tr rhobarold = s.rhobar;
    const Tr zetaold = s.zeta;
    const Tr thetabar = s.sbar * s.rho;
    const Tr rhotemp = s.cbar * s.rho;
    s.rhobar = (Tr)sqrt((double)((s.cbar * s.rho) * (s.cbar * s.rho) + thetanew * thetanew));
    s.cbar = (s.cbar * s.rho) / s.rhobar;
    s.sbar = thetanew / s.rhobar;
    s.zeta = s.cbar * s.zetabar;
    s.zetabar = (-s.sbar) * s.zetabar;
    // h, hbar, x coefficients (:138-142)
    s.c_hbar = ((-thetabar) * s.rho) / (rhoold * rhobarold);
    s.c_x = s.zeta / (s.rho * s.rhobar);
    s.c_h = (-thetanew) / s.rho;
    // ||r|| estimate (:151-170)
    const Tr betaacute = chat * s.betadd;
    const Tr betacheck = (-shat) * s.betadd;
    const Tr betahat = cc * betaacute;
    s.betadd = (-ss) * betaacute;
    const Tr thetatildeold = s.thetatilde;
    const Tr rhotildeold = (Tr)sqrt((double)(s.rhodold * s.rhodold + thetabar * thetabar));
    const Tr ctildeold = s.rhodold / rhotildeold;
    const Tr stildeold = thetabar / rhotildeold;
    s.thetatilde = stildeold * s.rhobar;
    s.rhodold = ctildeold * s.rhobar;
    s.betad = (-stildeold) * s.betad + ctildeold * betahat;
    s.tautildeold = (zetaold - thetatildeold * s.tautildeold) / rhotildeold;
    const Tr taud = (s.zeta - s.thetatilde * s.tautildeold) / s.rhodold;
    s.d = s.d + betacheck * betacheck;
    s.normr = (Tr)sqrt((double)((s.d + (s.betad - taud) * (s.betad - taud)) + s.betadd * s.betadd));
    // ||A|| (:173-175)
    s.normA2 = s.normA2 + beta * beta;
    s.normA = (Tr)sqrt((double)s.normA2);
    s.normA2 = s.normA2 + alpha * alpha;
    // cond(A) (:178-182)
    s.maxrbar = s.maxrbar > rhobarold ? s.maxrbar : rhobarold;
    if (s.iter > 1) s.minrbar = s.minrbar < rhobarold ? s.minrbar : rhobarold;
    s.condA = (s.maxrbar > rhotemp ? s.maxrbar : rhotemp) / (s.minrbar < rhotemp ? s.minrbar : rhotemp);
    s.normAr = s.zetabar < (Tr)0 ? -s.zetabar : s.zetabar;  // :191
    // the next fused pass's coefficients, u = A v - alpha u (:104); this
    // iteration's stopping test (:193-215) needs ||x|| after the vector
    // update and runs in the next beta finalize (fin_db), after that pass
    nx_db->a[c] = 1.0;
    nx_db->b[c] = (double)(-(s.alpha * s.inv_beta));
    nx_db->on[c] = 1;
    s.pending = 1;
}

// after the h/hbar/x update: ||x|| and the stopping rules (:192-215).
template <typename Tr>
__device__ void fin_after_update(LsmrState<Tr> &s, double xsq, int c, ColOps *nx_db) {
#pragma clang fp contract(off)
    s.pending = 0;
    if (s.done) {
        nx_db->on[c] = 0;
        return;
    }
    s.normx = (Tr)sqrt(xsq);
    const Tr test1 = s.normr / s.normb;
    const Tr test2 = s.normAr / (s.normA * s.normr);
    const Tr test3 = (Tr)1 / s.condA;
    const Tr t1 = test1 / ((Tr)1 + s.normA * s.normx / s.normb);
    const double rtol = s.btol + ((s.atol * (double)s.normA) * (double)s.normx) / (double)s.normb;
    int istop = 0;
    if (s.iter >= s.maxiter) istop = 7;
    else if ((Tr)1 + test3 <= (Tr)1) istop = 6;
    else if ((Tr)1 + test2 <= (Tr)1) istop = 5;
    else if ((Tr)1 + t1 <= (Tr)1) istop = 4;
    else if (test3 <= s.ctol) istop = 3;
    else if ((double)test2 <= s.atol) istop = 2;
    else if ((double)test1 <= rtol) istop = 1;
    if (istop) {
        s.istop = istop;
        s.done = 1;
        nx_db->on[c] = 0;
    }
}

// all columns done -> *all_done = 1 (threads 0..K-1 hold their column's flag)
template <int K, typename Tr>
__device__ __forceinline__ void publish_all_done(const LsmrState<Tr> *st, int32_t *all_done) {
    if (threadIdx.x == 0) {
        int all = 1;
        for (int c = 0; c < K; c++) all &= st[c].done ? 1 : 0;
        *all_done = all;
    }
}

// ---------------------------------------------------------------------------
// D*x: target[i,c] <- b[c]*target[i,c] + a[c] * sum_g cache_g[i]*coef_g[ref_g[i], c]
// (reference scatter!/mul!, src/SolverCPU.jl:29-34, src/AbstractFixedEffectLinearMap.jl:37-44)
// ---------------------------------------------------------------------------
template <typename T>
struct DbParams {
    int64_t N;
    int32_t G;
    const int32_t *refs[kMaxG];
    const T *cache[kMaxG];
    int64_t coef_off[kMaxG];
    const T *coef;  // p x K
    T *target;      // N x K
    const ColOps *ops;
    double *partials;  // gridDim.x x K (||target||^2 per column, MODE_INIT/ITER)
    const int32_t *all_done;  // early exit when set (iterating launches)
    int mode;
};

template <typename T, int K>
__global__ __launch_bounds__(kBlock) void k_db(DbParams<T> p) {
    if (p.all_done && *p.all_done) return;
    constexpr int P = kWave / K;
    constexpr int U = 4;  // wave steps in flight per lane
    const int lane = threadIdx.x & (kWave - 1);
    const int psub = lane / K, c = lane % K;
    const bool on = p.ops->on[c] != 0;
    const T a = (T)p.ops->a[c], b = (T)p.ops->b[c];
    const bool keep = p.ops->b[c] != 0.0;
    const int64_t nrounds = (p.N + P - 1) / P;
    const int64_t gw = (int64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x / kWave);
    const int64_t nw = (int64_t)gridDim.x * kWavesPerBlock;
    double sq = 0.0;
    // loads are unconditional (positions past N read position 0), so no
    // load waits inside a branch; only the stores are predicated
    for (int64_t r0 = gw * U; r0 < nrounds; r0 += nw * U) {
        int64_t idx[U], ic[U];
        T acc[U], uo[U];
#pragma unroll
        for (int q = 0; q < U; q++) {
            idx[q] = (r0 + q) * P + psub;
            ic[q] = idx[q] < p.N ? idx[q] : 0;
            acc[q] = (T)0;
            uo[q] = p.target[ic[q] * K + c];
        }
        for (int g = 0; g < p.G; g++) {
            int32_t r[U];
            T cv[U];
#pragma unroll
            for (int q = 0; q < U; q++) {
                r[q] = __builtin_nontemporal_load(p.refs[g] + ic[q]);
                cv[q] = __builtin_nontemporal_load(p.cache[g] + ic[q]);
            }
            const T *cf = p.coef + p.coef_off[g] * K + c;
#pragma unroll
            for (int q = 0; q < U; q++) acc[q] += cv[q] * cf[(int64_t)r[q] * K];
        }
#pragma unroll
        for (int q = 0; q < U; q++) {
            T out = a * acc[q];
            out = keep ? b * uo[q] + out : out;
            if (on && idx[q] < p.N) {
                p.target[idx[q] * K + c] = out;
                sq += (double)out * (double)out;
            }
        }
    }
    if (p.mode == MODE_NONE) return;
    __shared__ double red[kBlock];
    const double s = block_column_sum<K>(sq, red);
    if (threadIdx.x < K) p.partials[(int64_t)blockIdx.x * K + threadIdx.x] = s;
}

// ---------------------------------------------------------------------------
// D'y: coef[k,c] <- b[c]*coef[k,c] + a[c] * sum_{i in level k} cache[i]*y[i,c]
// (reference gather!/mul!', src/SolverCPU.jl:21-27, src/AbstractFixedEffectLinearMap.jl:26-35)
//
// Segmented FEs (FE s and CSR FEs) are cut into tiles of CSR positions, one
// wave per tile: FE s into tiles of a fixed multiple of 256 positions (the
// fused pass's partition), a CSR FE at segment heads near a target size, so
// that only segments longer than a tile cross tiles.  A segment that lies
// inside a tile is reduced and written by that wave.  A segment crossing tile
// boundaries leaves a partial in its first tile (tileB) and in each later tile
// it covers (tileA); the fix-up launch combines them in tile order.
// ---------------------------------------------------------------------------
template <typename T>
struct SegFE {
    const int32_t *perm;      // CSR position -> internal observation (0..N-1 for the sorted FE)
    const T *cache;           // cache in CSR order
    const uint64_t *heads;    // segment head bits, 1 beyond N
    const uint32_t *wprefix;  // heads before each 64-position word
    const int32_t *seg_level; // ordinal of non-empty segment -> level
    const int32_t *plev;      // PLEV form: the level of each position (replaces the ordinal and seg_level)
    const T *lscale;          // RAW form: per-level scale applied to each level's sum (cache not read)
    const int64_t *tstart;    // tile i covers positions [tstart[i], tstart[i+1])
    int64_t coef_off;
    int32_t tile0, ntiles;
    int32_t nsegs, pad;       // non-empty segments
};

template <typename T>
struct DtxSegParams {
    int64_t N;
    int32_t nseg, first_tile, total_tiles, pad;  // tiles [first_tile, total_tiles)
    SegFE<T> fe[kMaxG];
    const T *u;
    T *v;
    double *tileA, *tileB;  // [tile*K + c]
    const ColOps *ops;
    double *partials;  // gridDim.x x K
    const int32_t *all_done;
    int mode;
};

// A wave's walk over one tile of segmented positions.
template <typename Acc>
struct SegCursor {
    Acc carry;  // open run carried into the next step
    bool cont;  // no segment head seen yet in this tile
    bool open;  // a run is still open after the last step
};

// One wave step at position `base` (P = 64/K positions, base a multiple of
// P; sh = base mod 64, valid = this lane's position lies in the tile): the
// run carried in is added and the step is scanned in place (x).  Returns true
// on the lane where a segment that started in this tile ends (the caller
// writes it); a segment that began in an earlier tile ends in *tileA_tc.
// w = heads[base >> 6], wnext = heads[(base >> 6) + 1], loaded by the caller
// together with the step's data, so no load sits on the scan's chain.
template <int K, typename Acc>
__device__ __forceinline__ bool seg_step(uint64_t w, uint64_t wnext, int sh, bool valid, int psub, int c, bool on,
                                         Acc &x, SegCursor<Acc> &cur, double *tileA_tc) {
    constexpr int P = kWave / K;
    const uint64_t hb = (P == 64) ? w : ((w >> sh) & ((1ull << (P & 63)) - 1ull));
    const uint64_t nb = (sh + P < 64) ? ((w >> (sh + P)) & 1ull) : (wnext & 1ull);
    const uint64_t hle = hb & lanes_le(psub);
    const int rs = hle ? (63 - __clzll(hle)) : 0;
    if (psub == 0 && !(hb & 1ull)) x += cur.carry;
    x = seg_scan<K>(x, psub, rs);
    const bool isend = (psub == P - 1) ? (nb != 0) : (((hb >> (psub + 1)) & 1ull) != 0);
    const bool mycont = cur.cont && (hle == 0);
    bool em = false;
    if (valid && on && isend) {
        if (mycont) *tileA_tc = acc_to_double(x);
        else em = true;
    }
    Acc last;
    if constexpr (K == 1) {
        last = lane_value(x, kWave - 1);
    } else if constexpr (K == 32) {
        const Acc other = half_swap(x);
        last = psub == 0 ? other : x;  // lane 32 + c for every lane of column c
    } else {
        last = __shfl(x, (P - 1) * K + c, kWave);
    }
    cur.open = (nb == 0);
    cur.carry = cur.open ? last : Acc(0.f);
    cur.cont = cur.cont && (hb == 0);
    return em;
}

// Wide lane map of the fused pass: every lane holds C adjacent columns of
// one position (lane = psub*KL + cg, columns cg*C .. cg*C+C-1, KL = K/C
// lanes and P = 64/KL positions per wave step), so the per-position work
// (refs, head words, level, addresses) is done once for C columns.  C = 1
// is the plain map (lane = psub*K + c).
template <typename T, int C>
struct alignas(sizeof(T) * C) VecC {
    T v[C];
};
// C adjacent elements at element offset i from a wave-uniform base (one
// 8- to 32-byte access; the host keeps every row of K elements aligned)
template <int C, typename T>
__device__ __forceinline__ VecC<T, C> ldv(const T *base, uint32_t i) {
    return *reinterpret_cast<const VecC<T, C> *>(reinterpret_cast<const char *>(base) + i * (uint32_t)sizeof(T));
}
template <int C, typename T>
__device__ __forceinline__ void stv(T *base, uint32_t i, const VecC<T, C> &v) {
    *reinterpret_cast<VecC<T, C> *>(reinterpret_cast<char *>(base) + i * (uint32_t)sizeof(T)) = v;
}

template <typename Acc, int C>
struct SegCursorC {
    Acc carry[C];  // open runs carried into the next step, per column
    bool cont;     // no segment head seen yet in this tile
    bool open;     // a run is still open after the last step
};

// seg_step for C columns per lane (KL lanes per position): the head-bit
// logic once per position, the carry, scan and hand-over per column.  `on`
// is whether any of the lane's columns is on; values of columns that are
// off are zero and their sums are never read.  tileA_tc points at the
// lane's first column.
template <int KL, int C, typename Acc>
__device__ __forceinline__ bool seg_step_c(uint64_t w, uint64_t wnext, int sh, bool valid, int psub, int cg, bool on,
                                           Acc (&x)[C], SegCursorC<Acc, C> &cur, double *tileA_tc) {
    constexpr int P = kWave / KL;
    const uint64_t hb = (P == 64) ? w : ((w >> sh) & ((1ull << (P & 63)) - 1ull));
    const uint64_t nb = (sh + P < 64) ? ((w >> (sh + P)) & 1ull) : (wnext & 1ull);
    const uint64_t hle = hb & lanes_le(psub);
    const int rs = hle ? (63 - __clzll(hle)) : 0;
    if (psub == 0 && !(hb & 1ull)) {
#pragma unroll
        for (int j = 0; j < C; j++) x[j] += cur.carry[j];
    }
#pragma unroll
    for (int j = 0; j < C; j++) x[j] = seg_scan<KL>(x[j], psub, rs);
    const bool isend = (psub == P - 1) ? (nb != 0) : (((hb >> (psub + 1)) & 1ull) != 0);
    const bool mycont = cur.cont && (hle == 0);
    bool em = false;
    if (valid && on && isend) {
        if (mycont) {
#pragma unroll
            for (int j = 0; j < C; j++) tileA_tc[j] = acc_to_double(x[j]);
        } else {
            em = true;
        }
    }
    cur.open = (nb == 0);
#pragma unroll
    for (int j = 0; j < C; j++) {
        Acc last;
        if constexpr (KL == 1) {
            last = lane_value(x[j], kWave - 1);
        } else if constexpr (KL == 32) {
            const Acc other = half_swap(x[j]);
            last = psub == 0 ? other : x[j];
        } else {
            last = __shfl(x[j], (P - 1) * KL + cg, kWave);
        }
        cur.carry[j] = cur.open ? last : Acc(0.f);
    }
    cur.cont = cur.cont && (hb == 0);
    return em;
}

// acc[slot[q]*64 + lane] += x[q] for q = 0..U-1 in this order, with one LDS
// round trip: all slots are read first, repeated slots are chained in
// registers, and the writes go out in order (the last write to a repeated
// slot carries every addition, in the same order as one-by-one updates).
template <int U, typename Acc>
__device__ __forceinline__ void lds_accumulate(Acc *acc, const int32_t (&slot)[U], const Acc (&x)[U], int lane) {
    Acc old[U], nv[U];
#pragma unroll
    for (int q = 0; q < U; q++) old[q] = acc[slot[q] * kWave + lane];
#pragma unroll
    for (int q = 0; q < U; q++) {
        Acc base = old[q];
#pragma unroll
        for (int q2 = 0; q2 < q; q2++) base = slot[q2] == slot[q] ? nv[q2] : base;
        nv[q] = base + x[q];
    }
#pragma unroll
    for (int q = 0; q < U; q++) acc[slot[q] * kWave + lane] = nv[q];
}

// level of the segment holding position pos (w = head word of pos,
// wp = wprefix of that word)
__device__ __forceinline__ int32_t seg_level_of(uint32_t wp, const int32_t *seg_level, int64_t pos, uint64_t w) {
    return seg_level[wp + (uint32_t)__popcll(w & lanes_le((int)(pos & 63))) - 1u];
}

// RAW (the factored form's iteration, where u is stored as sqrt(w) .* u):
// the segment sums are of the stored vector as it is, and each level's sum
// takes the FE's scale (lscale; crossing segments take it in the fix-up), so
// the CSR-ordered cache is not read.
// PLEV: each position's level is read from plev with its permutation entry
// (one memory round trip less per step than ordinal -> seg_level; used by the
// piece form, whose position counts are small).
// the body of one launch block (bid of nblk): launched alone (k_dtx_seg) or
// as the later blocks of k_dtx_seg_fix
template <typename T, typename Acc, int K, bool RAW, bool PLEV = false>
__device__ __forceinline__ void dtx_seg_body(const DtxSegParams<T> &p, int bid, int nblk) {
    constexpr int P = kWave / K;
    // wave steps whose loads are issued together; the piece launch (PLEV)
    // takes two (measured at K = 1: 26.1 / 25.3 us against 27.3 / 28.5 with
    // four, 35 with eight)
    constexpr int U = PLEV ? 2 : 4;
    const int lane = threadIdx.x & (kWave - 1);
    const int psub = lane / K, c = lane % K;
    const int on = p.ops->on[c];
    const T a = (T)p.ops->a[c], b = (T)p.ops->b[c];
    const bool keep = p.ops->b[c] != 0.0;
    const int gw = bid * kWavesPerBlock + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);  // wave-uniform
    const int nw = nblk * kWavesPerBlock;
    double sq = 0.0;
    for (int t = p.first_tile + gw; t < p.total_tiles; t += nw) {
        int f = 0;
        while (f + 1 < p.nseg && t >= p.fe[f + 1].tile0) f++;
        const SegFE<T> fe = p.fe[f];
        const int64_t start = fe.tstart[t - fe.tile0];
        const int64_t end = fe.tstart[t - fe.tile0 + 1];
        // the walk starts at the word holding start; the scan sees the head
        // bits of positions [start, end] only (end's bit ends the tile's last
        // segment), so positions outside the tile neither open nor close runs
        const int64_t wbeg = start & ~(int64_t)63;
        auto clip = [&](uint64_t w, int64_t wi) -> uint64_t {
            const int64_t lo = start - (wi << 6), hi = end - (wi << 6);
            if (hi < 0 || lo > 63) return 0ull;
            uint64_t m = lo > 0 ? ~0ull << lo : ~0ull;
            if (hi < 63) m &= (2ull << hi) - 1ull;
            return w & m;
        };
        T *v = p.v + fe.coef_off * K;
        SegCursor<Acc> cur{Acc(0.f), true, false};
        // per lane, one loop step ahead: the observation each position reads
        // and the head-bit words, so that a step's u gathers, its cache loads
        // and the v of the levels its positions belong to go out together
        int32_t in[U], lvn[U];
        uint64_t hwn[U], hnn[U];
        uint32_t wpn[U];
        auto meta = [&](int64_t b0) {
#pragma unroll
            for (int q = 0; q < U; q++) {
                const int64_t base = b0 + q * P;
                const int64_t pos = base + psub;
                const int64_t pc = (pos >= start && pos < end) ? pos : start;
                in[q] = __builtin_nontemporal_load(fe.perm + pc);
                if constexpr (PLEV) lvn[q] = fe.plev[pc];
                hwn[q] = hnn[q] = 0;
                wpn[q] = 0;
                if (base < end) {  // wave-uniform
                    hwn[q] = ld_const(fe.heads, base >> 6);
                    hnn[q] = ld_const(fe.heads, (base >> 6) + 1);
                    if constexpr (!PLEV) wpn[q] = ld_const(fe.wprefix, base >> 6);
                }
            }
        };
        meta(wbeg);
        for (int64_t base0 = wbeg; base0 < end; base0 += U * P) {
            int32_t iq[U];
            uint64_t hw[U], hn[U];
            int32_t lev[U];
#pragma unroll
            for (int q = 0; q < U; q++) {
                iq[q] = in[q];
                hw[q] = hwn[q];
                hn[q] = hnn[q];
                if constexpr (PLEV) {
                    lev[q] = lvn[q];
                } else {
                    // ordinal of the segment holding this lane's position,
                    // clamped for lanes past the tile's end
                    const int64_t pos = base0 + q * P + psub;
                    uint32_t ord = wpn[q] + (uint32_t)__popcll(hw[q] & lanes_le((int)(pos & 63))) - 1u;
                    ord = ord < (uint32_t)fe.nsegs ? ord : (uint32_t)fe.nsegs - 1u;
                    lev[q] = fe.seg_level[ord];
                }
            }
            // unconditional loads (positions past the tile read its first
            // position): every load of the step is in flight together
            T cv[U], uv[U], ov[U];
#pragma unroll
            for (int q = 0; q < U; q++) {
                const int64_t pos = base0 + q * P + psub;
                const int64_t pc = (pos >= start && pos < end) ? pos : start;
                cv[q] = RAW ? fe.lscale[lev[q]] : __builtin_nontemporal_load(fe.cache + pc);
                uv[q] = p.u[(int64_t)iq[q] * K + c];
                ov[q] = v[(int64_t)lev[q] * K + c];
            }
            meta(base0 + U * P);
            Acc val[U];
#pragma unroll
            for (int q = 0; q < U; q++) {
                const int64_t pos = base0 + q * P + psub;
                const Acc pr = RAW ? Acc(uv[q]) : mul_acc<Acc, T>(cv[q], uv[q]);
                val[q] = (pos >= start && pos < end && on) ? pr : Acc(0.f);
            }
            // every em[q] is defined before the scan, which stops at the
            // first step past the tile's end
            bool em[U];
#pragma unroll
            for (int q = 0; q < U; q++) em[q] = false;
#pragma unroll
            for (int q = 0; q < U; q++) {
                const int64_t base = base0 + q * P;
                if (base >= end) break;  // wave-uniform
                const int64_t wi = base >> 6;
                em[q] = seg_step<K>(clip(hw[q], wi), clip(hn[q], wi + 1), (int)(base & 63),
                                    base + psub >= start && base + psub < end, psub, c, on, val[q], cur,
                                    &p.tileA[(int64_t)t * K + c]);
            }
#pragma unroll
            for (int q = 0; q < U; q++) {
                if (em[q]) {
                    T out = a * (RAW ? cv[q] * acc_to_T<T, Acc>(val[q]) : acc_to_T<T, Acc>(val[q]));
                    if (keep) out = b * ov[q] + out;
                    v[(int64_t)lev[q] * K + c] = out;
                    sq += (double)out * (double)out;
                }
            }
        }
        if (cur.open && on && psub == 0) {
            if (cur.cont) p.tileA[(int64_t)t * K + c] = acc_to_double(cur.carry);
            else p.tileB[(int64_t)t * K + c] = acc_to_double(cur.carry);
        }
    }
    if (p.mode == MODE_NONE) return;
    __shared__ double red[kBlock];
    const double s = block_column_sum<K>(sq, red);
    if (threadIdx.x < K) p.partials[(int64_t)bid * K + threadIdx.x] = s;
}

template <typename T, typename Acc, int K, bool RAW = false>
__global__ __launch_bounds__(kBlock) void k_dtx_seg(DtxSegParams<T> p) {
    if (p.all_done && *p.all_done) return;
    dtx_seg_body<T, Acc, K, RAW>(p, blockIdx.x, gridDim.x);
}

// FEs with few levels: one wave per chunk of observations (internal order,
// fully sequential reads); every lane owns a private LDS accumulator column
// per level, so adds never collide.  The WPB waves of a workgroup cover
// consecutive chunks; their level sums are combined in wave order in LDS and
// written as one row per workgroup to dpart ([level][workgroup] rows of K),
// which the fix-up launch sums in workgroup order.
template <typename T>
struct DenseFE {
    const int32_t *refs;
    const T *cache;
    int32_t nlev, blk0, nblk, pad;
    int64_t coef_off;
    int64_t dpart_row0;  // first row (of K doubles) of this FE in dpart
};

template <typename T>
struct DtxDenseParams {
    int64_t N;
    int32_t ndense, chunk, pad0, pad1;
    DenseFE<T> fe[kMaxG];
    const T *u;
    double *dpart;
    const ColOps *ops;
    const int32_t *all_done;
};

// Few-level FE accumulators of the fused pass and of k_dtx_dense: per level
// and column, added to in ordered phases, for FP64 batches of 16 or more
// columns (P <= 4 phases); lane-private otherwise.  The host sizes the LDS
// with the same rule.
template <int K, typename Acc>
__host__ __device__ constexpr bool dense_phased() {
    return K >= 16 && std::is_same<Acc, double>::value;
}

// Columns per lane of the fused pass's wide lane map (k_fused, C): batches
// of 16 or more columns; 1 = the plain map.  The fused pass then sums
// few-level FEs in phased per-column FP64 accumulators whatever the
// precision.  The host uses it unless FEHIP_WIDE=0 (fused_cols in fehip.hip
// applies the same rule).  Measured on the fused pass (FP64, default
// schedule): the factored piece form at K = 32 (N = 1e7) 1.91 ms with C = 1,
// 1.44 ms with C = 2 and two wave steps in flight, 1.59 ms with C = 4; the
// cache form at K = 16 (N = 1e8, an interacted FE) 11.97 ms with C = 2, 8.13
// ms with C = 4; the FP32 streams at K = 32 1.30 ms with C = 2, 1.07 ms with
// C = 4.  C = 8, or C = 4 with two steps, spill registers (2.8-6.4 ms).
#ifndef FUSED_WIDE_C
#define FUSED_WIDE_C 2
#endif
#ifndef FUSED_WIDE_C_CACHE
#define FUSED_WIDE_C_CACHE 4
#endif
#ifndef FUSED_WIDE_C32
#define FUSED_WIDE_C32 4
#endif
template <typename Acc, int K, bool FACT>
__host__ __device__ constexpr int fused_wide_c() {
    return K < 16                              ? 1
           : !std::is_same<Acc, double>::value ? FUSED_WIDE_C32
           : FACT                              ? FUSED_WIDE_C
                                               : FUSED_WIDE_C_CACHE;
}

// Lane-private few-level accumulators (the form that is not phased): each
// sums the positions of one lane and level within a chunk, a small share of
// the level, and the lanes' sums are combined in FP64.  FP32 modes keep them
// as plain floats, half the LDS of a compensated pair.  The share is below
// chunk/64 positions, so the rounding stays far below that of one FP32 sum
// over the whole level.
template <typename Acc>
struct LaneAccOf {
    using type = Acc;
};
template <>
struct LaneAccOf<F32x2> {
    using type = float;
};
template <int K, typename Acc>
using DenseAcc = typename std::conditional<dense_phased<K, Acc>(), Acc, typename LaneAccOf<Acc>::type>::type;

// One phased add of U steps into per-wave accumulators acc[slot*K + c] (FP64):
// the lanes of phase ph (psub == ph) hold distinct columns, so no two lanes of
// an instruction meet, and the adds go out phase by phase, steps in order
// within a phase.  The phase loop is kept rolled: unrolled, the compiler
// merged the P guarded adds into one unguarded instruction at K = 32 (lanes
// c and 32 + c then meet in one ds_add_f64, whose order among them the
// hardware picks), so only program order fixes the summation order here.
template <int U, int K, typename Acc>
__device__ __forceinline__ void phased_add(Acc *acc, const int32_t (&slot)[U], const Acc (&x)[U], int psub, int c) {
    static_assert(std::is_same<Acc, double>::value, "phased accumulators are FP64 only");
    constexpr int P = kWave / K;
#pragma unroll 1
    for (int ph = 0; ph < P; ph++) {
        if (psub == ph) {
#pragma unroll
            for (int q = 0; q < U; q++) atomicAdd(acc + slot[q] * K + c, x[q]);
        }
    }
}

// phased_add for C columns per lane (KL = K/C lanes per position, P = 64/KL
// phases): the lanes of a phase hold distinct column groups, so again no two
// lanes of an instruction meet; within a phase, steps in order, then columns.
// FP64 accumulators in every precision (k_fused).
template <int U, int K, int C, typename Acc>
__device__ __forceinline__ void phased_add_c(Acc *acc, const int32_t (&slot)[U], const Acc (&x)[U][C], int psub,
                                             int c0) {
    static_assert(std::is_same<Acc, double>::value, "FP64 LDS adds");
    constexpr int P = kWave / (K / C);
#pragma unroll 1
    for (int ph = 0; ph < P; ph++) {
        if (psub == ph) {
#pragma unroll
            for (int q = 0; q < U; q++)
#pragma unroll
                for (int j = 0; j < C; j++) atomicAdd(acc + slot[q] * K + c0 + j, x[q][j]);
        }
    }
}


// RAW: sums of the stored sqrt(w) .* u without the cache (the fix-up scales them)
template <typename T, typename Acc, int K, bool RAW = false>
__global__ __launch_bounds__(kBlock) void k_dtx_dense(DtxDenseParams<T> p) {
    if (p.all_done && *p.all_done) return;
    extern __shared__ __align__(16) unsigned char dense_smem[];
    constexpr int P = kWave / K;
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int wpb = blockDim.x / kWave;
    const int psub = lane / K, c = lane % K;
    int f = 0;
    while (f + 1 < p.ndense && (int)blockIdx.x >= p.fe[f + 1].blk0) f++;
    const DenseFE<T> fe = p.fe[f];
    const int bl = blockIdx.x - fe.blk0;
    constexpr bool PHASED = dense_phased<K, Acc>();  // per-column accumulators, as in k_fused
    using DA = DenseAcc<K, Acc>;
    DA *acc = reinterpret_cast<DA *>(dense_smem) + (size_t)wave * fe.nlev * (PHASED ? K : kWave);
    if constexpr (PHASED) {
        for (int e = lane; e < fe.nlev * K; e += kWave) acc[e] = DA(0.f);
    } else {
        for (int k = 0; k < fe.nlev; k++) acc[k * kWave + lane] = DA(0.f);
    }
    const int on = p.ops->on[c];
    const int64_t start = ((int64_t)bl * wpb + wave) * p.chunk;
    const int64_t end = start + p.chunk < p.N ? start + p.chunk : p.N;
    if (on && start < end) {
        constexpr int U = 4;
        for (int64_t base0 = start; base0 < end; base0 += U * P) {
            int32_t r[U];
            DA x[U];
#pragma unroll
            for (int q = 0; q < U; q++) {
                const int64_t pos = base0 + q * P + psub;
                const int64_t pc = pos < end ? pos : start;
                r[q] = __builtin_nontemporal_load(fe.refs + pc);
                const T uv = p.u[pc * K + c];
                if constexpr (RAW) {
                    x[q] = pos < end ? DA(uv) : DA(0.f);
                } else {
                    const T cv = __builtin_nontemporal_load(fe.cache + pc);
                    x[q] = pos < end ? mul_acc<DA, T>(cv, uv) : DA(0.f);
                }
            }
            if constexpr (PHASED) phased_add<U, K, DA>(acc, r, x, psub, c);
            else lds_accumulate<U, DA>(acc, r, x, lane);
        }
    }
    if constexpr (!PHASED) {
        // Per-wave level sums over the lanes of each column.  The double result
        // of level k goes to slot k*K + c of the wave's own region, which only
        // overlaps accumulators of levels <= k, already consumed (K <= 32).
        double *wsum = reinterpret_cast<double *>(acc);
        for (int k = 0; k < fe.nlev; k++) {
            double x = acc_to_double(acc[k * kWave + lane]);
            x = column_allreduce<K>(x);
            if (psub == 0) wsum[k * K + c] = x;
        }
    }
    __syncthreads();
    // combine the waves in order; thread t handles entries t, t + blockDim, ...
    double *out = p.dpart + fe.dpart_row0 * K;
    const size_t region = (size_t)fe.nlev * (PHASED ? K : kWave) * sizeof(DA) / sizeof(double);
    const double *base_ws = reinterpret_cast<const double *>(dense_smem);
    for (int e = threadIdx.x; e < fe.nlev * K; e += blockDim.x) {
        double s = 0.0;
        for (int w = 0; w < wpb; w++) s += base_ws[w * region + e];
        const int k = e / K, cc = e % K;
        out[((int64_t)k * fe.nblk + bl) * K + cc] = s;
    }
}

// ---------------------------------------------------------------------------
// One pass over the internal order for an LSMR iteration:
//   u <- b*u + a*A v                       (D*x of every FE, as k_db)
//   wsum[k] = sum_{i in level k of FE s} cache_s[i]*u[i]   (raw D_s'u)
//   dense FEs: per-workgroup level sums of cache_g*u  -> dpart rows
//   ||u||^2 per workgroup                  -> partials
// D'u is linear, so its raw sums can be formed before beta = ||u|| is known;
// the fix-up launch applies v <- a*wsum + b*v once beta is.  This reads the
// index data, caches and u once per iteration for FE s and the dense FEs
// instead of three times.  One wave per tile of R positions (tile t = wave t
// of the grid); tiles crossing FE s segments leave tileA/tileB pieces exactly
// as k_dtx_seg with tp = R does, and the dense rows are those k_dtx_dense
// writes with chunk = R and 4 waves per workgroup, so one set of fix-up items
// serves both paths.
// ---------------------------------------------------------------------------
template <typename T>
struct FusedParams {
    int64_t N, R;
    int32_t G, s, ntiles, ndense, nslot;
    // PK kernels: levels of the few-level FE at position GT-1 whose gathered
    // vector gsrc[GT-1] (nvl x K) the workgroup stages in LDS after the
    // accumulators
    int32_t nvl;
    const int32_t *refs[kMaxG];
    const T *cache[kMaxG];  // internal order
    int64_t coef_off[kMaxG];
    int32_t dense_g[kMaxG];      // dense FE ordinal -> FE index
    int32_t dense_slot0[kMaxG];  // its first LDS slot
    int64_t dense_row0[kMaxG];   // its first dpart row
    const uint64_t *heads;       // FE s
    const uint32_t *wprefix;
    const int32_t *seg_level;    // ordinal of an observed FE s level -> level (loads stay unconditional)
    int32_t nsegs_s;             // observed levels of FE s
    int32_t packed;              // GT = 3: refs[1] == refs[2] hold ref_1 | ref_2 << 26
    const T *coef;     // v
    T *target;         // u
    // per FE: the vector gathered at ref_g (scale .* v under FACT, else v,
    // from the FE's first level); the projected iteration points FE s at
    // its correction vector instead
    const T *gsrc[kMaxG];
    // PM = 1: the FE s correction -scale_s^2 * sum (factored) or -sum of each
    // segment inside a tile goes straight to zsw (levels of FE s x K);
    // gsrc[0] is then the per-level scale_s^2 (read at the level, not x K)
    T *zsw;
    // PM = 3: heads/wprefix/seg_level describe pieces (runs of one (FE s,
    // FE t) level pair inside a tile; seg_level maps a piece to its FE s
    // level), and each piece's sum of sqrt(w) .* u goes to psum[piece]
    T *psum;
    // FACT: cache_g = sqrt(w) * scale_g[ref] for every FE (no interactions),
    // so D*v = sqrt(w) * sum_g vt_g[ref_g] with vt = scale .* v (kept by the
    // vector update) and the raw D'u sums are of sqrt(w)*u (scaled in the
    // fix-up); the caches are not read
    const T *sqrtw;    // internal order; nullptr = unit weights
    const T *vt;       // p x K
    double *wsum;      // p_s x K
    double *tileA, *tileB;
    double *dpart;
    const ColOps *ops;  // the D*x coefficients
    double *partials;   // gridDim.x x K
    const int32_t *all_done;
};

#ifndef FUSED_U64
#define FUSED_U64 2
#endif
#ifndef FUSED_U_C2
#define FUSED_U_C2 2
#endif
// waves per SIMD the fused pass is compiled for (VGPR budget 512 / FUSED_WPE)
#ifndef FUSED_WPE
#define FUSED_WPE 4
#endif
#ifndef FUSED_U_C4
#define FUSED_U_C4 1
#endif
#ifndef FUSED_U_C4F
#define FUSED_U_C4F 1
#endif
// GT > 0 (the FE count, compile-time): a three-stage software pipeline per
// wave.  While loop step j is computed, the loads of step j+1 (u, caches,
// head words and the coefficient gathers, whose refs arrived one step
// earlier) and the refs of step j+2 are in flight.  GT = 0: any FE count,
// one FE at a time, no pipelining.
// Loads and stores at a 32-bit byte offset from a wave-uniform base (the
// scalar-base + 32-bit-offset addressing form): the host keeps every offset
// below 2^32 (tile of R positions: R*K*8; coefficient vectors: p*K*8).
template <typename V>
__device__ __forceinline__ V ldo(const V *base, uint32_t i) {
    return *reinterpret_cast<const V *>(reinterpret_cast<const char *>(base) + i * (uint32_t)sizeof(V));
}
template <typename V>
__device__ __forceinline__ void sto(V *base, uint32_t i, V v) {
    *reinterpret_cast<V *>(reinterpret_cast<char *>(base) + i * (uint32_t)sizeof(V)) = v;
}

// The factored iteration with fewer than 8 columns stores u as sqrt(w) .* u:
// the CSR and few-level D'u passes then read no cache (80 MB per column at
// N = 1e7).  Wider batches keep u (their fused pass is VALU-bound, and the
// division back to u would cost more than the cache read saves).  The host
// applies the same rule (raw D'u forms, fix-up scales, the conversion at
// LSMR start).
template <int K, bool FACT>
__host__ __device__ constexpr bool stores_scaled_u() {
    return FACT && K < 8;
}

// FE order (host-arranged): FE s is index 0; with GT > 0 and few-level FEs in
// the pass, the first of them (dense_g[0]) is index GT-1 -- compile-time
// positions, so no load sits behind a runtime FE test.
// PM (projected iteration, FE s eliminated): 0 = plain pass; 1 = sums only:
// the FE s segment sums of cache_s * (D*v over the other FEs), no u load,
// store or norm, no dense sums; 2 = the u pass without FE s segment sums.
// PM = 3: the plain pass in the piece form (GT > 0; factored, or the cache
// form with FE s and FE t not interacted): the segments are pieces, every
// tile starts one, and each piece's raw sum of sqrt(w) .* u goes to psum (its
// ordinal in position order); u is stored as it is.  PM = 4: the
// projected iteration's u pass (as PM = 2) in the piece form (sums as PM = 3,
// for the FE t half of the piece launch).
// C (columns per lane, wide lane map, seg_step_c): 1, or 2/4 for FP64
// batches of 16 or more columns with compile-time FE positions, whose
// few-level sums are phased.
// PK (GT = 3, the launch's packed flag set, nvl > 0): the packed ref stream
// is loaded once and split at the level stage, and the few-level FE's
// gathered vector is read from LDS.
template <typename T, typename Acc, int K, int GT, bool FACT, int PM = 0, int C = 1, bool PK = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(FUSED_WPE))) void k_fused(FusedParams<T> p) {
    if (*p.all_done) return;
    extern __shared__ __align__(16) unsigned char fused_smem[];
    static_assert(K % C == 0, "C divides K");
    constexpr int KL = K / C;  // lanes per position
    constexpr int P = kWave / KL;
    // wave steps per loop step (their loads are in flight together); FP64
    // and batched columns take fewer to stay within 128 VGPRs (4 waves per
    // SIMD) without spills (measured: U = 4 is slower for FP64 even without
    // spills); a wide lane map holds C columns' values per step
    constexpr int U = C >= 4 ? (sizeof(T) == 8 ? FUSED_U_C4 : FUSED_U_C4F)
                      : C == 2 ? FUSED_U_C2 : (sizeof(T) == 8 || K > 1) ? FUSED_U64 : 4;
    constexpr uint32_t UP = U * P;
    // the head-word indexing (jw = q*P/64 per step, NW words per loop step)
    // needs a loop step to cover whole words or to divide one
    static_assert(UP % kWave == 0 || kWave % UP == 0, "U * P: whole words or a divisor of one");
    // head words one loop step touches: steps q use word (q*P)/64 and the next
    constexpr int NW = UP / kWave + 1 > 2 ? UP / kWave + 1 : 2;
    constexpr int GA = GT > 0 ? GT : 1;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);  // wave-uniform
    const int psub = lane / KL, cg = lane % KL;
    const int c0 = cg * C;  // the lane's first column
    bool on[C], keep[C];
    T a[C], b[C];
    bool anyon = false;
#pragma unroll
    for (int j = 0; j < C; j++) {
        const double bj = p.ops->b[c0 + j];
        on[j] = p.ops->on[c0 + j] != 0;
        a[j] = (T)p.ops->a[c0 + j];
        b[j] = (T)bj;
        keep[j] = bj != 0.0;
        anyon = anyon || on[j];
    }
    // few-level FE sums: lane-private accumulators (nslot x 64 per wave), or
    // for wide batches in FP64 (dense_phased) and in the wide lane map one
    // accumulator per level and column (nslot x K per wave) that the P lanes
    // of a column add to in P ordered phases with LDS adds -- a quarter to a
    // half of the LDS
    constexpr bool PHASED = dense_phased<K, Acc>() || C > 1;
    constexpr bool PIECES = PM == 3 || PM == 4;
    constexpr bool UT = stores_scaled_u<K, FACT>() && !PIECES;
    static_assert(!PIECES || GT > 0, "the piece forms have compile-time FE positions");
    static_assert(C == 1 || (GT > 0 && !UT), "the wide lane map: compile-time FE positions, u stored as it is");
    // the wide map's phased accumulators are FP64 in every precision (the
    // exact products of floats summed in doubles): measured on MI355X, LDS
    // float atomic adds (ds_add_f32) cost 1.4 ms per K = 32 pass where the
    // same adds in FP64 (ds_add_f64) cost about 30 us
    using DA = typename std::conditional<(C > 1), double, DenseAcc<K, Acc>>::type;
    using VT = VecC<T, C>;
    // segment (FE s level or piece) sums: the mixed mode's FP32 phase adds
    // the exact products of its floats in FP64 (full-rate FP64 arithmetic,
    // fewer instructions than a compensated pair and at least as accurate)
    using SA = typename std::conditional<std::is_same<Acc, F32x2>::value, double, Acc>::type;
    DA *acc = reinterpret_cast<DA *>(fused_smem) + (size_t)wave * p.nslot * (PHASED ? K : kWave);
    if constexpr (PHASED) {
        for (int e = lane; e < p.nslot * K; e += kWave) acc[e] = DA(0.f);
    } else {
        for (int k = 0; k < p.nslot; k++) acc[k * kWave + lane] = DA(0.f);
    }
    // PK: the few-level FE's vector (nvl x K) staged after the accumulators;
    // its gathers become LDS reads.  Every wave of the workgroup reaches the
    // barrier (before the tile test).
    T *vl = reinterpret_cast<T *>(
        fused_smem + (((size_t)kWavesPerBlock * p.nslot * (PHASED ? K : kWave) * sizeof(DA) + 15) & ~(size_t)15));
    if constexpr (PK) {
        const T *src = p.gsrc[GA - 1];
        for (int e = threadIdx.x; e < p.nvl * K; e += kBlock) vl[e] = src[e];
        __syncthreads();
    }
    auto dense_add = [&](const int32_t (&sl)[U], const DA (&x)[U][C]) {
        if constexpr (C > 1) {
            phased_add_c<U, K, C, DA>(acc, sl, x, psub, c0);
        } else {
            DA x1[U];
#pragma unroll
            for (int q = 0; q < U; q++) x1[q] = x[q][0];
            if constexpr (PHASED) phased_add<U, K, DA>(acc, sl, x1, psub, c0);
            else lds_accumulate<U, DA>(acc, sl, x1, lane);
        }
    };
    const int t = blockIdx.x * kWavesPerBlock + wave;
    double sq[C];
#pragma unroll
    for (int j = 0; j < C; j++) sq[j] = 0.0;
    if (t < p.ntiles) {  // wave-uniform
        const int64_t start = (int64_t)t * p.R;
        const int64_t end = start + p.R < p.N ? start + p.R : p.N;
        const uint32_t len = (uint32_t)(end - start);
        // tile-relative bases; positions below are offsets from start
        T *ub = p.target + start * K;
        // sqrt(w): the factored form's multiplier, and the piece sums' in the
        // cache form (FE s and FE t carry no interaction there, so a piece's
        // sum of sqrt(w) .* u serves both, each scaled later)
        const T *swb = ((FACT || PIECES) && p.sqrtw) ? p.sqrtw + start : nullptr;
        const T *swl = swb ? swb : ub;  // any in-bounds source (offsets below len)
        SegCursorC<SA, C> cur;
#pragma unroll
        for (int j = 0; j < C; j++) cur.carry[j] = SA(0.f);
        cur.cont = true;
        cur.open = false;
        // one loop step, ready to compute: D*v sums, old u, FE s and first
        // dense FE cache/ref, head words
        struct Step {
            T sum[U][C], uo[U][C], cs[U], cd[U], g0[U];
            int32_t rs[U], rd[U];
            uint64_t hw[NW];
        };
        // offset of lane, step q of the loop step at lb, clamped into the
        // tile (loads are unconditional, so no load waits inside a branch)
        auto loc = [&](uint32_t lb, int q) -> uint32_t {
            const uint32_t o = lb + (uint32_t)(q * P + psub);
            return o < len ? o : 0u;
        };
        // start is a multiple of 256 and lb of U*P; the words up to
        // (end-1)/64 + 1 exist (the head bits run one word past N)
        auto load_words = [&](uint32_t lb, uint64_t (&hw)[NW]) {
            const int64_t b0 = start + lb;
#pragma unroll
            for (int j = 0; j < NW; j++) {
                const int64_t wi = (b0 >> 6) + j;
                hw[j] = (lb < len && wi <= ((end - 1) >> 6) + 1) ? ld_const(p.heads, wi) : 0ull;
            }
        };
        // FE s level of lane position lb + q*P + psub: the ordinal of its
        // segment (heads up to it, minus one), clamped for positions past N,
        // through seg_level when some level has no observation.  Replaces a
        // load of ref_s: positions are sorted by FE s.
        auto level_s = [&](uint32_t lb, int q, const uint64_t (&hw)[NW], const uint32_t (&wp)[NW],
                           int32_t &ordv) -> int32_t {
            const uint32_t ob = lb + (uint32_t)(q * P);
            const int jw = (q * P) / kWave;  // constant after unrolling
            const int bit = (int)((ob + (uint32_t)psub) & 63u);
            uint32_t ord = wp[jw] + (uint32_t)__popcll(hw[jw] & lanes_le(bit)) - 1u;
            ord = ord < (uint32_t)p.nsegs_s ? ord : (uint32_t)p.nsegs_s - 1u;
            ordv = (int32_t)ord;
            return p.seg_level[ord];
        };
        auto compute = [&](const Step &st, uint32_t lb) {
            T un[U][C];
#pragma unroll
            for (int q = 0; q < U; q++) {
                const uint32_t o = lb + (uint32_t)(q * P + psub);
                VT w;
#pragma unroll
                for (int j = 0; j < C; j++) {
                    const bool ok = on[j] && o < len;
                    T out = a[j] * st.sum[q][j];
                    // UT: the stored vector is sqrt(w) .* u (st.cs = sqrt(w)),
                    // which the D'u passes then sum without caches
                    if constexpr (PM != 1)
                        out = keep[j] ? b[j] * (UT ? st.uo[q][j] / st.cs[q] : st.uo[q][j]) + out : out;
                    un[q][j] = ok ? out : (T)0;
                    const T sv = UT ? st.cs[q] * out : out;
                    if constexpr (PM != 1) {
                        if constexpr (C == 1) {
                            if (ok) sto(ub, o * K + c0, sv);
                        } else {
                            w.v[j] = ok ? sv : st.uo[q][j];  // a column that is off keeps its old value
                        }
                        sq[j] += (double)un[q][j] * (double)un[q][j];
                    }
                }
                if constexpr (PM != 1 && C > 1) {
                    if (o < len) stv<C>(ub, o * K + c0, w);
                }
            }
            // FE s: segmented sums of cache_s * u; a segment ending at a
            // position belongs to that position's level
#pragma unroll
            for (int q = 0; q < (PM == 2 ? 0 : U); q++) {
                const uint32_t ob = lb + (uint32_t)(q * P);
                if (ob >= len) break;  // wave-uniform
                SA x[C];
#pragma unroll
                for (int j = 0; j < C; j++) x[j] = mul_acc<SA, T>(st.cs[q], un[q][j]);
                const int jw = (q * P) / kWave;  // constant after unrolling
                if (seg_step_c<KL, C>(st.hw[jw], st.hw[jw + 1], (int)(ob & 63u), ob + (uint32_t)psub < len, psub, cg,
                                      anyon, x, cur, &p.tileA[(int64_t)t * K + c0])) {
                    const uint32_t e = (uint32_t)st.rs[q] * K + c0;
                    if constexpr (PM == 1) {
                        VT z;
#pragma unroll
                        for (int j = 0; j < C; j++)
                            z.v[j] = (T)(-(FACT ? (double)st.g0[q] : 1.0) * acc_to_double(x[j]));
                        stv<C>(p.zsw, e, z);
                    } else if constexpr (PIECES) {
                        VT z;
#pragma unroll
                        for (int j = 0; j < C; j++) z.v[j] = acc_to_T<T, SA>(x[j]);
                        stv<C>(p.psum, e, z);
                    } else {
                        VecC<double, C> z;
#pragma unroll
                        for (int j = 0; j < C; j++) z.v[j] = acc_to_double(x[j]);
                        stv<C>(p.wsum, e, z);
                    }
                }
            }
            // dense FEs: lane-private LDS accumulators
            if (PM != 1 && p.ndense) {
                const int s0 = p.dense_slot0[0];
                int32_t sl[U];
                DA x[U][C];
#pragma unroll
                for (int q = 0; q < U; q++) {
                    sl[q] = s0 + st.rd[q];
#pragma unroll
                    for (int j = 0; j < C; j++) x[q][j] = mul_acc<DA, T>(st.cd[q], un[q][j]);
                }
                dense_add(sl, x);
                for (int d = 1; d < p.ndense; d++) {
                    const int32_t *rf = p.refs[p.dense_g[d]] + start;
                    const T *cf = p.cache[p.dense_g[d]] + start;
                    const int sd = p.dense_slot0[d];
#pragma unroll
                    for (int q = 0; q < U; q++) {
                        const uint32_t i = loc(lb, q);
                        sl[q] = sd + ldo(rf, i);
                        const T cv = FACT ? st.cs[q] : ldo(cf, i);
#pragma unroll
                        for (int j = 0; j < C; j++) x[q][j] = mul_acc<DA, T>(cv, un[q][j]);
                    }
                    dense_add(sl, x);
                }
            }
        };
        if constexpr (GT > 0) {
            // A four-stage software pipeline over loop steps, so that every
            // wait targets loads issued one loop step earlier (vector loads
            // complete in order, so a wait for a load also waits for every
            // older one):
            //   A  refs of step +3 (FE s: none, its level comes from the head
            //      words, loaded here with their prefixes)
            //   L  FE s levels of step +2 (seg_level loads; packed refs split)
            //   B  every other load of step +1 (gathers use its refs, levels)
            //   compute step 0
            struct Refs {
                int32_t r[GA][U];
                int32_t o[U];  // the piece forms: the piece ordinal of each position
                uint64_t hw[NW];
                uint32_t wp[NW];
            };
            auto issue_refs = [&](uint32_t lb, Refs &rr) {
#pragma unroll
                for (int g = 0; g < GT; g++) {
                    const int32_t *rb = p.refs[g] + start;
#pragma unroll
                    for (int q = 0; q < U; q++) rr.r[g][q] = (g == 0 || (PK && g == 2)) ? 0 : ldo(rb, loc(lb, q));
                }
                load_words(lb, rr.hw);
            };
            // heads before the first word of the next issue_level call (its
            // calls run in position order): the prefixes come from the
            // words' own counts, one scalar load per tile instead of one per
            // word (measured at K = 32: the fused pass 7% faster)
            uint32_t wpc = ld_const(p.wprefix, start >> 6);
            auto issue_level = [&](uint32_t lb, Refs &rr) {
                {
                    uint32_t acc = wpc;
#pragma unroll
                    for (int j = 0; j < NW; j++) {
                        rr.wp[j] = acc;
                        acc += (uint32_t)__popcll(rr.hw[j]);
                    }
                    const int adv = (int)(((start + lb + UP) >> 6) - ((start + lb) >> 6));
#pragma unroll
                    for (int j = 0; j < NW; j++)
                        if (j < adv) wpc += (uint32_t)__popcll(rr.hw[j]);
                }
#pragma unroll
                for (int q = 0; q < U; q++) rr.r[0][q] = level_s(lb, q, rr.hw, rr.wp, rr.o[q]);
                if constexpr (GT == 3 && PK) {
                    // one packed stream: ref_1 in the low 26 bits, ref_2 above
#pragma unroll
                    for (int q = 0; q < U; q++) {
                        const uint32_t raw = (uint32_t)rr.r[1][q];
                        rr.r[1][q] = (int32_t)(raw & 0x3FFFFFFu);
                        rr.r[2][q] = (int32_t)(raw >> 26);
                    }
                } else if constexpr (GT == 3) {
                    // packed pair: branch-free unpack (the same loads either way)
                    const uint32_t mask = p.packed ? 0x3FFFFFFu : 0xFFFFFFFFu;
                    const uint32_t shift = p.packed ? 26u : 0u;
#pragma unroll
                    for (int q = 0; q < U; q++) {
                        rr.r[1][q] = (int32_t)((uint32_t)rr.r[1][q] & mask);
                        rr.r[2][q] = (int32_t)((uint32_t)rr.r[2][q] >> shift);
                    }
                }
            };
            // a step's loads in flight, with the refs and head words that
            // compute needs (copied, so the Refs can move on)
            struct InFlight {
                VT uo[U], gv[GA][U];
                T cv[FACT ? 1 : GA][U], sw[U];
                int32_t rs[U], rd[U];
                uint64_t hw[NW];
            };
            auto issue_data = [&](uint32_t lb, const Refs &rr, InFlight &d) {
#pragma unroll
                for (int q = 0; q < U; q++) {
                    const uint32_t i = loc(lb, q);
                    if constexpr (PM == 1) {
#pragma unroll
                        for (int j = 0; j < C; j++) d.uo[q].v[j] = (T)0;
                    } else {
                        d.uo[q] = ldv<C>(ub, i * K + c0);
                    }
                    d.sw[q] = (T)1;
                    if constexpr (FACT || PIECES) {
                        // loaded either way (from u's tile under unit
                        // weights): a load in a branch makes the compiler's
                        // waits count it as absent, so the wait for the
                        // previous step also waited for this step's refs
                        const T x = ldo(swl, i);
                        d.sw[q] = swb ? x : (T)1;
                    }
                }
#pragma unroll
                for (int g = 0; g < GT; g++) {
                    const T *cf = p.gsrc[g];
#pragma unroll
                    for (int q = 0; q < U; q++) {
                        if (PM == 1 && g == 0) d.gv[g][q].v[0] = ldo(cf, (uint32_t)rr.r[g][q]);
                        else if (PK && g == GA - 1)
                            d.gv[g][q] = *reinterpret_cast<const VT *>(vl + (uint32_t)rr.r[g][q] * K + c0);
                        else d.gv[g][q] = ldv<C>(cf, (uint32_t)rr.r[g][q] * K + c0);
                        if constexpr (!FACT) d.cv[g][q] = ldo(p.cache[g] + start, loc(lb, q));
                    }
                }
#pragma unroll
                for (int q = 0; q < U; q++) {
                    d.rs[q] = PIECES ? rr.o[q] : rr.r[0][q];  // where a segment ending here is written
                    d.rd[q] = rr.r[GT - 1][q];
                }
#pragma unroll
                for (int j = 0; j < NW; j++) d.hw[j] = rr.hw[j];
            };
            // the step's loads have arrived: form the D*v sums (FEs in order)
            auto settle = [&](const InFlight &d, Step &st) {
#pragma unroll
                for (int q = 0; q < U; q++) {
#pragma unroll
                    for (int j = 0; j < C; j++) {
                        st.sum[q][j] = (T)0;
                        st.uo[q][j] = d.uo[q].v[j];
                    }
                    st.g0[q] = d.gv[0][q].v[0];
                    st.cs[q] = st.cd[q] = (T)0;
                    st.rs[q] = d.rs[q];
                    st.rd[q] = p.ndense ? d.rd[q] : 0;
                }
                if constexpr (FACT) {
#pragma unroll
                    for (int g = 0; g < GT; g++)
#pragma unroll
                        for (int q = 0; q < U; q++)
#pragma unroll
                            for (int j = 0; j < C; j++) st.sum[q][j] += (PM == 1 && g == 0) ? (T)0 : d.gv[g][q].v[j];
#pragma unroll
                    for (int q = 0; q < U; q++) {
#pragma unroll
                        for (int j = 0; j < C; j++) st.sum[q][j] = d.sw[q] * st.sum[q][j];
                        st.cs[q] = st.cd[q] = d.sw[q];
                    }
                } else {
#pragma unroll
                    for (int g = 0; g < GT; g++)
#pragma unroll
                        for (int q = 0; q < U; q++)
#pragma unroll
                            for (int j = 0; j < C; j++)
                                st.sum[q][j] += (PM == 1 && g == 0) ? (T)0 : d.cv[g][q] * d.gv[g][q].v[j];
#pragma unroll
                    for (int q = 0; q < U; q++) {
                        st.cs[q] = PIECES ? d.sw[q] : d.cv[0][q];
                        if (p.ndense) st.cd[q] = d.cv[GT - 1][q];
                    }
                }
#pragma unroll
                for (int j = 0; j < NW; j++) st.hw[j] = d.hw[j];
            };
            // prologue: data of step 0 in flight, levels of step 1 issued,
            // refs of step 2 issued
            Refs s0, s1, s2;
            InFlight d;
            Step st;
            issue_refs(0u, s0);
            issue_refs(UP, s1);
            issue_refs(2 * UP, s2);
            issue_level(0u, s0);
            issue_level(UP, s1);
            issue_data(0u, s0, d);
            // one loop step: refs of step +3 into ra, levels of step +2 from
            // rb, data of step +1 from rc
            auto step = [&](uint32_t lb, Refs &ra, Refs &rb, Refs &rc) {
                issue_refs(lb + 3 * UP, ra);
                issue_level(lb + 2 * UP, rb);
                settle(d, st);
                issue_data(lb + UP, rc, d);
                compute(st, lb);
            };
            if constexpr (C == 1) {
                // the three slots take the three roles in turn, the loop
                // unrolled three times instead of copying the slots along
                // (register moves every step): measured K = 1 fused pass
                // 77.9 -> 75.0 us; the wide map (C > 1) was slower so
                for (uint32_t lb = 0; lb < len; lb += 3 * UP) {
                    step(lb, s0, s2, s1);
                    if (lb + UP >= len) break;
                    step(lb + UP, s1, s0, s2);
                    if (lb + 2 * UP >= len) break;
                    step(lb + 2 * UP, s2, s1, s0);
                }
            } else {
                for (uint32_t lb = 0; lb < len; lb += UP) {
                    step(lb, s0, s2, s1);
                    s1 = s2;  // the slot of step +1 next
                    s2 = s0;  // of step +2 next
                }
            }
        } else {
            // any FE count, one FE at a time (C = 1)
            const int c = c0;
            for (uint32_t lb = 0; lb < len; lb += UP) {
                Step st;
#pragma unroll
                for (int q = 0; q < U; q++) {
                    st.sum[q][0] = (T)0;
                    st.g0[q] = (T)0;
                    st.uo[q][0] = PM == 1 ? (T)0 : ldo(ub, loc(lb, q) * K + c);
                    st.cs[q] = st.cd[q] = (T)0;
                    st.rs[q] = st.rd[q] = 0;
                }
                load_words(lb, st.hw);
                T sw[U];
#pragma unroll
                for (int q = 0; q < U; q++) {
                    sw[q] = (T)1;
                    if constexpr (FACT) {
                        if (swb) sw[q] = ldo(swb, loc(lb, q));  // wave-uniform test
                    }
                }
                for (int g = 0; g < p.G; g++) {
                    int32_t r[U];
                    T cv[U];
                    const int32_t *rb = p.refs[g] + start;
#pragma unroll
                    for (int q = 0; q < U; q++) {
                        r[q] = ldo(rb, loc(lb, q));
                        cv[q] = FACT ? (T)1 : ldo(p.cache[g] + start, loc(lb, q));
                    }
                    const T *cf = p.gsrc[g];
                    if (PM == 1 && g == p.s) {
#pragma unroll
                        for (int q = 0; q < U; q++) st.g0[q] = ldo(cf, (uint32_t)r[q]);
                    }
                    if (PM != 1 || g != p.s) {
#pragma unroll
                        for (int q = 0; q < U; q++) {
                            const T gv = ldo(cf, (uint32_t)r[q] * K + c);
                            st.sum[q][0] += FACT ? gv : cv[q] * gv;
                        }
                    }
                    if (g == p.s) {
#pragma unroll
                        for (int q = 0; q < U; q++) {
                            st.cs[q] = FACT ? sw[q] : cv[q];
                            st.rs[q] = r[q];
                        }
                    }
                    if (p.ndense && g == p.dense_g[0]) {
#pragma unroll
                        for (int q = 0; q < U; q++) {
                            st.cd[q] = FACT ? sw[q] : cv[q];
                            st.rd[q] = r[q];
                        }
                    }
                }
                if constexpr (FACT) {
#pragma unroll
                    for (int q = 0; q < U; q++) st.sum[q][0] = sw[q] * st.sum[q][0];
                }
                compute(st, lb);
            }
        }
        // (the piece forms: every piece ends inside its tile)
        if (PM != 2 && !PIECES && cur.open && anyon && psub == 0) {
#pragma unroll
            for (int j = 0; j < C; j++) {
                if (cur.cont) p.tileA[(int64_t)t * K + c0 + j] = acc_to_double(cur.carry[j]);
                else p.tileB[(int64_t)t * K + c0 + j] = acc_to_double(cur.carry[j]);
            }
        }
    }
    if (PM != 1 && p.ndense) {
        // per-wave slot sums (written in place, see k_dtx_dense), then the
        // waves of the workgroup in order -> one dpart row per workgroup
        if constexpr (!PHASED) {
            double *ws = reinterpret_cast<double *>(acc);
            for (int k = 0; k < p.nslot; k++) {
                double x = acc_to_double(acc[k * kWave + lane]);
                x = column_allreduce<K>(x);
                if (psub == 0) ws[k * K + c0] = x;
            }
        }
        __syncthreads();
        // phased: the waves' accumulators as they are (DA); lane-private: the
        // per-wave doubles written in place above
        const size_t region = PHASED ? (size_t)p.nslot * K : (size_t)p.nslot * kWave * sizeof(DA) / sizeof(double);
        const DA *base_acc = reinterpret_cast<const DA *>(fused_smem);
        const double *base_ws = reinterpret_cast<const double *>(fused_smem);
        for (int e = threadIdx.x; e < p.nslot * K; e += blockDim.x) {
            double s = 0.0;
            for (int w = 0; w < kWavesPerBlock; w++)
                s += PHASED ? acc_to_double(base_acc[w * region + e]) : base_ws[w * region + e];
            const int slot = e / K, cc = e % K;
            int d = 0;
            while (d + 1 < p.ndense && slot >= p.dense_slot0[d + 1]) d++;
            const int k = slot - p.dense_slot0[d];
            p.dpart[(p.dense_row0[d] + (int64_t)k * gridDim.x + blockIdx.x) * K + cc] = s;
        }
    }
    __shared__ double red[kBlock];
#pragma unroll
    for (int j = 0; j < C; j++) {
        // column c0 + j: the lanes of column group cg, then the waves
        const double s = block_column_sum<KL>(sq[j], red);
        if (threadIdx.x < KL) p.partials[(int64_t)blockIdx.x * K + threadIdx.x * C + j] = s;
    }
}

// Fix-up items: kind 0 = a segment crossing tiles (total = tileB[first] +
// sum_{q=1..count} tileA[first+q]); kind 1 = a dense level (total = sum of
// count dpart rows from row first); kind 2 = a level no observation has
// (total = 0: only the b*coef scaling of rmul!, src/AbstractFixedEffectLinearMap.jl:30).
// One wave per item, fixed-order sums.
struct FixItem {
    int32_t kind, level;
    int64_t coef_off;
    int64_t first;
    int32_t count, fe;  // fe: the FE the level belongs to
};

template <typename T>
struct FixParams {
    int32_t nitems, pad;
    const FixItem *items;
    const double *tileA, *tileB, *dpart;
    T *v;
    const ColOps *ops;
    double *partials;  // gridDim.x x K
    const int32_t *all_done;
    int mode;
    // after k_fused: v_s[k] <- b*v_s[k] + a*wsum[k] for the levels of FE s
    // whose segment lies inside one tile (sdirect[k] = 1); nullptr = none
    const double *wsum;
    const uint8_t *sdirect;
    int64_t ns, s_off;
    int32_t s, pad2;
    // after a FACT fused pass: the sums of FE g's levels left by it (FE s
    // segments, few-level FEs) are of sqrt(w)*u and take scale_g here;
    // nullptr for FEs whose sums already carry their caches
    const T *fscale[kMaxG];
};

template <typename T, int K>
__device__ __forceinline__ void dtx_fixup_body(const FixParams<T> &p, int bid, int nblk) {
    constexpr int P = kWave / K;
    const int lane = threadIdx.x & (kWave - 1);
    const int psub = lane / K, c = lane % K;
    const int gw = bid * kWavesPerBlock + (threadIdx.x / kWave);
    const int nw = nblk * kWavesPerBlock;
    const int on = p.ops->on[c];
    const T a = (T)p.ops->a[c], b = (T)p.ops->b[c];
    const bool keep = p.ops->b[c] != 0.0;
    double sq = 0.0;
    if (p.wsum && on) {
        // stride is a multiple of K, so element e belongs to column c.  The
        // loads of kU elements are issued together, none waiting on sdirect
        // (every index is in range; values of levels not taken are dropped),
        // so a thread pays one memory round trip per kU elements.
        constexpr int kU = 4;
        const int64_t nsK = p.ns * K, stride = (int64_t)nblk * kBlock;
        const T *fs = p.fscale[p.s];
        T *vs = p.v + p.s_off * K;
        int64_t e = (int64_t)bid * kBlock + threadIdx.x;
        for (; e < nsK; e += kU * stride) {
            uint8_t dir[kU];
            double ws[kU];
            T f[kU], old[kU];
#pragma unroll
            for (int u = 0; u < kU; u++) {
                const int64_t q = e + u * stride < nsK ? e + u * stride : e;
                dir[u] = e + u * stride < nsK ? p.sdirect[q / K] : 0;
                ws[u] = p.wsum[q];
                f[u] = fs ? fs[q / K] : (T)1;
                old[u] = keep ? vs[q] : (T)0;
            }
#pragma unroll
            for (int u = 0; u < kU; u++) {
                if (!dir[u]) continue;
                const double tot = fs ? (double)f[u] * ws[u] : ws[u];
                T out = a * (T)tot;
                if (keep) out = b * old[u] + out;
                vs[e + u * stride] = out;
                sq += (double)out * (double)out;
            }
        }
    }
    for (int item = gw; item < p.nitems; item += nw) {
        const FixItem it = p.items[item];
        // the loads that do not depend on the sum go out first; each level
        // is one item's, so no other wave writes dst in this launch
        T *dst = p.v + (it.coef_off + it.level) * K + c;
        const T *fsp = it.kind != 2 ? p.fscale[it.fe] : nullptr;
        const double tb = on && it.kind == 0 ? p.tileB[it.first * K + c] : 0.0;
        const T fsv = on && fsp ? fsp[it.level] : (T)1;
        const T old = on && keep ? *dst : (T)0;
        double s = 0.0;
        if (on && it.kind != 2) {
            // rows [lo, hi) of `src`, lane-strided, four independent chains
            const double *src = it.kind == 0 ? p.tileA : p.dpart;
            const int64_t lo = it.kind == 0 ? it.first + 1 : it.first;
            const int64_t hi = lo + it.count;
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
            int64_t q = lo + psub;
            for (; q + 3 * P < hi; q += 4 * P) {
                s0 += src[q * K + c];
                s1 += src[(q + P) * K + c];
                s2 += src[(q + 2 * P) * K + c];
                s3 += src[(q + 3 * P) * K + c];
            }
            for (; q < hi; q += P) s0 += src[q * K + c];
            s = (s0 + s1) + (s2 + s3);
        }
        s = column_allreduce<K>(s);
        if (on && psub == 0) {
            if (it.kind == 0) s = tb + s;
            if (fsp) s = (double)fsv * s;
            T out = a * (T)s;
            if (keep) out = b * old + out;
            *dst = out;
            sq += (double)out * (double)out;
        }
    }
    if (p.mode == MODE_NONE) return;
    __shared__ double red[kBlock];
    const double s = block_column_sum<K>(sq, red);
    if (threadIdx.x < K) p.partials[(int64_t)bid * K + threadIdx.x] = s;
}

template <typename T, int K>
__global__ __launch_bounds__(kBlock) void k_dtx_fixup(FixParams<T> p) {
    if (p.all_done && *p.all_done) return;
    dtx_fixup_body<T, K>(p, blockIdx.x, gridDim.x);
}

// The fix-up items that need only the fused pass (FE s, few-level FEs) and
// the CSR pass in one launch: blocks [0, nfix) run the fix-up, the rest the
// CSR tiles.  The fix-up's short dependent loads then overlap the CSR
// pass's stream instead of running after it.  The two parts touch disjoint
// levels of v and leave their norm partials in separate arrays.
// The piece launch's tiles in the wide lane map (C adjacent columns per
// lane, KL = K/C lanes and P = 64/KL positions per wave step; seg_step_c):
// each position's order entry, level, head bits and scale are loaded once
// for C columns, the piece sums and each level's v move as C-element
// vectors.  RAW and PLEV (the piece form) only.  The sums of the FP32
// streams' piece sums are FP64 (exact for floats).  A level's v is written
// once (in the tile where its pieces end, when they do not cross tiles); a
// column that is off keeps its old value.
template <typename T, typename Acc, int K, int C>
__device__ __forceinline__ void dtx_seg_body_wide(const DtxSegParams<T> &p, int bid, int nblk) {
    constexpr int KL = K / C;
    constexpr int P = kWave / KL;
    constexpr int U = 2;  // wave steps whose loads are issued together
    using SA = typename std::conditional<std::is_same<Acc, F32x2>::value, double, Acc>::type;
    using VT = VecC<T, C>;
    const int lane = threadIdx.x & (kWave - 1);
    const int psub = lane / KL, cg = lane % KL, c0 = cg * C;
    bool on[C], keep[C];
    T a[C], b[C];
    bool anyon = false;
#pragma unroll
    for (int j = 0; j < C; j++) {
        const double bj = p.ops->b[c0 + j];
        on[j] = p.ops->on[c0 + j] != 0;
        a[j] = (T)p.ops->a[c0 + j];
        b[j] = (T)bj;
        keep[j] = bj != 0.0;
        anyon = anyon || on[j];
    }
    const int gw = bid * kWavesPerBlock + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);  // wave-uniform
    const int nw = nblk * kWavesPerBlock;
    double sq[C];
#pragma unroll
    for (int j = 0; j < C; j++) sq[j] = 0.0;
    for (int t = p.first_tile + gw; t < p.total_tiles; t += nw) {
        int f = 0;
        while (f + 1 < p.nseg && t >= p.fe[f + 1].tile0) f++;
        const SegFE<T> fe = p.fe[f];
        const int64_t start = fe.tstart[t - fe.tile0];
        const int64_t end = fe.tstart[t - fe.tile0 + 1];
        const int64_t wbeg = start & ~(int64_t)63;
        auto clip = [&](uint64_t w, int64_t wi) -> uint64_t {
            const int64_t lo = start - (wi << 6), hi = end - (wi << 6);
            if (hi < 0 || lo > 63) return 0ull;
            uint64_t m = lo > 0 ? ~0ull << lo : ~0ull;
            if (hi < 63) m &= (2ull << hi) - 1ull;
            return w & m;
        };
        T *v = p.v + fe.coef_off * K;
        SegCursorC<SA, C> cur;
#pragma unroll
        for (int j = 0; j < C; j++) cur.carry[j] = SA(0.f);
        cur.cont = true;
        cur.open = false;
        int32_t in[U], lvn[U];
        uint64_t hwn[U], hnn[U];
        auto meta = [&](int64_t b0) {
#pragma unroll
            for (int q = 0; q < U; q++) {
                const int64_t base = b0 + q * P;
                const int64_t pos = base + psub;
                const int64_t pc = (pos >= start && pos < end) ? pos : start;
                in[q] = __builtin_nontemporal_load(fe.perm + pc);
                lvn[q] = fe.plev[pc];
                hwn[q] = hnn[q] = 0;
                if (base < end) {  // wave-uniform
                    hwn[q] = ld_const(fe.heads, base >> 6);
                    hnn[q] = ld_const(fe.heads, (base >> 6) + 1);
                }
            }
        };
        meta(wbeg);
        for (int64_t base0 = wbeg; base0 < end; base0 += U * P) {
            int32_t iq[U], lev[U];
            uint64_t hw[U], hn[U];
#pragma unroll
            for (int q = 0; q < U; q++) {
                iq[q] = in[q];
                lev[q] = lvn[q];
                hw[q] = hwn[q];
                hn[q] = hnn[q];
            }
            // unconditional loads (positions past the tile read its first
            // position): every load of the step is in flight together
            T cv[U];
            VT uv[U], ov[U];
#pragma unroll
            for (int q = 0; q < U; q++) {
                cv[q] = fe.lscale[lev[q]];
                uv[q] = ldv<C>(p.u, (uint32_t)iq[q] * K + c0);
                ov[q] = ldv<C>(v, (uint32_t)lev[q] * K + c0);
            }
            meta(base0 + U * P);
            SA val[U][C];
#pragma unroll
            for (int q = 0; q < U; q++) {
                const int64_t pos = base0 + q * P + psub;
                const bool inb = pos >= start && pos < end;
#pragma unroll
                for (int j = 0; j < C; j++) val[q][j] = (inb && on[j]) ? SA(uv[q].v[j]) : SA(0.f);
            }
            bool em[U];
#pragma unroll
            for (int q = 0; q < U; q++) em[q] = false;
#pragma unroll
            for (int q = 0; q < U; q++) {
                const int64_t base = base0 + q * P;
                if (base >= end) break;  // wave-uniform
                const int64_t wi = base >> 6;
                em[q] = seg_step_c<KL, C>(clip(hw[q], wi), clip(hn[q], wi + 1), (int)(base & 63),
                                          base + psub >= start && base + psub < end, psub, cg, anyon, val[q], cur,
                                          &p.tileA[(int64_t)t * K + c0]);
            }
#pragma unroll
            for (int q = 0; q < U; q++) {
                if (em[q]) {
                    VT o;
#pragma unroll
                    for (int j = 0; j < C; j++) {
                        T out = a[j] * (cv[q] * acc_to_T<T, SA>(val[q][j]));
                        if (keep[j]) out = b[j] * ov[q].v[j] + out;
                        o.v[j] = on[j] ? out : ov[q].v[j];
                        if (on[j]) sq[j] += (double)out * (double)out;
                    }
                    stv<C>(v, (uint32_t)lev[q] * K + c0, o);
                }
            }
        }
        if (cur.open && anyon && psub == 0) {
#pragma unroll
            for (int j = 0; j < C; j++) {
                if (cur.cont) p.tileA[(int64_t)t * K + c0 + j] = acc_to_double(cur.carry[j]);
                else p.tileB[(int64_t)t * K + c0 + j] = acc_to_double(cur.carry[j]);
            }
        }
    }
    if (p.mode == MODE_NONE) return;
    __shared__ double red[kBlock];
#pragma unroll
    for (int j = 0; j < C; j++) {
        const double s = block_column_sum<KL>(sq[j], red);
        if (threadIdx.x < KL) p.partials[(int64_t)bid * K + threadIdx.x * C + j] = s;
    }
}

// C > 1: the tiles in the wide lane map (the piece form)
template <typename T, typename Acc, int K, bool RAW, bool PLEV = false, int C = 1>
__global__ __launch_bounds__(kBlock) void k_dtx_seg_fix(DtxSegParams<T> p, FixParams<T> f, int nfix) {
    static_assert(C == 1 || (RAW && PLEV), "the wide lane map serves the piece launch");
    if (p.all_done && *p.all_done) return;
    if ((int)blockIdx.x < nfix) dtx_fixup_body<T, K>(f, blockIdx.x, nfix);
    else if constexpr (C > 1) dtx_seg_body_wide<T, Acc, K, C>(p, (int)blockIdx.x - nfix, (int)gridDim.x - nfix);
    else dtx_seg_body<T, Acc, K, RAW, PLEV>(p, (int)blockIdx.x - nfix, (int)gridDim.x - nfix);
}

// ---------------------------------------------------------------------------
// Fused LSMR vector update (src/utils/lsmr.jl:110, :138-142, :192):
//   v <- v/alpha ; hbar <- c_hbar*hbar + h ; x <- x + c_x*hbar ;
//   h <- c_h*h + v ; ||x||^2 partials (the next beta finalize, k_fin_db, runs the
//   stopping rules on them).
// Replaces rmul!/axpy!/norm on FixedEffectCoefficients
// (src/FixedEffectCoefficients.jl:21-55).
// ---------------------------------------------------------------------------
template <typename T>
struct UpdParams {
    int64_t pK;
    T *v, *h, *hbar, *x;
    double *partials;
    const int32_t *all_done;
    const void *state;
    const T *scale;  // p: with vt, the fused pass's scale .* v (FACT)
    T *vt;
    // elements [skip0, skip0 + skipn) are left alone (the coefficients of
    // the eliminated FE in the projected iteration, identically zero);
    // both multiples of K
    int64_t skip0, skipn;
};

template <typename T, typename Tr, int K>
__global__ __launch_bounds__(kBlock) void k_update(UpdParams<T> p) {
    if (*p.all_done) return;
    const LsmrState<Tr> *st = (const LsmrState<Tr> *)p.state;
    const int c = threadIdx.x % K;  // stride is a multiple of K
    const bool act = !st[c].done;
    const bool vs = st[c].vskip != 0;
    const T ia = (T)st[c].inv_alpha, chb = (T)st[c].c_hbar, cx = (T)st[c].c_x, ch = (T)st[c].c_h;
    double sq = 0.0;
    if (act) {
        // kU grid-stride elements per trip, their loads issued together (a
        // thread's elements past n read element e0 and are not stored)
        constexpr int kU = 4;
        const int64_t n = p.pK - p.skipn, stride = (int64_t)gridDim.x * kBlock;
        for (int64_t e0 = (int64_t)blockIdx.x * kBlock + threadIdx.x; e0 < n; e0 += kU * stride) {
            int64_t ee[kU];
            T vv[kU], hv[kU], hbv[kU], xv[kU], sc[kU];
#pragma unroll
            for (int u = 0; u < kU; u++) {
                const int64_t e2 = e0 + u * stride < n ? e0 + u * stride : e0;
                ee[u] = e2 < p.skip0 ? e2 : e2 + p.skipn;
                vv[u] = p.v[ee[u]];
                hv[u] = p.h[ee[u]];
                hbv[u] = p.hbar[ee[u]];
                xv[u] = p.x[ee[u]];
                sc[u] = p.vt ? p.scale[ee[u] / K] : (T)0;
            }
#pragma unroll
            for (int u = 0; u < kU; u++) {
                if (e0 + u * stride >= n) continue;
                const int64_t e = ee[u];
                T vn = vv[u];
                if (!vs) {
                    vn = vn * ia;
                    p.v[e] = vn;
                }
                if (p.vt) p.vt[e] = sc[u] * vn;
                const T hb = hbv[u] * chb + hv[u];
                p.hbar[e] = hb;
                const T xn = xv[u] + cx * hb;
                p.x[e] = xn;
                p.h[e] = hv[u] * ch + vn;
                sq += (double)xn * (double)xn;
            }
        }
    }
    __shared__ double red[kBlock];
    const double s = block_column_sum<K>(sq, red);
    if (threadIdx.x < K) p.partials[(int64_t)blockIdx.x * K + threadIdx.x] = s;
}

// ---------------------------------------------------------------------------
// Single-workgroup finalize launches: sum a reduction's per-workgroup partials
// in index order and run the LSMR scalar step for each column.  A separate
// launch (instead of a last-workgroup election inside the reduction) keeps
// device-scope fences out of the streaming kernels: the kernel boundary
// publishes the partials.
// ---------------------------------------------------------------------------
// MODE_ITER: first the stopping test of the previous iteration (its ||x||
// partials from the vector update; the fused pass that ran since changes
// nothing the test or the solution reads), which may finish columns; then
// beta of this iteration for the columns still running.  Moving the test
// here saves a launch per iteration.
template <typename Tr, int K>
__global__ __launch_bounds__(kFinBlock) void k_fin_db(const double *partials, int64_t nparts, int mode, void *state,
                                                   ColOps *ops_next_dtx, const int32_t *all_done,
                                                   const double *upd_partials, int64_t nupd, ColOps *ops_db,
                                                   int32_t *all_done_w) {
    if (all_done && *all_done) return;
    __shared__ double red[2 * kFinBlock];
    LsmrState<Tr> *st = (LsmrState<Tr> *)state;
    double tot;
    if (mode == MODE_ITER && upd_partials) {
        double xsq;
        block_sum_partials_pair<K, kFinBlock>(upd_partials, nupd, partials, nparts, red, xsq, tot);
        if (threadIdx.x < K && st[threadIdx.x].pending) fin_after_update<Tr>(st[threadIdx.x], xsq, threadIdx.x, ops_db);
        __syncthreads();
        publish_all_done<K, Tr>(st, all_done_w);
    } else {
        tot = block_sum_partials<K, kFinBlock>(partials, nparts, nullptr, 0, red);
    }
    if (threadIdx.x < K) fin_after_db<Tr>(st[threadIdx.x], tot, mode, threadIdx.x, ops_next_dtx);
}

template <typename Tr, int K>
__global__ __launch_bounds__(kFinBlock) void k_fin_dtx(const double *p0, int64_t n0, const double *p1, int64_t n1,
                                                    int mode, void *state, ColOps *ops_next_db, int32_t *all_done_w,
                                                    const int32_t *all_done) {
    if (all_done && *all_done) return;
    __shared__ double red[kFinBlock];
    const double tot = block_sum_partials<K, kFinBlock>(p0, n0, p1, n1, red);
    LsmrState<Tr> *st = (LsmrState<Tr> *)state;
    if (threadIdx.x < K) fin_after_dtx<Tr>(st[threadIdx.x], tot, mode, threadIdx.x, ops_next_db);
    __syncthreads();
    if (mode == MODE_INIT) publish_all_done<K, Tr>(st, all_done_w);
}


// LSMR set-up of the p-length vectors (:68, :77-78): v /= alpha; h = v; hbar = 0.
template <typename T, typename Tr, int K>
__global__ __launch_bounds__(kBlock) void k_init_vectors(int64_t pK, T *v, T *h, T *hbar, const void *state,
                                                        const T *scale, T *vt) {
    const LsmrState<Tr> *st = (const LsmrState<Tr> *)state;
    const int c = threadIdx.x % K;
    const T ia = (T)st[c].inv_alpha;
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < pK; e += (int64_t)gridDim.x * kBlock) {
        const T vn = v[e] * ia;
        v[e] = vn;
        h[e] = vn;
        hbar[e] = (T)0;
        if (vt) vt[e] = scale[e / K] * vn;
    }
}

}  // namespace fehip
