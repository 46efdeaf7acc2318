// fehip.hip -- libfehip.so: host side of the MI355X fixed-effect solver and
// its C-ABI (include/fehip.h).
//
// One handle = one set of fixed effects on one GPU, with its own HIP stream.
// Set-up (fehip_create / fehip_update_weights) builds on the device:
//   * int32 0-based refs, range-checked;
//   * the stable CSR bucketing of every FE (radix sort, csr_build.h);
//   * the internal observation order (stable sort by the FE with most levels,
//     then by the largest other CSR FE; the former's levels numbered by their
//     main level of the latter) and, where it applies, the piece form (runs of
//     one level pair cut at the fused pass's tiles);
//   * scale (serial per level, in the reference's summation order) and cache.
// A solve (fehip_solve_residuals*) runs the whole LSMR loop on the device:
// each iteration is the fused pass (D*v, the new u, the raw D'u sums of the
// sorted and few-level FEs or of the pieces), the D'u launch of the other FEs
// (+ fix-up) and the fused vector update; the scalar recurrences run in a
// single-workgroup finalize launch after each reduction, so the host only
// replays a captured hipGraph of several iterations and polls one flag.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fehip.h"
#include "components.h"
#include "csr_build.h"
#include "grouping.h"
#include "fehip_common.h"
#include "linmap.h"

using namespace fehip;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

namespace fehip {
int set_error(int code, const std::string &msg) { return fail(code, msg); }
int mlp_predict_device(fehip_mlp *m, const float *X_dev, int64_t n, float *y_dev);  // mlp.hip
}  // namespace fehip

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(FEHIP_EHIP, std::string(#expr) + " failed: " + hipGetErrorString(e_)); \
    } while (0)

#define CHK(expr)                 \
    do {                          \
        int rc_ = (expr);         \
        if (rc_ != 0) return rc_; \
    } while (0)

// ---------------------------------------------------------------------------
// handle
// ---------------------------------------------------------------------------
enum FeKind { KIND_SORTED = 0, KIND_CSR = 1, KIND_DENSE = 2 };

struct FeInfo {
    int64_t n = 0;
    int kind = KIND_CSR;
    int32_t *ref_int = nullptr;    // internal order, 0-based
    int32_t *operm = nullptr;      // CSR over the caller's order (bit-exact bucketing)
    int32_t *ooff = nullptr;
    int32_t *iperm = nullptr;      // CSR over the internal order (KIND_CSR)
    int32_t *ioff = nullptr;
    uint64_t *heads = nullptr;
    uint32_t *wprefix = nullptr;
    int32_t *seg_level = nullptr;
    int64_t nwords = 0;
    double *z_orig = nullptr;      // interaction, caller's order (nullptr: UnitWeights)
    // per storage precision: [0] double, [1] float (both for FEHIP_MIXED)
    void *scale_T[2] = {nullptr, nullptr};    // n
    double *scale_d = nullptr;                // n, widened
    void *cache_int[2] = {nullptr, nullptr};  // N, internal order
    void *cache_csr[2] = {nullptr, nullptr};  // N, CSR order (== cache_int for KIND_SORTED)
    std::vector<int32_t> h_ioff;   // host copy of segment offsets (tile schedule)
    std::vector<int32_t> big_levels;  // levels with > kScaleBig observations
    int32_t *d_big = nullptr;
};

struct Plan;  // per-K buffers and schedule

struct fehip_solver {
    int64_t N = 0;
    int32_t G = 0;
    int64_t p = 0;
    int precision = FEHIP_FP64;
    bool has64 = true, has32 = false;  // storage precisions present
    double mixed_tol = 1e-6;           // FEHIP_MIXED: tolerance of the FP32-stream phase
    int s = 0;  // FE defining the internal order
    // FE s's levels in an internal numbering (renumber_s): every device array
    // indexed by an FE s level uses it; coefficient blocks, scales and the
    // CSR cross the C-ABI in the caller's numbering.  nullptr = identity.
    int32_t *s_int_of = nullptr;  // caller level -> internal level
    int32_t *s_ext_of = nullptr;  // internal level -> caller level
    std::vector<int32_t> h_s_int_of;
    double *s_tmp = nullptr;      // n_s doubles: a permuted FE s block
    std::vector<int64_t> n, off;
    hipStream_t stream = nullptr;
    std::vector<FeInfo> fe;
    int32_t *perm_s = nullptr;  // internal position -> caller's observation
    int32_t *pos_s = nullptr;   // caller's observation -> internal position (perm_s inverted)
    int32_t *ident = nullptr;   // 0..N-1: the "permutation" of the sorted FE (kernels load it unconditionally)
    // the fused pass's last two FEs in one ref stream (G = 3): the first
    // few-level FE (pack_d, <= 64 levels) in the top 6 bits over the other
    // FE (pack_g, < 2^26 levels); nullptr when the shape does not allow it
    int32_t *packed = nullptr;
    int pack_g = -1, pack_d = -1;
    double *w_orig = nullptr;   // caller's order (nullptr: UnitWeights)
    double *sqrtw_int = nullptr;
    // no FE has an interaction: cache_g = sqrt(w) * scale_g[ref], and the
    // fused pass reads sqrt(w) once instead of every cache (k_fused FACT)
    bool factored = false;
    void *sqrtw_T[2] = {nullptr, nullptr};    // N, internal order, [0] double (= sqrtw_int), [1] float
    void *scale_all[2] = {nullptr, nullptr};  // p, every FE's scale, [0] double, [1] float
    int32_t *d_flag = nullptr;  // range-check flag
    int32_t *h_flag = nullptr;  // pinned
    std::vector<void *> allocs;
    std::map<int, Plan *> plans;
    // sort scratch
    int32_t *sk[2] = {nullptr, nullptr}, *sv[2] = {nullptr, nullptr};
    uint32_t *hist = nullptr, *sums = nullptr;
    int64_t sort_cap = 0;
    int poll_iters = 8;
    // the iteration's CSR pass and the fused-pass fix-up items in one launch
    // (FEHIP_SEG_FIX=0: separate launches)
    bool seg_fix = true;
    // FP64 batches of 16 or more columns: the fused pass in the wide lane
    // map (several adjacent columns per lane, FUSED_WIDE_C)
    bool wide = true;
    // the projected iteration sums FE t from the piece form's sums where the
    // plan has one (FEHIP_PIECES_PROJ=0: the CSR pass)
    bool pieces_proj = true;
    // schedule overrides (fehip_set_schedule / fehip_autotune); 0 or -1 =
    // the default (environment, else sized by occupancy)
    int64_t sched_tiles_s = 0, sched_tiles_c = 0;
    int sched_fuse_dense = -1;
    hipEvent_t poll_ev[2] = {nullptr, nullptr};
    // second stream of the captured iteration: the FE s fix-up runs beside
    // the CSR D'u pass (fork after beta, join before alpha)
    hipStream_t stream2 = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    std::vector<double> last_info;  // LSMR scalars of the last solve, 9 per column
    // warm start: per-level degrees (lazy), feature / prediction scratch
    int32_t *degree = nullptr;   // p
    float *ws_feats = nullptr;   // p x 16
    float *ws_pred = nullptr;    // p
    double *ws_sig = nullptr;    // 1
    double *ws_coef = nullptr;   // p (host-provided initial coefficients)
    double *ws_x0 = nullptr;     // p x ws_x0_cols (the network's starts)
    int ws_x0_cols = 0;
    int t_sec = -1;              // FE giving the secondary order key (-1: none)
    // projected LSMR (FE s eliminated): requested by fehip_set_algorithm,
    // in effect for the residual solves (fwl_now is set per call)
    bool fwl_req = false, fwl_now = false;
    void *scale2_T[2] = {nullptr, nullptr};  // FE s: scale_s^2 (factored form), per storage precision
    // connected components for normalize! (lazy, see ensure_components)
    bool cc_ready = false;
    int cc_nu = 0;                // FEs without interaction
    std::vector<int> cc_role;     // per FE: -1 not rescaled, else position among them
    int32_t *cc_nodes = nullptr;  // p
    int32_t *cc_start = nullptr;  // p + 2
    uint32_t *cc_ncomp = nullptr; // 1
};

// With separate CSR and fix-up launches (FEHIP_SEG_FIX=0), from this batch
// width on the plain iteration's FE s fix-up runs on a second branch of the
// captured graph, beside the CSR pass.  Measured on N = 1e7: +4..7% at
// K = 32, while at K = 1 the cross-stream join costs about 18 us per
// iteration, more than the overlap gains.  The default merged launch
// (k_dtx_seg_fix) overlaps the two without a second branch, at every width.
constexpr int kSplitFixK = 16;

// the projected LSMR applies to the residual solves of a call (G >= 2)
struct FwlScope {
    fehip_solver *h;
    explicit FwlScope(fehip_solver *hh) : h(hh) { h->fwl_now = h->fwl_req && h->G >= 2; }
    ~FwlScope() { h->fwl_now = false; }
};

template <typename P>
static int dalloc(fehip_solver *h, P **ptr, size_t count) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(P));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(FEHIP_EOOM, "hipMalloc of " + std::to_string(count * sizeof(P)) + " bytes failed: " +
                                    hipGetErrorString(e));
    }
    h->allocs.push_back(q);
    *ptr = (P *)q;
    return 0;
}

static int dalloc_bytes(fehip_solver *h, void **ptr, size_t bytes) {
    return dalloc<unsigned char>(h, (unsigned char **)ptr, bytes);
}

static void dfree(fehip_solver *h, void *q) {
    if (!q) return;
    auto it = std::find(h->allocs.begin(), h->allocs.end(), q);
    if (it != h->allocs.end()) h->allocs.erase(it);
    (void)hipFree(q);
}

// integer tuning knob from the environment, clamped to [lo, hi]
static int64_t env_int(const char *name, int64_t dflt, int64_t lo, int64_t hi) {
    const char *e = getenv(name);
    if (!e || !*e) return dflt;
    const long long v = atoll(e);
    return v < lo ? lo : (v > hi ? hi : (int64_t)v);
}

static inline unsigned grid_for(int64_t n, int per_block) {
    int64_t g = (n + per_block - 1) / per_block;
    return (unsigned)std::max<int64_t>(g, 1);
}

// ---------------------------------------------------------------------------
// device sort: stable radix sort of (key, index) -> CSR
// ---------------------------------------------------------------------------
static int ensure_sort_scratch(fehip_solver *h, int64_t N) {
    if (h->sort_cap >= N) return 0;
    for (int b = 0; b < 2; b++) {
        dfree(h, h->sk[b]);
        dfree(h, h->sv[b]);
        CHK(dalloc(h, &h->sk[b], N));
        CHK(dalloc(h, &h->sv[b], N));
    }
    dfree(h, h->hist);
    dfree(h, h->sums);
    const int64_t ntiles = (N + kSortTile - 1) / kSortTile;
    CHK(dalloc(h, &h->hist, 256 * ntiles + 1));
    CHK(dalloc(h, &h->sums, (256 * ntiles + kSortTile - 1) / kSortTile + 1));
    h->sort_cap = N;
    return 0;
}

static void free_sort_scratch(fehip_solver *h) {
    for (int b = 0; b < 2; b++) {
        dfree(h, h->sk[b]);
        dfree(h, h->sv[b]);
        h->sk[b] = h->sv[b] = nullptr;
    }
    dfree(h, h->hist);
    dfree(h, h->sums);
    h->hist = h->sums = nullptr;
    h->sort_cap = 0;
}

// sums: (n + kSortTile - 1) / kSortTile + 1 entries
static int exclusive_scan_u32(hipStream_t st, const uint32_t *in, int64_t n, uint32_t *out, bool write_total,
                              uint32_t *sums) {
    const unsigned nb = grid_for(n, kSortTile);
    hipLaunchKernelGGL(k_scan_chunk_sums, dim3(nb), dim3(kBlock), 0, st, in, n, sums);
    hipLaunchKernelGGL(k_scan_sums_single, dim3(1), dim3(kBlock), 0, st, sums, (int64_t)nb);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kBlock), 0, st, in, n, sums, out, write_total ? 1 : 0);
    HIPCHK(hipGetLastError());
    return 0;
}
static int exclusive_scan_u32(fehip_solver *h, const uint32_t *in, int64_t n, uint32_t *out, bool write_total,
                              uint32_t *sums) {
    return exclusive_scan_u32(h->stream, in, n, out, write_total, sums);
}

// keys (0..nlev-1) -> sorted keys + the stably permuted values (vals_in, or
// the source index when vals_in is nullptr); vals_in must not alias vals_out
static int radix_sort(fehip_solver *h, const int32_t *keys, int64_t N, int64_t nlev, int32_t *keys_out,
                      int32_t *vals_out, const int32_t *vals_in = nullptr) {
    CHK(ensure_sort_scratch(h, N));
    int bits = 0;
    while ((int64_t(1) << bits) < nlev) bits++;
    const int passes = std::max(1, (bits + 7) / 8);
    const int64_t ntiles = (N + kSortTile - 1) / kSortTile;
    const int32_t *kin = keys, *vin = vals_in;
    for (int ps = 0; ps < passes; ps++) {
        const bool lastp = ps == passes - 1;
        int32_t *kout = lastp ? keys_out : h->sk[ps & 1];
        int32_t *vout = lastp ? vals_out : h->sv[ps & 1];
        hipLaunchKernelGGL(k_radix_hist<int32_t>, dim3((unsigned)ntiles), dim3(kBlock), 0, h->stream, kin, N, 8 * ps,
                           h->hist, ntiles);
        CHK(exclusive_scan_u32(h, h->hist, 256 * ntiles, h->hist, false, h->sums));
        hipLaunchKernelGGL(k_radix_scatter<int32_t>, dim3((unsigned)ntiles), dim3(kBlock), 0, h->stream, kin, vin, kout,
                           vout, N, 8 * ps, h->hist, ntiles);
        HIPCHK(hipGetLastError());
        kin = kout;
        vin = vout;
    }
    return 0;
}

// offsets, head bits, word prefix and segment->level map of a sorted key array
static int csr_products(fehip_solver *h, const int32_t *skeys, int64_t N, int64_t nlev, int32_t *off,
                        FeInfo *fi) {
    if (off) {
        hipLaunchKernelGGL(k_offsets_from_sorted, dim3(grid_for(N + 1, kBlock)), dim3(kBlock), 0, h->stream, skeys,
                           N, (int32_t)nlev, off);
        HIPCHK(hipGetLastError());
    }
    if (!fi) return 0;
    fi->nwords = N / 64 + 2;
    CHK(dalloc(h, &fi->heads, fi->nwords));
    CHK(dalloc(h, &fi->wprefix, fi->nwords + 1));
    CHK(dalloc(h, &fi->seg_level, nlev));
    uint32_t *cnt = nullptr;
    CHK(dalloc(h, &cnt, fi->nwords + 1));
    hipLaunchKernelGGL(k_head_bits, dim3(grid_for(fi->nwords, kWavesPerBlock)), dim3(kBlock), 0, h->stream, skeys,
                       N, fi->heads, fi->nwords, cnt);
    HIPCHK(hipGetLastError());
    uint32_t *sums = nullptr;
    CHK(dalloc(h, &sums, (fi->nwords + kSortTile - 1) / kSortTile + 2));
    CHK(exclusive_scan_u32(h, cnt, fi->nwords, fi->wprefix, true, sums));
    hipLaunchKernelGGL(k_seg_level, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, h->stream, skeys, N, fi->heads,
                       fi->wprefix, fi->seg_level);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    dfree(h, cnt);
    dfree(h, sums);
    return 0;
}


// FE s's internal numbering: its levels ordered by their main level of FE t
// (the t level holding the longest run of the s level's observations, ties
// to the smaller t; levels without observations last), then by caller
// number.  Workers sharing a main firm are then adjacent in the internal
// order, so a firm's pieces lie close together in the piece order and the
// piece launch's gathers of their sums share cache lines.  ref_s (the
// caller's order, 0-based) is relabeled in place.
static int renumber_s(fehip_solver *h, int32_t *ref_s, const int32_t *ref_t, int t) {
    const int64_t N = h->N, ns = h->n[h->s], nt = h->n[t];
    int32_t *sk = nullptr, *p2 = nullptr, *k1 = nullptr, *ss = nullptr, *perm = nullptr, *tt = nullptr;
    CHK(dalloc(h, &sk, N));
    CHK(dalloc(h, &p2, N));
    CHK(dalloc(h, &k1, N));
    CHK(dalloc(h, &ss, N));
    CHK(dalloc(h, &perm, N));
    CHK(dalloc(h, &tt, N));
    // observations sorted by (s, t)
    CHK(radix_sort(h, ref_t, N, nt, sk, p2));
    hipLaunchKernelGGL(k_gather_perm<int32_t>, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, h->stream, ref_s, p2, N,
                       k1);
    CHK(radix_sort(h, k1, N, ns, ss, perm, p2));
    hipLaunchKernelGGL(k_gather_perm<int32_t>, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, h->stream, ref_t, perm, N,
                       tt);
    HIPCHK(hipGetLastError());
    dfree(h, sk);
    dfree(h, p2);
    dfree(h, k1);
    dfree(h, perm);
    // runs of one (s, t) pair
    const int64_t nwords = N / 64 + 2;
    uint64_t *heads = nullptr;
    uint32_t *cnt = nullptr, *wp = nullptr, *sums = nullptr;
    CHK(dalloc(h, &heads, nwords));
    CHK(dalloc(h, &cnt, nwords + 1));
    CHK(dalloc(h, &wp, nwords + 1));
    CHK(dalloc(h, &sums, (nwords + kSortTile - 1) / kSortTile + 2));
    hipLaunchKernelGGL(k_piece_heads, dim3(grid_for(nwords, kWavesPerBlock)), dim3(kBlock), 0, h->stream,
                       (const int32_t *)ss, (const int32_t *)tt, N, N + 1, heads, nwords, cnt);
    HIPCHK(hipGetLastError());
    CHK(exclusive_scan_u32(h, cnt, nwords, wp, true, sums));
    uint32_t nruns = 0;
    HIPCHK(hipMemcpyAsync(&nruns, wp + nwords, sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    int32_t *run_s = nullptr, *run_t = nullptr, *run_start = nullptr, *roff = nullptr, *mainl = nullptr,
            *skeys = nullptr;
    CHK(dalloc(h, &run_s, std::max<int64_t>(nruns, 1)));
    CHK(dalloc(h, &run_t, std::max<int64_t>(nruns, 1)));
    CHK(dalloc(h, &run_start, (int64_t)nruns + 1));
    hipLaunchKernelGGL(k_run_starts, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, h->stream, (const int32_t *)ss,
                       (const int32_t *)tt, N, (const uint64_t *)heads, (const uint32_t *)wp, run_s, run_t,
                       run_start);
    hipLaunchKernelGGL(k_set_i32, dim3(1), dim3(64), 0, h->stream, run_start + nruns, (int32_t)N);
    HIPCHK(hipGetLastError());
    dfree(h, ss);
    dfree(h, tt);
    dfree(h, heads);
    dfree(h, cnt);
    dfree(h, wp);
    dfree(h, sums);
    // each s level's main t level, then the levels stably sorted by it
    CHK(dalloc(h, &roff, ns + 1));
    CHK(dalloc(h, &mainl, ns));
    CHK(dalloc(h, &skeys, ns));
    CHK(dalloc(h, &h->s_ext_of, ns));
    CHK(dalloc(h, &h->s_int_of, ns));
    CHK(dalloc(h, &h->s_tmp, ns));
    hipLaunchKernelGGL(k_offsets_from_sorted, dim3(grid_for((int64_t)nruns + 1, kBlock)), dim3(kBlock), 0, h->stream,
                       (const int32_t *)run_s, (int64_t)nruns, (int32_t)ns, roff);
    hipLaunchKernelGGL(k_main_level, dim3(grid_for(ns, kBlock)), dim3(kBlock), 0, h->stream, (const int32_t *)roff,
                       (const int32_t *)run_start, (const int32_t *)run_t, ns, (int32_t)nt, mainl);
    HIPCHK(hipGetLastError());
    CHK(radix_sort(h, mainl, ns, nt + 1, skeys, h->s_ext_of));
    hipLaunchKernelGGL(k_invert_perm, dim3(grid_for(ns, kBlock)), dim3(kBlock), 0, h->stream,
                       (const int32_t *)h->s_ext_of, ns, h->s_int_of);
    hipLaunchKernelGGL(k_relabel, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, h->stream, ref_s, N,
                       (const int32_t *)h->s_int_of);
    HIPCHK(hipGetLastError());
    h->h_s_int_of.resize(ns);
    HIPCHK(hipMemcpyAsync(h->h_s_int_of.data(), h->s_int_of, sizeof(int32_t) * ns, hipMemcpyDeviceToHost,
                          h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    dfree(h, run_s);
    dfree(h, run_t);
    dfree(h, run_start);
    dfree(h, roff);
    dfree(h, mainl);
    dfree(h, skeys);
    return 0;
}

// ---------------------------------------------------------------------------
// scale! / cache!  (src/SolverCPU.jl:78-95)
// ---------------------------------------------------------------------------

// One thread per level, serial over the level's observations in ascending
// caller order -- the reference's summation order (:81-83) -- in double,
// rounded into T after every add exactly as the Julia promotion does.
// Levels with more than kScaleBig observations are left to k_scale_big.
constexpr int32_t kScaleBig = 2048;

template <typename T>
__global__ void k_scale(const int32_t *__restrict__ off, const int32_t *__restrict__ perm, int64_t nlev,
                        const double *__restrict__ w, const double *__restrict__ z, T *__restrict__ scale_T,
                        double *__restrict__ scale_d) {
#pragma clang fp contract(off)
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nlev) return;
    constexpr int B = 16;  // loads in flight; the adds stay in observation order
    T s = (T)0;
    const int32_t q1 = off[k + 1];
    int32_t q = off[k];
    if (q1 - q > kScaleBig) return;
    for (; q + B <= q1; q += B) {
        double t[B];
#pragma unroll
        for (int j = 0; j < B; j++) {
            const int32_t i = perm ? perm[q + j] : q + j;
            const double zi = z ? z[i] : 1.0;
            const double wi = w ? w[i] : 1.0;
            t[j] = (zi * zi) * wi;
        }
#pragma unroll
        for (int j = 0; j < B; j++) s = (T)((double)s + t[j]);
    }
    for (; q < q1; q++) {
        const int32_t i = perm ? perm[q] : q;
        const double zi = z ? z[i] : 1.0;
        const double wi = w ? w[i] : 1.0;
        s = (T)((double)s + (zi * zi) * wi);
    }
    const T r = s > (T)0 ? (T)1 / (T)sqrt((double)s) : (T)0;
    scale_T[k] = r;
    scale_d[k] = (double)r;
}

// scale! of one large level per 64-lane workgroup, in the same order and
// rounding as k_scale: the wave loads 1024 terms (z_i^2 w_i) at a time into
// LDS and lane 0 adds them in observation order; the next 1024 terms are in
// flight while lane 0 adds.
template <typename T>
__global__ __launch_bounds__(kWave) void k_scale_big(const int32_t *__restrict__ off, const int32_t *__restrict__ perm,
                                                     const int32_t *__restrict__ levels, const double *__restrict__ w,
                                                     const double *__restrict__ z, T *__restrict__ scale_T,
                                                     double *__restrict__ scale_d) {
#pragma clang fp contract(off)
    constexpr int B = 16, CH = kWave * B;
    __shared__ double buf[CH];
    const int lane = threadIdx.x;
    const int32_t k = levels[blockIdx.x];
    const int32_t q0 = off[k], q1 = off[k + 1];
    double t[B];
    auto load = [&](int32_t base) {
#pragma unroll
        for (int j = 0; j < B; j++) {
            const int32_t q = base + j * kWave + lane;
            t[j] = 0.0;
            if (q < q1) {
                const int32_t i = perm ? perm[q] : q;
                const double zi = z ? z[i] : 1.0;
                const double wi = w ? w[i] : 1.0;
                t[j] = (zi * zi) * wi;
            }
        }
    };
    T s = (T)0;
    load(q0);
    for (int32_t base = q0; base < q1; base += CH) {
#pragma unroll
        for (int j = 0; j < B; j++) buf[j * kWave + lane] = t[j];
        __syncthreads();
        if (base + CH < q1) load(base + CH);
        if (lane == 0) {
            const int32_t n = q1 - base < CH ? q1 - base : CH;
            for (int32_t e = 0; e < n; e++) s = (T)((double)s + buf[e]);
        }
        __syncthreads();
    }
    if (lane == 0) {
        const T r = s > (T)0 ? (T)1 / (T)sqrt((double)s) : (T)0;
        scale_T[k] = r;
        scale_d[k] = (double)r;
    }
}

// cache[j] = z_i * sqrt(w_i) * scale[ref]  with i = perm_s[j]  (:93)
template <typename T>
__global__ void k_cache(const int32_t *__restrict__ perm_s, int64_t N, const int32_t *__restrict__ ref_int,
                        const double *__restrict__ w, const double *__restrict__ z, const T *__restrict__ scale,
                        T *__restrict__ cache) {
#pragma clang fp contract(off)
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    const int32_t i = perm_s[j];
    const double zi = z ? z[i] : 1.0;
    const double wi = w ? w[i] : 1.0;
    cache[j] = (T)((zi * sqrt(wi)) * (double)scale[ref_int[j]]);
}

// packed[i] = lo[i] | hi[i] << 26 (lo < 2^26, hi < 64)
__global__ void k_pack_refs(const int32_t *__restrict__ lo, const int32_t *__restrict__ hi, int64_t n,
                            int32_t *__restrict__ packed) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) packed[i] = (int32_t)((uint32_t)lo[i] | ((uint32_t)hi[i] << 26));
}

__global__ void k_iota(int32_t *__restrict__ a, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = (int32_t)i;
}

template <typename T>
__global__ void k_copy_range(const T *__restrict__ a, T *__restrict__ b, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) b[e] = a[e];
}

__global__ void k_round_f32(const double *__restrict__ a, float *__restrict__ b, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) b[e] = (float)a[e];
}

__global__ void k_sqrtw(const double *__restrict__ w, const int32_t *__restrict__ perm_s, int64_t N,
                        double *__restrict__ out) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < N) out[j] = sqrt(w[perm_s[j]]);
}

template <typename T>
constexpr int pidx() { return sizeof(T) == 8 ? 0 : 1; }

// sqrtw_T[q] follows sqrtw_int: an alias for double, a rounded copy for float
static int set_sqrtw_T(fehip_solver *h, int q) {
    if (q == 0) {
        h->sqrtw_T[0] = h->sqrtw_int;
        return 0;
    }
    if (!h->sqrtw_int) {
        dfree(h, h->sqrtw_T[1]);
        h->sqrtw_T[1] = nullptr;
        return 0;
    }
    if (!h->sqrtw_T[1]) CHK(dalloc_bytes(h, &h->sqrtw_T[1], sizeof(float) * h->N));
    hipLaunchKernelGGL(k_round_f32, dim3(grid_for(h->N, kBlock)), dim3(kBlock), 0, h->stream,
                       (const double *)h->sqrtw_int, (float *)h->sqrtw_T[1], h->N);
    HIPCHK(hipGetLastError());
    return 0;
}

template <typename T>
static int compute_scale_cache(fehip_solver *h) {
    constexpr int q = pidx<T>();
    // scale! sums each level in the caller's observation order (the
    // reference's order), through the caller-order bucketing of every FE
    for (int g = 0; g < h->G; g++) {
        FeInfo &f = h->fe[g];
        hipLaunchKernelGGL(k_scale<T>, dim3(grid_for(f.n, 64)), dim3(64), 0, h->stream, f.ooff, f.operm, f.n,
                           h->w_orig, f.z_orig, (T *)f.scale_T[q], f.scale_d);
        HIPCHK(hipGetLastError());
        if (f.big_levels.empty()) continue;
        if (!f.d_big) {
            CHK(dalloc(h, &f.d_big, f.big_levels.size()));
            HIPCHK(hipMemcpyAsync(f.d_big, f.big_levels.data(), sizeof(int32_t) * f.big_levels.size(),
                                  hipMemcpyHostToDevice, h->stream));
        }
        hipLaunchKernelGGL(k_scale_big<T>, dim3((unsigned)f.big_levels.size()), dim3(kWave), 0, h->stream, f.ooff,
                           f.operm, f.d_big, h->w_orig, f.z_orig, (T *)f.scale_T[q], f.scale_d);
        HIPCHK(hipGetLastError());
    }
    for (int g = 0; g < h->G; g++) {
        FeInfo &f = h->fe[g];
        hipLaunchKernelGGL(k_cache<T>, dim3(grid_for(h->N, kBlock)), dim3(kBlock), 0, h->stream, h->perm_s, h->N,
                           f.ref_int, h->w_orig, f.z_orig, (const T *)f.scale_T[q], (T *)f.cache_int[q]);
        if (f.kind == KIND_CSR)
            hipLaunchKernelGGL(k_gather_perm<T>, dim3(grid_for(h->N, kBlock)), dim3(kBlock), 0, h->stream,
                               (const T *)f.cache_int[q], f.iperm, h->N, (T *)f.cache_csr[q]);
        HIPCHK(hipGetLastError());
    }
    if (h->w_orig) {
        if (!h->sqrtw_int) CHK(dalloc(h, &h->sqrtw_int, h->N));
        hipLaunchKernelGGL(k_sqrtw, dim3(grid_for(h->N, kBlock)), dim3(kBlock), 0, h->stream, h->w_orig, h->perm_s,
                           h->N, h->sqrtw_int);
        HIPCHK(hipGetLastError());
    }
    // every FE's scale in one p-vector, and sqrt(w) in T (the fused pass's
    // factored form)
    if (!h->scale_all[q]) CHK(dalloc_bytes(h, &h->scale_all[q], sizeof(T) * h->p));
    for (int g = 0; g < h->G; g++)
        hipLaunchKernelGGL(k_copy_range<T>, dim3(grid_for(h->n[g], kBlock)), dim3(kBlock), 0, h->stream,
                           (const T *)h->fe[g].scale_T[q], (T *)h->scale_all[q] + h->off[g], h->n[g]);
    CHK(set_sqrtw_T(h, q));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

static int do_update_weights(fehip_solver *h, const double *w) {
    if (w) {
        for (int64_t i = 0; i < h->N; i++)
            if (!(w[i] > 0.0) || !std::isfinite(w[i]))
                return fail(FEHIP_EARG, "weights must be positive and finite (observation " + std::to_string(i + 1) +
                                            ")");
        if (!h->w_orig) CHK(dalloc(h, &h->w_orig, h->N));
        HIPCHK(hipMemcpyAsync(h->w_orig, w, sizeof(double) * h->N, hipMemcpyHostToDevice, h->stream));
    } else {
        dfree(h, h->w_orig);
        dfree(h, h->sqrtw_int);
        h->w_orig = nullptr;
        h->sqrtw_int = nullptr;
        h->sqrtw_T[0] = nullptr;
    }
    if (h->precision == FEHIP_FP32) return compute_scale_cache<float>(h);
    CHK(compute_scale_cache<double>(h));
    if (h->precision == FEHIP_MIXED) {
        // the FP32 streams hold the FP64 scales and caches rounded once
        for (int g = 0; g < h->G; g++) {
            FeInfo &f = h->fe[g];
            hipLaunchKernelGGL(k_round_f32, dim3(grid_for(f.n, kBlock)), dim3(kBlock), 0, h->stream,
                               (const double *)f.scale_T[0], (float *)f.scale_T[1], f.n);
            hipLaunchKernelGGL(k_round_f32, dim3(grid_for(h->N, kBlock)), dim3(kBlock), 0, h->stream,
                               (const double *)f.cache_int[0], (float *)f.cache_int[1], h->N);
            if (f.kind == KIND_CSR)
                hipLaunchKernelGGL(k_round_f32, dim3(grid_for(h->N, kBlock)), dim3(kBlock), 0, h->stream,
                                   (const double *)f.cache_csr[0], (float *)f.cache_csr[1], h->N);
            HIPCHK(hipGetLastError());
        }
        if (!h->scale_all[1]) CHK(dalloc_bytes(h, &h->scale_all[1], sizeof(float) * h->p));
        hipLaunchKernelGGL(k_round_f32, dim3(grid_for(h->p, kBlock)), dim3(kBlock), 0, h->stream,
                           (const double *)h->scale_all[0], (float *)h->scale_all[1], h->p);
        CHK(set_sqrtw_T(h, 1));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    return 0;
}

// ---------------------------------------------------------------------------
// create
// ---------------------------------------------------------------------------
static int dense_limit(int precision) {
    // levels whose per-lane LDS accumulators fit 32 KB for one wave
    return precision == FEHIP_FP64 ? 64 : 64;
}

template <typename R>
static int create_impl(fehip_t **out, int64_t N, int32_t G, const R *const *refs, const int64_t *nlevels,
                       const double *const *interaction, const double *weights, int32_t precision) {
    if (!out) return fail(FEHIP_EARG, "out is NULL");
    *out = nullptr;
    if (N < 1) return fail(FEHIP_EDIM, "N must be >= 1");
    if (N > (int64_t)INT32_MAX - 4096) return fail(FEHIP_EDIM, "N exceeds the int32 observation index range");
    if (G < 1 || G > kMaxG) return fail(FEHIP_EARG, "G must be in 1.." + std::to_string(kMaxG));
    if (!refs || !nlevels) return fail(FEHIP_EARG, "refs and nlevels are required");
    if (precision != FEHIP_FP64 && precision != FEHIP_FP32 && precision != FEHIP_MIXED)
        return fail(FEHIP_EARG, "unknown precision");
    for (int g = 0; g < G; g++) {
        if (!refs[g]) return fail(FEHIP_EARG, "refs[" + std::to_string(g) + "] is NULL");
        if (nlevels[g] < 1 || nlevels[g] > (int64_t)INT32_MAX - 2)
            return fail(FEHIP_EDIM, "nlevels[" + std::to_string(g) + "] out of range");
    }
    std::unique_ptr<fehip_solver> hp(new fehip_solver());
    fehip_solver *h = hp.get();
    if (const char *e = std::getenv("FEHIP_POLL_ITERS")) h->poll_iters = std::max(1, std::atoi(e));
    h->N = N;
    h->G = G;
    h->precision = precision;
    h->has64 = precision != FEHIP_FP32;
    h->has32 = precision != FEHIP_FP64;
    h->factored = env_int("FEHIP_FACTORED", 1, 0, 1) != 0;
    h->seg_fix = env_int("FEHIP_SEG_FIX", 1, 0, 1) != 0;
    h->wide = env_int("FEHIP_WIDE", 1, 0, 1) != 0;
    h->pieces_proj = env_int("FEHIP_PIECES_PROJ", 1, 0, 1) != 0;
    for (int g = 0; g < G; g++)
        if (interaction && interaction[g]) h->factored = false;
    if (const char *e = std::getenv("FEHIP_MIXED_TOL")) h->mixed_tol = std::atof(e);
    h->n.assign(nlevels, nlevels + G);
    h->off.resize(G + 1);
    h->off[0] = 0;
    for (int g = 0; g < G; g++) h->off[g + 1] = h->off[g] + h->n[g];
    h->p = h->off[G];
    h->fe.resize(G);
    h->s = 0;
    for (int g = 1; g < G; g++)
        if (h->n[g] > h->n[h->s]) h->s = g;
    struct Cleanup {
        fehip_solver *h;
        bool armed = true;
        ~Cleanup() {
            if (!armed) return;
            for (void *q : h->allocs) (void)hipFree(q);
            h->allocs.clear();
            if (h->h_flag) (void)hipHostFree(h->h_flag);
            for (int e = 0; e < 2; e++)
                if (h->poll_ev[e]) (void)hipEventDestroy(h->poll_ev[e]);
            if (h->fork_ev) (void)hipEventDestroy(h->fork_ev);
            if (h->join_ev) (void)hipEventDestroy(h->join_ev);
            if (h->stream2) (void)hipStreamDestroy(h->stream2);
            if (h->stream) (void)hipStreamDestroy(h->stream);
        }
    } cleanup{h};
    HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    HIPCHK(hipHostMalloc((void **)&h->h_flag, sizeof(int32_t) * 4, hipHostMallocDefault));
    for (int e = 0; e < 2; e++) HIPCHK(hipEventCreateWithFlags(&h->poll_ev[e], hipEventDisableTiming));
    CHK(dalloc(h, &h->d_flag, 4));
    HIPCHK(hipMemsetAsync(h->d_flag, 0, sizeof(int32_t) * 4, h->stream));

    // refs -> int32 0-based, range-checked (caller's order)
    std::vector<int32_t *> ref_orig(G, nullptr);
    {
        R *tmp = nullptr;
        CHK(dalloc(h, &tmp, N));
        for (int g = 0; g < G; g++) {
            CHK(dalloc(h, &ref_orig[g], N));
            HIPCHK(hipMemcpyAsync(tmp, refs[g], sizeof(R) * N, hipMemcpyHostToDevice, h->stream));
            hipLaunchKernelGGL(k_refs_to_i32<R>, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, h->stream, tmp, N,
                               h->n[g], ref_orig[g], h->d_flag);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(h->h_flag, h->d_flag, sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
            if (h->h_flag[0])
                return fail(FEHIP_EREF, "refs[" + std::to_string(g) + "] has a value outside 1.." +
                                            std::to_string(h->n[g]));
        }
        dfree(h, tmp);
    }
    // Internal order: observations sorted stably by (level of FE s, level of
    // the largest other FE with a CSR, caller's index).  FE s's levels stay
    // contiguous runs; within a run, each (FE s, FE t) cell is contiguous, so
    // FE t's D'u reads come in runs.
    int t_sec = -1;
    for (int g = 0; g < G; g++)
        if (g != h->s && h->n[g] > dense_limit(precision) && (t_sec < 0 || h->n[g] > h->n[t_sec])) t_sec = g;
    h->t_sec = t_sec;
    // FE s in its internal numbering from here on (FEHIP_RENUMBER=0 keeps the caller's)
    if (t_sec >= 0 && env_int("FEHIP_RENUMBER", 1, 0, 1)) CHK(renumber_s(h, ref_orig[h->s], ref_orig[t_sec], t_sec));
    {
        FeInfo &f = h->fe[h->s];
        f.n = h->n[h->s];
        f.kind = KIND_SORTED;
        CHK(dalloc(h, &f.ref_int, N));
        CHK(dalloc(h, &h->perm_s, N));
        CHK(dalloc(h, &h->ident, N));
        hipLaunchKernelGGL(k_iota, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, h->stream, h->ident, N);
        if (t_sec >= 0) {
            int32_t *sk = nullptr, *p2 = nullptr, *k1 = nullptr;
            CHK(dalloc(h, &sk, N));
            CHK(dalloc(h, &p2, N));
            CHK(dalloc(h, &k1, N));
            CHK(radix_sort(h, ref_orig[t_sec], N, h->n[t_sec], sk, p2));
            hipLaunchKernelGGL(k_gather_perm<int32_t>, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, h->stream,
                               ref_orig[h->s], p2, N, k1);
            HIPCHK(hipGetLastError());
            CHK(radix_sort(h, k1, N, f.n, f.ref_int, h->perm_s, p2));
            HIPCHK(hipStreamSynchronize(h->stream));
            dfree(h, sk);
            dfree(h, p2);
            dfree(h, k1);
        } else {
            CHK(radix_sort(h, ref_orig[h->s], N, f.n, f.ref_int, h->perm_s));
        }
        CHK(dalloc(h, &h->pos_s, N));
        hipLaunchKernelGGL(k_invert_perm, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, h->stream,
                           (const int32_t *)h->perm_s, N, h->pos_s);
        HIPCHK(hipGetLastError());
        CHK(csr_products(h, f.ref_int, N, f.n, nullptr, &f));
        // the reference's bucketing over the caller's order (scale!, get_csr)
        CHK(dalloc(h, &f.operm, N));
        CHK(dalloc(h, &f.ooff, f.n + 1));
        int32_t *skeys = nullptr;
        CHK(dalloc(h, &skeys, N));
        CHK(radix_sort(h, ref_orig[h->s], N, f.n, skeys, f.operm));
        CHK(csr_products(h, skeys, N, f.n, f.ooff, nullptr));
        HIPCHK(hipStreamSynchronize(h->stream));
        dfree(h, skeys);
        f.ioff = f.ooff;  // same level counts, so the same offsets
    }
    for (int g = 0; g < G; g++) {
        if (g == h->s) continue;
        FeInfo &f = h->fe[g];
        f.n = h->n[g];
        f.kind = f.n <= dense_limit(precision) ? KIND_DENSE : KIND_CSR;
        CHK(dalloc(h, &f.ref_int, N));
        hipLaunchKernelGGL(k_gather_perm<int32_t>, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, h->stream,
                           ref_orig[g], h->perm_s, N, f.ref_int);
        HIPCHK(hipGetLastError());
        // bucketing over the caller's order (bit-exact with bucketize_refs; used by scale!)
        CHK(dalloc(h, &f.operm, N));
        CHK(dalloc(h, &f.ooff, f.n + 1));
        int32_t *skeys = nullptr;
        CHK(dalloc(h, &skeys, N));
        CHK(radix_sort(h, ref_orig[g], N, f.n, skeys, f.operm));
        CHK(csr_products(h, skeys, N, f.n, f.ooff, nullptr));
        if (f.kind == KIND_CSR) {
            // bucketing over the internal order (locality of the D'u reads)
            CHK(dalloc(h, &f.iperm, N));
            CHK(dalloc(h, &f.ioff, f.n + 1));
            CHK(radix_sort(h, f.ref_int, N, f.n, skeys, f.iperm));
            CHK(csr_products(h, skeys, N, f.n, f.ioff, &f));
        }
        HIPCHK(hipStreamSynchronize(h->stream));
        dfree(h, skeys);
    }
    for (int g = 0; g < G; g++) dfree(h, ref_orig[g]);
    free_sort_scratch(h);
    // interactions (caller's order), scales and caches
    for (int g = 0; g < G; g++) {
        FeInfo &f = h->fe[g];
        if (interaction && interaction[g]) {
            CHK(dalloc(h, &f.z_orig, N));
            HIPCHK(hipMemcpyAsync(f.z_orig, interaction[g], sizeof(double) * N, hipMemcpyHostToDevice, h->stream));
        }
        CHK(dalloc(h, &f.scale_d, f.n));
        for (int q = 0; q < 2; q++) {
            if (!(q == 0 ? h->has64 : h->has32)) continue;
            const size_t ts = q == 0 ? 8 : 4;
            CHK(dalloc_bytes(h, &f.scale_T[q], ts * f.n));
            CHK(dalloc_bytes(h, &f.cache_int[q], ts * N));
            if (f.kind == KIND_CSR) CHK(dalloc_bytes(h, &f.cache_csr[q], ts * N));
            else f.cache_csr[q] = f.cache_int[q];
        }
        if (f.kind != KIND_DENSE) {
            f.h_ioff.resize(f.n + 1);
            HIPCHK(hipMemcpyAsync(f.h_ioff.data(), f.ioff, sizeof(int32_t) * (f.n + 1), hipMemcpyDeviceToHost,
                                  h->stream));
        }
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    for (int g = 0; g < G; g++) {
        FeInfo &f = h->fe[g];
        std::vector<int32_t> o(f.n + 1);
        HIPCHK(hipMemcpy(o.data(), f.ooff, sizeof(int32_t) * (f.n + 1), hipMemcpyDeviceToHost));
        for (int64_t k = 0; k < f.n; k++)
            if (o[k + 1] - o[k] > kScaleBig) f.big_levels.push_back((int32_t)k);
    }
    if (G == 3) {
        int d = -1, g2 = -1;
        for (int g = 0; g < G; g++)
            if (h->fe[g].kind == KIND_DENSE && d < 0) d = g;
        for (int g = 0; g < G; g++)
            if (g != h->s && g != d) g2 = g;
        // (g2 not few-level itself: the fused pass reads a second few-level
        // FE's refs directly)
        if (d >= 0 && g2 >= 0 && h->fe[g2].kind != KIND_DENSE && h->n[d] <= 64 && h->n[g2] <= (int64_t(1) << 26)) {
            CHK(dalloc(h, &h->packed, N));
            hipLaunchKernelGGL(k_pack_refs, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, h->stream,
                               (const int32_t *)h->fe[g2].ref_int, (const int32_t *)h->fe[d].ref_int, N, h->packed);
            HIPCHK(hipGetLastError());
            h->pack_g = g2;
            h->pack_d = d;
        }
    }
    CHK(do_update_weights(h, weights));
    cleanup.armed = false;
    *out = hp.release();
    return 0;
}

// ---------------------------------------------------------------------------
// per-K plan: tile schedule, fix-up items, buffers
// ---------------------------------------------------------------------------
// Tiles: one wave per tile, about one tile per resident wave, so every wave
// of a launch does the same amount of work.  FE s (and the fused pass over
// it) uses ~4096 tiles, the CSR FEs ~8192 (their gathers want more waves in
// flight).  Both are multiples of 256 positions.  FEHIP_TILES_S and
// FEHIP_TILES_C override the counts.
// few-level FEs accumulated by the fused pass (when they fit its LDS)
static bool fuse_dense_on(const fehip_solver *h) {
    if (h->sched_fuse_dense >= 0) return h->sched_fuse_dense != 0;
    return env_int("FEHIP_FUSE_DENSE", 1, 0, 1) != 0;
}

// columns per lane of the fused pass (the wide lane map, fused_wide_c in
// linmap.h) for storage of ts bytes and K columns on this handle
static int fused_cols(const fehip_solver *h, size_t ts, int K) {
    if (!h->wide || h->G < 2 || h->G > 3 || K < 16) return 1;
    return ts != 8 ? FUSED_WIDE_C32 : h->factored ? FUSED_WIDE_C : FUSED_WIDE_C_CACHE;
}

static size_t align16(size_t b) { return (b + 15) & ~(size_t)15; }
// levels of the few-level FE whose gathered vector the fused pass's packed
// forms stage in LDS after acc bytes of accumulators (k_fused's nvl, PK):
// the packed pair present and everything within 64 KB, else 0 (those forms
// then read the packed stream twice and gather the vector from memory)
static int32_t staged_levels(const fehip_solver *h, size_t ts, int K, size_t acc) {
    if (h->G != 3 || !h->packed || h->pack_d < 0) return 0;
    const size_t b = (size_t)h->fe[h->pack_d].n * K * ts;
    return align16(acc) + b <= 65536 ? (int32_t)h->fe[h->pack_d].n : 0;
}

static int64_t tile_size(int64_t N, int64_t ntiles) {
    const int64_t t = (N + ntiles - 1) / ntiles;
    return std::max<int64_t>(256, (t + 255) / 256 * 256);
}

// Tile starts over positions [0, n) of a segmented FE with segment offsets
// off (nlev + 1 entries), ending with n: each tile is cut at the last segment
// head at or before its start + tp, unless that segment began at or before
// the start (a segment longer than a tile): then at start + tp.
static void cut_at_heads(const int32_t *off, int64_t nlev, int64_t n, int64_t tp, std::vector<int64_t> &tsh) {
    for (int64_t q = 0; q < n;) {
        tsh.push_back(q);
        const int64_t tgt = q + tp;
        if (tgt >= n) break;
        const int64_t head = *(std::upper_bound(off, off + nlev + 1, tgt) - 1);
        q = head > q ? head : tgt;
    }
    tsh.push_back(n);
}

// The piece form of the plain iteration (factored handles whose only CSR FE
// is the secondary sort key t): the fused pass sums sqrt(w) .* u over pieces
// -- runs of one (FE s level, FE t level) pair inside one fused tile -- and
// one launch then forms both FEs' level sums from the piece sums, FE s over
// its contiguous pieces and FE t through the pieces' CSR by FE t level.
// This replaces the CSR pass's gather of u over all N observations by a
// gather of the pieces' sums.
struct PieceForm {
    int64_t np = 0;                  // pieces (0: the form is off for this plan)
    uint64_t *heads = nullptr;       // piece heads over the internal order (+ sentinel bits)
    uint32_t *wprefix = nullptr;     // heads before each word
    int32_t *piece_s = nullptr;      // piece -> FE s level
    int32_t *pt_perm = nullptr;      // pieces ordered stably by FE t level
    int32_t *pt_lev = nullptr;       // the FE t level of each piece in that order
    FeInfo fs, ft;                   // head bits, word prefixes and level maps of the pieces by FE s / FE t level
    int64_t *tstart = nullptr;       // tiles over pieces: FE s's, then FE t's, each ending with np
    int64_t ts0_t = 0;               // FE t's first entry in tstart
    int32_t tile0_t = 0, total_tiles = 0;
    int32_t nsegs_s = 0, nsegs_t = 0;  // observed levels
    unsigned blocks = 0;             // workgroups of the piece launch's tiles
    unsigned fix1 = 1, fix2 = 1;     // workgroups of its fix-up parts (sized by their items)
    double *tileA = nullptr, *tileB = nullptr;
    // norm partials of one iteration's D'u, one span: part 1 [0, fix1), the
    // tiles [fix1, fix1 + blocks), part 2 after them (when it runs)
    double *part = nullptr;
    // part 1 (the piece launch's first blocks): few-level FEs' levels, levels
    // without observations; part 2 (after it): levels whose pieces cross tiles
    FixItem *items1 = nullptr, *items2 = nullptr;
    int32_t n1 = 0, n2 = 0;
    // the projected iteration (FE s eliminated) runs FE t's tiles only: its
    // parts 1 and 2 without FE s's items
    FixItem *items1p = nullptr, *items2p = nullptr;
    int32_t n1p = 0, n2p = 0;
    unsigned fix1p = 1, fix2p = 1;
    void *psum = nullptr;            // np x K, T
};

struct Plan {
    int K = 0;
    int64_t waves_f = 0, waves_c = 0;  // resident waves of the fused pass / the CSR launch
    int32_t total_tiles = 0, nseg = 0, ndense = 0, dense_blocks = 0, dense_chunk = 0, nitems = 0;
    int32_t dense_wpb = 1, dense_lds_wave = 0;
    std::vector<int> seg_fe, dense_fe;  // seg_fe[0] is FE s
    std::vector<int64_t> seg_ts0;       // per segmented FE: its first entry in tstart
    int64_t *tstart = nullptr;          // tile start positions (+ N), FE after FE
    std::vector<int32_t> seg_nsegs;     // non-empty levels, per segmented FE
    std::vector<int32_t> tile0, blk0;
    std::vector<int64_t> drow0;
    FixItem *items = nullptr;
    double *tileA = nullptr, *tileB = nullptr, *dpart = nullptr;
    unsigned db_blocks = 0, seg_blocks = 0, csr_blocks = 0, fix_blocks = 0, upd_blocks = 0;
    double *part_db = nullptr, *part_seg = nullptr, *part_fix = nullptr, *part_upd = nullptr;
    // fused pass (k_fused): FE s tiles of R positions, one per wave
    int64_t R = 0;
    int32_t ntiles_s = 0, nslot = 0;
    bool fused_dense = false;
    std::vector<int32_t> slot0;
    unsigned fused_blocks = 0;
    size_t fused_lds = 0;
    int32_t nvl = 0;  // levels of the vector the packed fused forms stage in LDS
    double *part_fused = nullptr, *wsum = nullptr;
    void *vt = nullptr;  // p x K, scale .* v (factored fused pass)
    uint8_t *sdirect = nullptr;
    void *state = nullptr;
    ColOps *ops_db = nullptr, *ops_dtx = nullptr, *ops_once = nullptr;
    int32_t *all_done = nullptr;
    void *u = nullptr, *r = nullptr, *x = nullptr, *v = nullptr, *hh = nullptr, *hbar = nullptr;
    double *dY = nullptr;  // staging for host columns (N x K doubles)
    hipGraphExec_t graph = nullptr;
    int graph_iters = 0;
    bool graph_fwl = false;
    // projected iteration: FE s items (crossing segments, into zs from level
    // 0), the other FEs' items, zs = the FE s correction (n_s x K)
    FixItem *items_s = nullptr, *items_o = nullptr;
    int32_t nitems_s = 0, nitems_o = 0;
    // plain iteration, split fix-up: FE s (and the fused few-level FEs')
    // items beside the CSR pass, the CSR FEs' (and a separate few-level
    // pass's) items after it
    FixItem *items_sd = nullptr, *items_c = nullptr, *items_d = nullptr;
    int32_t nitems_sd = 0, nitems_c = 0, nitems_d = 0;  // items_d: the fused few-level FEs' (projected part 1)
    void *zs = nullptr;
    ColOps *ops_one = nullptr, *ops_neg = nullptr, *ops_sub = nullptr;
    PieceForm pc;
};

static int build_pieces(fehip_solver *h, Plan *P, size_t ts, const std::vector<FixItem> &dense_items,
                        int64_t tiles);

// waves_f / waves_c: how many waves of the fused pass / of the CSR D'u launch
// the device holds at once; each gets one tile
static int get_plan(fehip_solver *h, int K, size_t ts, Plan **out, int64_t waves_f, int64_t waves_c) {
    const int key = K * 16 + (int)ts;
    auto it = h->plans.find(key);
    if (it != h->plans.end()) {
        *out = it->second;
        return 0;
    }
    std::unique_ptr<Plan> pl(new Plan());
    Plan *P = pl.get();
    P->K = K;
    const int64_t N = h->N;
    P->waves_f = waves_f;
    P->waves_c = waves_c;
    // the fused pass addresses a tile of u and the coefficient vectors with
    // 32-bit byte offsets: R*K*8 and p*K*8 stay below 2^32
    if ((double)h->p * K * sizeof(double) >= 4294967296.0)
        return fail(FEHIP_EDIM, "p * K * 8 bytes exceeds 2^32: use fewer columns per batch");
    P->R = tile_size(N, h->sched_tiles_s > 0 ? h->sched_tiles_s : env_int("FEHIP_TILES_S", waves_f, 1, 1 << 20));
    const int64_t rmax = ((int64_t)1 << 31) / ((int64_t)K * (int64_t)sizeof(double)) / 256 * 256;
    P->R = std::min<int64_t>(P->R, std::max<int64_t>(rmax, 256));
    const int64_t Rc =
        tile_size(N, h->sched_tiles_c > 0 ? h->sched_tiles_c : env_int("FEHIP_TILES_C", waves_c, 1, 1 << 20));
    P->ntiles_s = (int32_t)((N + P->R - 1) / P->R);
    P->fused_blocks = grid_for(P->ntiles_s, kWavesPerBlock);
    std::vector<FixItem> items;
    std::vector<uint8_t> sdirect(h->fe[h->s].n, 0);
    std::vector<int64_t> tsh;  // tile starts (+ N) of every segmented FE, in order
    int32_t t0 = 0;
    std::vector<int> order{h->s};
    for (int g = 0; g < h->G; g++)
        if (g != h->s && h->fe[g].kind == KIND_CSR) order.push_back(g);
    for (int g : order) {
        const FeInfo &f = h->fe[g];
        const int64_t tp = g == h->s ? P->R : Rc;
        P->seg_fe.push_back(g);
        P->seg_ts0.push_back((int64_t)tsh.size());
        P->tile0.push_back(t0);
        // FE s: fixed tiles of R positions (the fused pass's partition).  A
        // CSR FE: cut at the last segment head at or before each target
        // position start + Rc, unless that segment began at or before start
        // (a segment longer than a tile): then it is cut at the target.
        const size_t b0 = tsh.size();
        if (g == h->s) {
            for (int64_t q = 0; q < N; q += tp) tsh.push_back(q);
            tsh.push_back(N);
        } else {
            cut_at_heads(f.h_ioff.data(), f.n, N, tp, tsh);
        }
        const int64_t *tb = tsh.data() + b0;
        const int64_t ntl = (int64_t)(tsh.size() - b0) - 1;
        auto tile_of = [&](int64_t pos) -> int64_t { return (int64_t)(std::upper_bound(tb, tb + ntl + 1, pos) - tb) - 1; };
        int32_t nonempty = 0;
        for (int64_t k = 0; k < f.n; k++) nonempty += f.h_ioff[k + 1] > f.h_ioff[k] ? 1 : 0;
        P->seg_nsegs.push_back(nonempty);
        for (int64_t k = 0; k < f.n; k++) {
            const int64_t st = f.h_ioff[k], en = f.h_ioff[k + 1];
            if (en <= st) {  // level without observations: coef <- b*coef only
                FixItem fi{};
                fi.kind = 2;
                fi.fe = g;
                fi.level = (int32_t)k;
                fi.coef_off = h->off[g];
                items.push_back(fi);
                continue;
            }
            const int64_t tfirst = tile_of(st), tlast = tile_of(en - 1);
            if (tfirst == tlast) {
                if (g == h->s) sdirect[k] = 1;
                continue;
            }
            FixItem fi{};
            fi.kind = 0;
            fi.fe = g;
            fi.level = (int32_t)k;
            fi.coef_off = h->off[g];
            fi.first = t0 + tfirst;
            fi.count = (int32_t)(tlast - tfirst);
            items.push_back(fi);
        }
        t0 += (int32_t)ntl;
    }
    CHK(dalloc(h, &P->tstart, tsh.size()));
    HIPCHK(hipMemcpyAsync(P->tstart, tsh.data(), sizeof(int64_t) * tsh.size(), hipMemcpyHostToDevice, h->stream));
    P->nseg = (int32_t)P->seg_fe.size();
    P->total_tiles = t0;
    // dense FEs.  If their levels fit the fused pass's LDS (<= 64 KB per
    // workgroup of 4 waves), the fused pass accumulates them and k_dtx_dense
    // (used outside the iteration) runs on the same partition: chunk = R,
    // 4 waves per workgroup.  Otherwise k_dtx_dense does them in every
    // iteration with ~4096 chunks per FE and waves per workgroup bounded by
    // 64 KB of LDS.
    const int Pq = kWave / K;
    // accumulator bytes: FP64 double; FP32 modes a float per lane and level
    // (DenseAcc in linmap.h: k_dtx_dense phases FP64 only)
    const size_t accsz = ts == 8 ? sizeof(double) : sizeof(float);
    int maxlev = 1;
    for (int g = 0; g < h->G; g++)
        if (h->fe[g].kind == KIND_DENSE) {
            maxlev = std::max<int>(maxlev, (int)h->fe[g].n);
            P->slot0.push_back(P->nslot);
            P->nslot += (int32_t)h->fe[g].n;
        }
    // accumulators per slot and wave: one per lane, or one per column
    // (dense_phased: FP64, K >= 16), in k_fused and k_dtx_dense alike; the
    // fused pass in the wide lane map keeps one FP64 accumulator per column
    // in every precision
    const size_t per_slot = ts == 8 && K >= 16 ? (size_t)K * accsz : (size_t)kWave * accsz;
    const size_t per_slot_fused = fused_cols(h, ts, K) > 1 ? (size_t)K * sizeof(double) : per_slot;
    P->dense_lds_wave = (int32_t)(per_slot * maxlev);
    // (k_dtx_dense then runs on the fused partition, 4 waves per workgroup:
    // its accumulators must fit 64 KB as well)
    P->fused_dense = P->nslot > 0 && (size_t)P->nslot * per_slot * kWavesPerBlock <= 65536 &&
                     (size_t)P->dense_lds_wave * kWavesPerBlock <= 65536 && fuse_dense_on(h);
    int64_t chunk;
    if (P->fused_dense) {
        chunk = P->R;
        P->dense_wpb = kWavesPerBlock;
        const size_t acc = (size_t)P->nslot * per_slot_fused * kWavesPerBlock;
        P->nvl = staged_levels(h, ts, K, acc);
        P->fused_lds = P->nvl ? align16(acc) + (size_t)P->nvl * K * ts : acc;
    } else {
        chunk = (N + 4095) / 4096;
        chunk = std::max<int64_t>(chunk, 16 * Pq);
        chunk = (chunk + Pq - 1) / Pq * Pq;
        P->dense_wpb = (int32_t)std::max<int64_t>(1, std::min<int64_t>(kWavesPerBlock, 65536 / P->dense_lds_wave));
    }
    P->dense_chunk = (int32_t)chunk;
    int32_t b0 = 0;
    int64_t row0 = 0;
    for (int g = 0; g < h->G; g++) {
        const FeInfo &f = h->fe[g];
        if (f.kind != KIND_DENSE) continue;
        const int64_t nch = (N + chunk - 1) / chunk;
        const int32_t nblk = (int32_t)((nch + P->dense_wpb - 1) / P->dense_wpb);
        if (P->fused_dense && (unsigned)nblk != P->fused_blocks) return fail(FEHIP_EHIP, "dense partition mismatch");
        P->dense_fe.push_back(g);
        P->blk0.push_back(b0);
        P->drow0.push_back(row0);
        for (int64_t k = 0; k < f.n; k++) {
            FixItem fi{};
            fi.kind = 1;
            fi.fe = g;
            fi.level = (int32_t)k;
            fi.coef_off = h->off[g];
            fi.first = row0 + k * nblk;
            fi.count = nblk;
            items.push_back(fi);
        }
        b0 += nblk;
        row0 += f.n * nblk;
    }
    P->ndense = (int32_t)P->dense_fe.size();
    P->dense_blocks = b0;
    P->nitems = (int32_t)items.size();
    P->db_blocks = (unsigned)std::min<int64_t>(grid_for((N + Pq - 1) / Pq, 4 * kWavesPerBlock), 2048);
    P->seg_blocks = (unsigned)std::min<int64_t>(grid_for(P->total_tiles, kWavesPerBlock), 2048);
    P->csr_blocks = P->nseg > 1 ? (unsigned)std::min<int64_t>(grid_for(P->total_tiles - P->ntiles_s, kWavesPerBlock),
                                                             2048)
                                : 0u;
    P->fix_blocks = (unsigned)std::min<int64_t>(grid_for(std::max<int64_t>(P->nitems, h->n[h->s] * K / 64),
                                                         kWavesPerBlock),
                                                1024);
    P->upd_blocks = (unsigned)std::min<int64_t>(grid_for(h->p * K, kBlock), 1024);
    CHK(dalloc(h, &P->items, std::max<int32_t>(P->nitems, 1)));
    if (P->nitems)
        HIPCHK(hipMemcpyAsync(P->items, items.data(), sizeof(FixItem) * P->nitems, hipMemcpyHostToDevice, h->stream));
    std::vector<FixItem> its, ito, isd, ic, id;  // alive until the synchronize below
    {
        for (const FixItem &fi : items) {
            if (fi.fe != h->s) {
                ito.push_back(fi);
            } else if (fi.kind == 0) {
                FixItem z = fi;
                z.coef_off = 0;
                its.push_back(z);
            }
        }
        for (const FixItem &fi : items) {
            const bool dense = h->fe[fi.fe].kind == KIND_DENSE;
            if (fi.fe == h->s || (dense && P->fused_dense)) isd.push_back(fi);
            else ic.push_back(fi);
            if (fi.fe != h->s && dense && P->fused_dense) id.push_back(fi);
        }
        P->nitems_d = (int32_t)id.size();
        CHK(dalloc(h, &P->items_d, std::max<int32_t>(P->nitems_d, 1)));
        if (P->nitems_d)
            HIPCHK(hipMemcpyAsync(P->items_d, id.data(), sizeof(FixItem) * P->nitems_d, hipMemcpyHostToDevice,
                                  h->stream));
        P->nitems_sd = (int32_t)isd.size();
        P->nitems_c = (int32_t)ic.size();
        CHK(dalloc(h, &P->items_sd, std::max<int32_t>(P->nitems_sd, 1)));
        CHK(dalloc(h, &P->items_c, std::max<int32_t>(P->nitems_c, 1)));
        if (P->nitems_sd)
            HIPCHK(hipMemcpyAsync(P->items_sd, isd.data(), sizeof(FixItem) * P->nitems_sd, hipMemcpyHostToDevice,
                                  h->stream));
        if (P->nitems_c)
            HIPCHK(hipMemcpyAsync(P->items_c, ic.data(), sizeof(FixItem) * P->nitems_c, hipMemcpyHostToDevice,
                                  h->stream));
        P->nitems_s = (int32_t)its.size();
        P->nitems_o = (int32_t)ito.size();
        CHK(dalloc(h, &P->items_s, std::max<int32_t>(P->nitems_s, 1)));
        CHK(dalloc(h, &P->items_o, std::max<int32_t>(P->nitems_o, 1)));
        if (P->nitems_s)
            HIPCHK(hipMemcpyAsync(P->items_s, its.data(), sizeof(FixItem) * P->nitems_s, hipMemcpyHostToDevice,
                                  h->stream));
        if (P->nitems_o)
            HIPCHK(hipMemcpyAsync(P->items_o, ito.data(), sizeof(FixItem) * P->nitems_o, hipMemcpyHostToDevice,
                                  h->stream));
        CHK(dalloc_bytes(h, &P->zs, ts * h->n[h->s] * K));
        CHK(dalloc(h, &P->ops_one, 1));
        CHK(dalloc(h, &P->ops_neg, 1));
        CHK(dalloc(h, &P->ops_sub, 1));
    }
    CHK(dalloc(h, &P->sdirect, sdirect.size()));
    HIPCHK(hipMemcpyAsync(P->sdirect, sdirect.data(), sdirect.size(), hipMemcpyHostToDevice, h->stream));
    CHK(dalloc(h, &P->tileA, (size_t)std::max<int32_t>(P->total_tiles, 1) * K));
    CHK(dalloc(h, &P->tileB, (size_t)std::max<int32_t>(P->total_tiles, 1) * K));
    CHK(dalloc(h, &P->dpart, (size_t)std::max<int64_t>(row0, 1) * K));
    CHK(dalloc(h, &P->wsum, (size_t)h->n[h->s] * K));
    CHK(dalloc(h, &P->part_db, (size_t)P->db_blocks * K));
    CHK(dalloc(h, &P->part_fused, (size_t)P->fused_blocks * K));
    CHK(dalloc(h, &P->part_seg, (size_t)P->seg_blocks * K));
    CHK(dalloc(h, &P->part_fix, (size_t)2 * P->fix_blocks * K));  // two halves: the split fix-up
    CHK(dalloc(h, &P->part_upd, (size_t)P->upd_blocks * K));
    CHK(dalloc_bytes(h, &P->state, sizeof(LsmrState<double>) * K));
    CHK(dalloc(h, &P->ops_db, 1));
    CHK(dalloc(h, &P->ops_dtx, 1));
    CHK(dalloc(h, &P->ops_once, 1));
    CHK(dalloc(h, &P->all_done, 1));
    CHK(dalloc_bytes(h, &P->u, ts * N * K));
    CHK(dalloc_bytes(h, &P->r, ts * N * K));
    CHK(dalloc_bytes(h, &P->x, ts * h->p * K));
    CHK(dalloc_bytes(h, &P->v, ts * h->p * K));
    CHK(dalloc_bytes(h, &P->hh, ts * h->p * K));
    CHK(dalloc_bytes(h, &P->hbar, ts * h->p * K));
    if (h->factored) CHK(dalloc_bytes(h, &P->vt, ts * h->p * K));
    {
        std::vector<FixItem> dense_items;
        for (const FixItem &fi : items)
            if (fi.kind == 1) dense_items.push_back(fi);
        const int64_t tiles_c =
            h->sched_tiles_c > 0 ? h->sched_tiles_c : env_int("FEHIP_TILES_C", waves_c, 1, 1 << 20);
        CHK(build_pieces(h, P, ts, dense_items, tiles_c));
    }
    HIPCHK(hipStreamSynchronize(h->stream));
    h->plans[key] = pl.release();
    *out = P;
    return 0;
}


// Builds the piece form of plan P (see PieceForm) when the handle's shape
// allows it: factored (no interactions), the secondary sort key t is the only
// CSR FE, at most three FEs (compile-time FE positions in the fused pass), and
// pieces are at most half the observations (else the CSR pass over the
// observations reads less).  FEHIP_PIECES=0 keeps the CSR pass.
// dense_items: the plan's few-level FE items; tiles: target tile count of the
// piece launch (both FEs together).
static int build_pieces(fehip_solver *h, Plan *P, size_t ts, const std::vector<FixItem> &dense_items,
                        int64_t tiles) {
    PieceForm &pc = P->pc;
    const int s = h->s, t = h->t_sec;
    const int64_t N = h->N;
    // FE s and FE t without interactions (a piece's sum of sqrt(w) .* u then
    // serves both); other FEs may carry one (the cache form)
    if (t < 0 || h->fe[t].kind != KIND_CSR || h->G > 3 || h->fe[s].z_orig || h->fe[t].z_orig ||
        env_int("FEHIP_PIECES", 1, 0, 1) == 0)
        return 0;
    for (int g = 0; g < h->G; g++)
        if (g != t && h->fe[g].kind == KIND_CSR) return 0;
    const int64_t nwords = N / 64 + 2;
    uint64_t *heads = nullptr;
    uint32_t *cnt = nullptr, *wp = nullptr, *sums = nullptr;
    CHK(dalloc(h, &heads, nwords));
    CHK(dalloc(h, &cnt, nwords + 1));
    CHK(dalloc(h, &wp, nwords + 1));
    CHK(dalloc(h, &sums, (nwords + kSortTile - 1) / kSortTile + 2));
    hipLaunchKernelGGL(k_piece_heads, dim3(grid_for(nwords, kWavesPerBlock)), dim3(kBlock), 0, h->stream,
                       (const int32_t *)h->fe[s].ref_int, (const int32_t *)h->fe[t].ref_int, N, P->R, heads, nwords,
                       cnt);
    HIPCHK(hipGetLastError());
    CHK(exclusive_scan_u32(h, cnt, nwords, wp, true, sums));
    uint32_t total = 0;
    HIPCHK(hipMemcpyAsync(&total, wp + nwords, sizeof(uint32_t), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    dfree(h, cnt);
    dfree(h, sums);
    const int64_t np = (int64_t)total;
    // psum is addressed with 32-bit byte offsets in the fused pass
    if (np < 1 || 2 * np > N || (double)np * P->K * ts >= 4294967296.0) {
        dfree(h, heads);
        dfree(h, wp);
        return 0;
    }
    pc.np = np;
    pc.heads = heads;
    pc.wprefix = wp;
    int32_t *lev_t = nullptr, *off_s = nullptr, *off_t = nullptr;
    CHK(dalloc(h, &pc.piece_s, np));
    CHK(dalloc(h, &lev_t, np));
    hipLaunchKernelGGL(k_piece_levels, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, h->stream,
                       (const int32_t *)h->fe[s].ref_int, (const int32_t *)h->fe[t].ref_int, N,
                       (const uint64_t *)heads, (const uint32_t *)wp, pc.piece_s, lev_t);
    HIPCHK(hipGetLastError());
    // by FE s level: the pieces are already in that order
    const int64_t ns = h->n[s], nt = h->n[t];
    CHK(dalloc(h, &off_s, ns + 1));
    CHK(csr_products(h, pc.piece_s, np, ns, off_s, &pc.fs));
    // by FE t level: a stable sort of the pieces
    CHK(dalloc(h, &pc.pt_lev, np));
    CHK(dalloc(h, &pc.pt_perm, np));
    CHK(radix_sort(h, lev_t, np, nt, pc.pt_lev, pc.pt_perm));
    CHK(dalloc(h, &off_t, nt + 1));
    CHK(csr_products(h, pc.pt_lev, np, nt, off_t, &pc.ft));
    std::vector<int32_t> hs(ns + 1), ht(nt + 1);
    HIPCHK(hipMemcpyAsync(hs.data(), off_s, sizeof(int32_t) * (ns + 1), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(ht.data(), off_t, sizeof(int32_t) * (nt + 1), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    dfree(h, lev_t);
    dfree(h, off_s);
    dfree(h, off_t);
    free_sort_scratch(h);
    // tiles over the pieces, cut at level heads: half the CSR pass's target
    // tile count per FE, multiples of 64 pieces (the launch is short and
    // latency-bound: many short tiles keep more waves' loads in flight; the
    // autotuner's candidates span 64 to about 450 pieces per tile on the
    // benchmark panel); fix-up items
    const int64_t tgt = std::max<int64_t>(1, tiles / 2);
    const int64_t tp = std::max<int64_t>(64, ((np + tgt - 1) / tgt + 63) / 64 * 64);
    std::vector<int64_t> tsh;
    std::vector<FixItem> it1(dense_items), it2;
    int32_t t0 = 0;
    for (int q = 0; q < 2; q++) {
        const int g = q == 0 ? s : t;
        const std::vector<int32_t> &off = q == 0 ? hs : ht;
        const int64_t nlev = h->n[g];
        const size_t b0 = tsh.size();
        if (q == 1) {
            pc.ts0_t = (int64_t)b0;
            pc.tile0_t = t0;
        }
        cut_at_heads(off.data(), nlev, np, tp, tsh);
        const int64_t *tb = tsh.data() + b0;
        const int64_t ntl = (int64_t)(tsh.size() - b0) - 1;
        auto tile_of = [&](int64_t pos) -> int64_t {
            return (int64_t)(std::upper_bound(tb, tb + ntl + 1, pos) - tb) - 1;
        };
        int32_t nonempty = 0;
        for (int64_t k = 0; k < nlev; k++) {
            const int64_t st = off[k], en = off[k + 1];
            FixItem fi{};
            fi.fe = g;
            fi.level = (int32_t)k;
            fi.coef_off = h->off[g];
            if (en <= st) {
                fi.kind = 2;
                it1.push_back(fi);
                continue;
            }
            nonempty++;
            const int64_t tf = tile_of(st), tl = tile_of(en - 1);
            if (tf == tl) continue;
            fi.kind = 0;
            fi.first = t0 + tf;
            fi.count = (int32_t)(tl - tf);
            it2.push_back(fi);
        }
        (q == 0 ? pc.nsegs_s : pc.nsegs_t) = std::max<int32_t>(nonempty, 1);
        t0 += (int32_t)ntl;
    }
    pc.total_tiles = t0;
    pc.blocks = (unsigned)std::min<int64_t>(grid_for(t0, kWavesPerBlock), 4096);
    pc.fix1 = (unsigned)std::min<int64_t>(grid_for((int64_t)it1.size(), kWavesPerBlock), 1024);
    pc.fix2 = (unsigned)std::min<int64_t>(grid_for((int64_t)it2.size(), kWavesPerBlock), 1024);
    CHK(dalloc(h, &pc.tstart, tsh.size()));
    HIPCHK(hipMemcpyAsync(pc.tstart, tsh.data(), sizeof(int64_t) * tsh.size(), hipMemcpyHostToDevice, h->stream));
    pc.n1 = (int32_t)it1.size();
    pc.n2 = (int32_t)it2.size();
    CHK(dalloc(h, &pc.items1, std::max<int32_t>(pc.n1, 1)));
    CHK(dalloc(h, &pc.items2, std::max<int32_t>(pc.n2, 1)));
    if (pc.n1)
        HIPCHK(hipMemcpyAsync(pc.items1, it1.data(), sizeof(FixItem) * pc.n1, hipMemcpyHostToDevice, h->stream));
    if (pc.n2)
        HIPCHK(hipMemcpyAsync(pc.items2, it2.data(), sizeof(FixItem) * pc.n2, hipMemcpyHostToDevice, h->stream));
    // the projected iteration's lists: the same items without FE s's
    std::vector<FixItem> it1p, it2p;
    for (const FixItem &fi : it1)
        if (fi.fe != s) it1p.push_back(fi);
    for (const FixItem &fi : it2)
        if (fi.fe != s) it2p.push_back(fi);
    pc.n1p = (int32_t)it1p.size();
    pc.n2p = (int32_t)it2p.size();
    pc.fix1p = (unsigned)std::min<int64_t>(grid_for((int64_t)it1p.size(), kWavesPerBlock), 1024);
    pc.fix2p = (unsigned)std::min<int64_t>(grid_for((int64_t)it2p.size(), kWavesPerBlock), 1024);
    CHK(dalloc(h, &pc.items1p, std::max<int32_t>(pc.n1p, 1)));
    CHK(dalloc(h, &pc.items2p, std::max<int32_t>(pc.n2p, 1)));
    if (pc.n1p)
        HIPCHK(hipMemcpyAsync(pc.items1p, it1p.data(), sizeof(FixItem) * pc.n1p, hipMemcpyHostToDevice, h->stream));
    if (pc.n2p)
        HIPCHK(hipMemcpyAsync(pc.items2p, it2p.data(), sizeof(FixItem) * pc.n2p, hipMemcpyHostToDevice, h->stream));
    CHK(dalloc(h, &pc.tileA, (size_t)std::max<int32_t>(t0, 1) * P->K));
    CHK(dalloc(h, &pc.tileB, (size_t)std::max<int32_t>(t0, 1) * P->K));
    CHK(dalloc(h, &pc.part, (size_t)(pc.fix1 + pc.blocks + pc.fix2) * P->K));
    CHK(dalloc_bytes(h, &pc.psum, ts * np * P->K));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

static void destroy_pieces(fehip_solver *h, PieceForm &pc) {
    void *ptrs[] = {pc.heads, pc.wprefix, pc.piece_s, pc.pt_perm, pc.pt_lev, pc.fs.heads, pc.fs.wprefix, pc.fs.seg_level,
                    pc.ft.heads, pc.ft.wprefix, pc.ft.seg_level, pc.tstart, pc.tileA, pc.tileB, pc.part,
                    pc.items1, pc.items2, pc.items1p, pc.items2p, pc.psum};
    for (void *q : ptrs) dfree(h, q);
    pc = PieceForm{};
}

// LDS of the fused pass: the few-level FEs' accumulators when they fit 64 KB
// per workgroup (the rule get_plan applies)
static size_t fused_lds_bytes(const fehip_solver *h, size_t ts, int K) {
    const size_t accsz = ts == 8 ? sizeof(double) : sizeof(float);
    const size_t per_slot = ts == 8 && K >= 16 ? (size_t)K * accsz : (size_t)kWave * accsz;
    const size_t per_slot_fused = fused_cols(h, ts, K) > 1 ? (size_t)K * sizeof(double) : per_slot;
    int64_t nslot = 0;
    for (int g = 0; g < h->G; g++)
        if (h->fe[g].kind == KIND_DENSE) nslot += h->fe[g].n;
    int64_t maxlev = 1;
    for (int g = 0; g < h->G; g++)
        if (h->fe[g].kind == KIND_DENSE) maxlev = std::max<int64_t>(maxlev, h->fe[g].n);
    const size_t b = (size_t)nslot * per_slot * kWavesPerBlock;
    if (!(b <= 65536 && (size_t)maxlev * per_slot * kWavesPerBlock <= 65536 && fuse_dense_on(h))) return 0;
    const size_t acc = (size_t)nslot * per_slot_fused * kWavesPerBlock;
    const int32_t nvl = staged_levels(h, ts, K, acc);
    return nvl ? align16(acc) + (size_t)nvl * K * ts : acc;
}

// resident waves of a kernel on this device (all CUs, kBlock threads)
static int64_t resident_waves(const void *kernel, size_t lds) {
    int dev = 0, cus = 0, nb = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, kBlock, lds) != hipSuccess || nb < 1 || cus < 1) {
        (void)hipGetLastError();
        return 4096;
    }
    return (int64_t)nb * cus * kWavesPerBlock;
}

// the plan of one engine: tiles sized by the occupancy of its kernels
template <typename T, typename Acc, int K>
static int get_plan_for(fehip_solver *h, Plan **out) {
    const int key = K * 16 + (int)sizeof(T);
    auto it = h->plans.find(key);
    if (it != h->plans.end()) {
        *out = it->second;
        return 0;
    }
    const void *kf = h->factored ? (h->G == 3   ? (const void *)k_fused<T, Acc, K, 3, true>
                                    : h->G == 2 ? (const void *)k_fused<T, Acc, K, 2, true>
                                                : (const void *)k_fused<T, Acc, K, 0, true>)
                                 : (h->G == 3   ? (const void *)k_fused<T, Acc, K, 3, false>
                                    : h->G == 2 ? (const void *)k_fused<T, Acc, K, 2, false>
                                                : (const void *)k_fused<T, Acc, K, 0, false>);
    constexpr int WF = fused_wide_c<Acc, K, true>(), WN = fused_wide_c<Acc, K, false>();
    if constexpr (WF > 1) {
        if (h->factored && fused_cols(h, sizeof(T), K) == WF)
            kf = h->G == 3 ? (const void *)k_fused<T, Acc, K, 3, true, 0, WF>
                           : (const void *)k_fused<T, Acc, K, 2, true, 0, WF>;
    }
    if constexpr (WN > 1) {
        if (!h->factored && fused_cols(h, sizeof(T), K) == WN)
            kf = h->G == 3 ? (const void *)k_fused<T, Acc, K, 3, false, 0, WN>
                           : (const void *)k_fused<T, Acc, K, 2, false, 0, WN>;
    }
    const int64_t wf = resident_waves(kf, fused_lds_bytes(h, sizeof(T), K));
    const int64_t wc = resident_waves((const void *)k_dtx_seg<T, Acc, K>, 0);
    return get_plan(h, K, sizeof(T), out, wf, wc);
}

static void destroy_plan(fehip_solver *h, Plan *P) {
    if (P->graph) (void)hipGraphExecDestroy(P->graph);
    void *ptrs[] = {P->items, P->tileA, P->tileB, P->dpart, P->part_db, P->part_seg, P->part_fix, P->part_upd,
                    P->part_fused, P->wsum, P->sdirect, P->vt, P->state, P->ops_db, P->ops_dtx, P->ops_once, P->all_done, P->u, P->r, P->x,
                    P->v, P->hh, P->hbar, P->dY, P->tstart, P->items_s, P->items_o, P->items_sd, P->items_c,
                    P->zs, P->ops_one, P->ops_neg, P->ops_sub, P->items_d};
    for (void *q : ptrs) dfree(h, q);
    destroy_pieces(h, P->pc);
    delete P;
}

static void destroy_all_plans(fehip_solver *h) {
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (auto &kv : h->plans) destroy_plan(h, kv.second);
    h->plans.clear();
}

// ---------------------------------------------------------------------------
// small kernels of the solve
// ---------------------------------------------------------------------------
__global__ void k_set_ops(ColOps *ops, double a, double b, int K, int Kreal) {
    const int c = threadIdx.x;
    if (c < kMaxK) {
        ops->a[c] = a;
        ops->b[c] = b;
        ops->on[c] = c < Kreal ? 1 : 0;
    }
}

template <typename Tr>
__global__ void k_lsmr_setup(LsmrState<Tr> *st, int K, int Kreal, double atol, double btol, double conlim,
                             int64_t maxiter, int32_t *all_done) {
    const int c = threadIdx.x;
    if (c < K) {
        LsmrState<Tr> s{};
        s.atol = atol;
        s.btol = btol;
        s.ctol = conlim > 0 ? (Tr)(1.0 / conlim) : (Tr)0;
        s.maxiter = maxiter;
        s.active = c < Kreal ? 1 : 0;
        s.done = c < Kreal ? 0 : 1;
        st[c] = s;
    }
    if (c == 0) *all_done = 0;
}

// Y (column-major, caller's order, double) -> r (N x K, internal order, T),
// weighted by sqrt(w) (src/AbstractFixedEffectSolver.jl:51-55)
template <typename T>
__global__ void k_load_cols(const double *__restrict__ Y, int64_t ldY, int K, int Kreal,
                            const int32_t *__restrict__ perm_s, const double *__restrict__ sqrtw, int64_t N,
                            T *__restrict__ r, T *__restrict__ u) {
#pragma clang fp contract(off)
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N * K) return;
    const int64_t j = e / K;
    const int c = (int)(e % K);
    T val = (T)0;
    if (c < Kreal) {
        val = (T)Y[(int64_t)c * ldY + perm_s[j]];
        if (sqrtw) val = (T)((double)val * sqrtw[j]);
    }
    r[e] = val;
    if (u) u[e] = val;
}

// r -> Y, unweighted (:63-66)
template <typename T>
__global__ void k_store_cols(double *__restrict__ Y, int64_t ldY, int K, int Kreal,
                             const int32_t *__restrict__ perm_s, const double *__restrict__ sqrtw, int64_t N,
                             const T *__restrict__ r) {
#pragma clang fp contract(off)
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N * K) return;
    const int64_t j = e / K;
    const int c = (int)(e % K);
    if (c >= Kreal) return;
    T val = r[e];
    if (sqrtw) val = (T)((double)val / sqrtw[j]);
    Y[(int64_t)c * ldY + perm_s[j]] = (double)val;
}

// Tiled forms for K >= 8 (internal rows of 64 bytes or more): a workgroup
// takes 64 consecutive caller rows, reads each column's 64 values as one
// coalesced run into LDS, and writes each row's K values as one contiguous
// run at its internal position (pos_s); the store is the reverse.  The
// untiled forms gather every element of the column-major block through
// perm_s, one 8-byte read per fetched line at K = 32.  Same arithmetic.
constexpr int kTileRows = 64;
template <typename T>
__global__ __launch_bounds__(kBlock) void k_load_cols_t(const double *__restrict__ Y, int64_t ldY, int K, int Kreal,
                                                        const int32_t *__restrict__ pos_s,
                                                        const double *__restrict__ sqrtw, int64_t N,
                                                        T *__restrict__ r, T *__restrict__ u) {
#pragma clang fp contract(off)
    __shared__ double tile[kMaxK * (kTileRows + 1)];  // [column][row], padded against bank conflicts
    const int64_t j0 = (int64_t)blockIdx.x * kTileRows;
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int64_t jj = j0 + lane;
    for (int c = wave; c < Kreal; c += kWavesPerBlock) tile[c * (kTileRows + 1) + lane] = jj < N ? Y[c * ldY + jj] : 0.0;
    __syncthreads();
    for (int e = threadIdx.x; e < kTileRows * K; e += kBlock) {
        const int l = e / K, c = e % K;
        const int64_t j = j0 + l;
        if (j >= N) break;  // (rows in order)
        const int64_t i = pos_s[j];
        T val = (T)0;
        if (c < Kreal) {
            val = (T)tile[c * (kTileRows + 1) + l];
            if (sqrtw) val = (T)((double)val * sqrtw[i]);
        }
        r[i * K + c] = val;
        if (u) u[i * K + c] = val;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_store_cols_t(double *__restrict__ Y, int64_t ldY, int K, int Kreal,
                                                         const int32_t *__restrict__ pos_s,
                                                         const double *__restrict__ sqrtw, int64_t N,
                                                         const T *__restrict__ r) {
#pragma clang fp contract(off)
    __shared__ double tile[kMaxK * (kTileRows + 1)];
    const int64_t j0 = (int64_t)blockIdx.x * kTileRows;
    for (int e = threadIdx.x; e < kTileRows * K; e += kBlock) {
        const int l = e / K, c = e % K;
        const int64_t j = j0 + l;
        if (j >= N) break;
        if (c >= Kreal) continue;
        const int64_t i = pos_s[j];
        T val = r[i * K + c];
        if (sqrtw) val = (T)((double)val / sqrtw[i]);
        tile[c * (kTileRows + 1) + l] = (double)val;
    }
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int64_t jj = j0 + lane;
    if (jj < N)
        for (int c = wave; c < Kreal; c += kWavesPerBlock) Y[c * ldY + jj] = tile[c * (kTileRows + 1) + lane];
}

template <typename T>
__global__ void k_copy(const T *__restrict__ a, T *__restrict__ b, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) b[e] = a[e];
}

// y[i*K + c] *= d[i]
template <typename T>
__global__ void k_scale_rows(T *__restrict__ y, const T *__restrict__ d, int64_t N, int K) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < N * K) y[e] = y[e] * d[e / K];
}

template <typename T>
__global__ void k_square(const T *__restrict__ a, int64_t n, T *__restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i] * a[i];
}

template <typename T>
__global__ void k_fill(T *__restrict__ a, int64_t n, T val) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) a[e] = val;
}

// x (p doubles) -> T ; T -> double scaled (x .*= scale, :155-157)
template <typename T>
__global__ void k_to_T(const double *__restrict__ a, T *__restrict__ b, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) b[e] = (T)a[e];
}
template <typename T>
__global__ void k_to_double(const T *__restrict__ a, double *__restrict__ b, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) b[e] = (double)a[e];
}
template <typename T>
__global__ void k_scale_coef(const T *__restrict__ x, const T *__restrict__ scale, int64_t n, double *__restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) out[e] = (double)(x[e] * scale[e]);
}

// dst[i] = src[map[i]]: a block between two numberings
template <typename V>
__global__ void k_gather_map(const V *__restrict__ src, const int32_t *__restrict__ map, int64_t n, V *__restrict__ dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[map[i]];
}
// rows of W floats: dst[i*W + j] = src[map[i]*W + j]
__global__ void k_gather_rows(const float *__restrict__ src, const int32_t *__restrict__ map, int64_t n, int W,
                             float *__restrict__ dst) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n * W) dst[e] = src[(int64_t)map[e / W] * W + e % W];
}

// The FE s block of a device coefficient vector (p doubles), in place, from
// the internal numbering to the caller's (to_caller) or back.
static int s_block_permute(fehip_solver *h, double *dc, bool to_caller) {
    if (!h->s_int_of) return 0;
    double *blk = dc + h->off[h->s];
    const int64_t ns = h->n[h->s];
    hipLaunchKernelGGL(k_gather_map<double>, dim3(grid_for(ns, kBlock)), dim3(kBlock), 0, h->stream,
                       (const double *)blk, (const int32_t *)(to_caller ? h->s_int_of : h->s_ext_of), ns, h->s_tmp);
    HIPCHK(hipMemcpyAsync(blk, h->s_tmp, sizeof(double) * ns, hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(hipGetLastError());
    return 0;
}
static int s_block_to_caller(fehip_solver *h, double *dc) { return s_block_permute(h, dc, true); }
static int s_block_to_internal(fehip_solver *h, double *dc) { return s_block_permute(h, dc, false); }

// ---------------------------------------------------------------------------
// warm-start support: per-level features and initial iterates
// ---------------------------------------------------------------------------

// degree of level k of FE s: distinct levels of FE t among its observations
// (the internal order sorts each FE s run by FE t, so changes are counted)
__global__ void k_degree_sorted(const int32_t *__restrict__ off, const int32_t *__restrict__ ref_t, int64_t nlev,
                                int32_t *__restrict__ deg) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nlev) return;
    const int32_t a = off[k], b = off[k + 1];
    int32_t d = b > a ? 1 : 0;
    if (ref_t)
        for (int32_t q = a + 1; q < b; q++) d += ref_t[q] != ref_t[q - 1];
    deg[k] = d;
}

// degree of level k of a CSR FE: distinct FE s levels among its observations
// (its internal CSR lists observations in internal order, i.e. by FE s)
__global__ void k_degree_csr(const int32_t *__restrict__ off, const int32_t *__restrict__ perm,
                             const int32_t *__restrict__ ref_s, int64_t nlev, int32_t *__restrict__ deg) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nlev) return;
    const int32_t a = off[k], b = off[k + 1];
    int32_t d = b > a ? 1 : 0;
    for (int32_t q = a + 1; q < b; q++) d += ref_s[perm[q]] != ref_s[perm[q - 1]];
    deg[k] = d;
}

__global__ void k_degree_count(const int32_t *__restrict__ off, int64_t nlev, int32_t *__restrict__ deg) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nlev) deg[k] = off[k + 1] - off[k];
}

struct WsFE {
    const double *scale;     // n
    const int32_t *off;      // caller-order offsets (counts)
    int64_t off_p;           // first level in p
    int32_t n, interacted;
};
struct WsParams {
    int32_t G, pad;
    WsFE fe[kMaxG];
    const int32_t *degree;   // p
};

// features of level e (FP32, 16 per level):
//  0 log1p(count)   1 scale*xref/sigma (weighted group mean / sigma)
//  2 log1p(degree)  3 scale*xap/sigma (one alternating-projection pass)
//  4 interacted     5 -log(scale) = log(sum w z^2)/2   6..13 one-hot FE
//  14 empty level   15 0
template <typename T>
__global__ void k_ws_features(WsParams prm, int64_t p, const T *__restrict__ xref, const T *__restrict__ xap,
                              const double *__restrict__ sig, float *__restrict__ feats, double *__restrict__ coef_ap) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= p) return;
    int g = 0;
    while (g + 1 < prm.G && e >= prm.fe[g + 1].off_p) g++;
    const WsFE f = prm.fe[g];
    const int64_t k = e - f.off_p;
    const double sc = f.scale[k];
    const double sigma = sig[0] > 0 ? sig[0] : 1.0;
    const int32_t cnt = f.off[k + 1] - f.off[k];
    float *o = feats + e * 16;
    o[0] = log1pf((float)cnt);
    o[1] = (float)(sc * (double)xref[e] / sigma);
    o[2] = log1pf((float)prm.degree[e]);
    o[3] = (float)(sc * (double)xap[e] / sigma);
    o[4] = (float)f.interacted;
    o[5] = sc > 0 ? (float)(-log(sc)) : 0.f;
    for (int j = 0; j < 8; j++) o[6 + j] = (j == (g < 8 ? g : 7)) ? 1.f : 0.f;
    o[14] = cnt == 0 ? 1.f : 0.f;
    o[15] = 0.f;
    if (coef_ap) coef_ap[e] = sc * (double)xap[e];
}

// x[e*K + c] = initial iterate in scaled coordinates from unscaled
// coefficients b0 (x = b0 / scale; 0 on levels with scale 0)
template <typename T>
__global__ void k_x0_from_coef(const double *__restrict__ b0, WsParams prm, int64_t p, T *__restrict__ x, int K,
                               int c) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= p) return;
    int g = 0;
    while (g + 1 < prm.G && e >= prm.fe[g + 1].off_p) g++;
    const double sc = prm.fe[g].scale[e - prm.fe[g].off_p];
    x[e * K + c] = sc > 0 ? (T)(b0[e] / sc) : (T)0;
}

// x[e*K + c] = xap[e] + pred[e] * sigma / scale  (learned correction of the
// alternating-projection start, in scaled coordinates)
template <typename T>
__global__ void k_x0_from_pred(const T *__restrict__ xap, const float *__restrict__ pred, const double *__restrict__ sig,
                               WsParams prm, int64_t p, double *__restrict__ x0) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= p) return;
    int g = 0;
    while (g + 1 < prm.G && e >= prm.fe[g + 1].off_p) g++;
    const double sc = prm.fe[g].scale[e - prm.fe[g].off_p];
    const double sigma = sig[0] > 0 ? sig[0] : 1.0;
    x0[e] = sc > 0 ? (double)xap[e] + (double)pred[e] * sigma / sc : 0.0;
}

// column c of the solve's x (p x K, level-major) <- x0
template <typename T>
__global__ void k_put_col(const double *__restrict__ x0, int64_t p, T *__restrict__ x, int K, int c) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < p) x[e * K + c] = (T)x0[e];
}

// sig[0] = sqrt(sum b^2 / n) over one column (single workgroup, fixed order)
template <typename T>
__global__ __launch_bounds__(kFinBlock) void k_col_rms(const T *__restrict__ b, int64_t n, double *__restrict__ sig) {
    __shared__ double red[kFinBlock];
    double s0 = 0, s1 = 0;
    int64_t j = threadIdx.x;
    for (; j + kFinBlock < n; j += 2 * kFinBlock) {
        s0 += (double)b[j] * (double)b[j];
        s1 += (double)b[j + kFinBlock] * (double)b[j + kFinBlock];
    }
    for (; j < n; j += kFinBlock) s0 += (double)b[j] * (double)b[j];
    red[threadIdx.x] = s0 + s1;
    __syncthreads();
    for (int w = kFinBlock / 2; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) sig[0] = n > 0 ? sqrt(red[0] / (double)n) : 0.0;
}


// ---------------------------------------------------------------------------
// engine: launches of one precision / column count
// ---------------------------------------------------------------------------
template <typename T, typename Acc, typename Tr, int K>
struct Engine {
    using TT = T;
    using AccT = Acc;
    using TrT = Tr;
    static constexpr int kK = K;

    // target <- b*target + a*A*coef ; mode != NONE leaves ||target||^2 partials
    static void db(fehip_solver *h, Plan *P, int mode, const ColOps *ops, const T *coef, T *target,
                   const int32_t *early_exit) {
        DbParams<T> p{};
        p.N = h->N;
        p.G = h->G;
        for (int g = 0; g < h->G; g++) {
            p.refs[g] = h->fe[g].ref_int;
            p.cache[g] = (const T *)h->fe[g].cache_int[pidx<T>()];
            p.coef_off[g] = h->off[g];
        }
        p.coef = coef;
        p.target = target;
        p.ops = ops;
        p.partials = P->part_db;
        p.all_done = early_exit;
        p.mode = mode;
        hipLaunchKernelGGL((k_db<T, K>), dim3(P->db_blocks), dim3(kBlock), 0, h->stream, p);
    }

    // beta = ||u|| and the D'u coefficients (partials of k_db, or of k_fused)
    // in the iteration (MODE_ITER) it first runs the previous iteration's
    // stopping test on the vector update's ||x|| partials (k_fin_db)
    static void fin_db(fehip_solver *h, Plan *P, int mode, const int32_t *early_exit, bool fused = false) {
        const bool it = mode == MODE_ITER;
        hipLaunchKernelGGL((k_fin_db<Tr, K>), dim3(1), dim3(kFinBlock), 0, h->stream,
                           (const double *)(fused ? P->part_fused : P->part_db),
                           (int64_t)(fused ? P->fused_blocks : P->db_blocks), mode, P->state, P->ops_dtx, early_exit,
                           it ? (const double *)P->part_upd : (const double *)nullptr, (int64_t)(it ? P->upd_blocks : 0),
                           P->ops_db, P->all_done);
    }

    // alpha = ||v|| and the plane rotations (partials of k_dtx_seg over all
    // segmented FEs, or over the CSR FEs only after k_fused, and the fix-up)
    // split: the partials of the split fix-up (both halves)
    static void fin_dtx(fehip_solver *h, Plan *P, int mode, const int32_t *early_exit, bool fused = false,
                        bool split = false) {
        const int64_t n0 = fused ? P->csr_blocks : (P->nseg ? P->seg_blocks : 0);
        hipLaunchKernelGGL((k_fin_dtx<Tr, K>), dim3(1), dim3(kFinBlock), 0, h->stream, (const double *)P->part_seg,
                           n0, (const double *)P->part_fix, (int64_t)P->fix_blocks * (split ? 2 : 1), mode, P->state,
                           P->ops_db, P->all_done, early_exit);
    }

    // ||x|| and the stopping rules

    // coef <- b*coef + a*A'y : segmented tiles, few-level chunks, fix-up.
    // csr_only: the tiles of the CSR FEs only (FE s is done by k_fused).
    // raw: y is the iteration's stored sqrt(w) .* u (factored form): sums of y
    // as it is, scaled per level (k_dtx_seg RAW)
    static DtxSegParams<T> seg_params(fehip_solver *h, Plan *P, int mode, const ColOps *ops, const T *y, T *coef,
                                      const int32_t *early_exit, bool csr_only) {
        DtxSegParams<T> p{};
        p.N = h->N;
        p.nseg = P->nseg;
        p.first_tile = csr_only ? P->ntiles_s : 0;
        p.total_tiles = P->total_tiles;
        for (int q = 0; q < P->nseg; q++) {
            const FeInfo &f = h->fe[P->seg_fe[q]];
            SegFE<T> &s = p.fe[q];
            s.perm = f.kind == KIND_SORTED ? h->ident : f.iperm;
            s.cache = (const T *)f.cache_csr[pidx<T>()];
            s.heads = f.heads;
            s.wprefix = f.wprefix;
            s.nsegs = std::max<int32_t>(P->seg_nsegs[q], 1);
            s.seg_level = f.seg_level;
            s.lscale = (const T *)f.scale_T[pidx<T>()];
            s.coef_off = h->off[P->seg_fe[q]];
            s.tstart = P->tstart + P->seg_ts0[q];
            s.tile0 = P->tile0[q];
            s.ntiles = (q + 1 < P->nseg ? P->tile0[q + 1] : P->total_tiles) - P->tile0[q];
        }
        p.u = y;
        p.v = coef;
        p.tileA = P->tileA;
        p.tileB = P->tileB;
        p.ops = ops;
        p.partials = P->part_seg;
        p.all_done = early_exit;
        p.mode = mode;
        return p;
    }

    static void dtx_seg(fehip_solver *h, Plan *P, int mode, const ColOps *ops, const T *y, T *coef,
                        const int32_t *early_exit, bool csr_only = false, bool raw = false) {
        if (!P->nseg || (csr_only && !P->csr_blocks)) return;
        const DtxSegParams<T> p = seg_params(h, P, mode, ops, y, coef, early_exit, csr_only);
        const dim3 gr(csr_only ? P->csr_blocks : P->seg_blocks);
        if (raw) hipLaunchKernelGGL((k_dtx_seg<T, Acc, K, true>), gr, dim3(kBlock), 0, h->stream, p);
        else hipLaunchKernelGGL((k_dtx_seg<T, Acc, K>), gr, dim3(kBlock), 0, h->stream, p);
    }

    // u stored as sqrt(w) .* u in this engine's iteration (k_fused's rule)
    static bool ut(const fehip_solver *h) { return h->factored && stores_scaled_u<K, true>(); }
    // the iteration runs in the piece form (plain LSMR on a plan that has it)
    static bool use_pieces(const fehip_solver *h, const Plan *P) { return P->pc.np > 0 && !h->fwl_now; }
    // the projected iteration runs in the piece form (its FE t half)
    static bool use_pieces_proj(const fehip_solver *h, const Plan *P) {
        return P->pc.np > 0 && h->fwl_now && h->pieces_proj;
    }
    // ... and stores u scaled (never in the piece forms)
    static bool ut_now(const fehip_solver *h, const Plan *P) {
        return ut(h) && !use_pieces(h, P) && !use_pieces_proj(h, P);
    }

    // the fused pass on this handle: factored or cache form, then its lane
    // map (fused_cols / fused_wide_c: the columns per lane), then the FE count
    template <int PM>
    static void launch_fused(fehip_solver *h, Plan *P, const FusedParams<T> &p) {
        if (h->factored) launch_fused_f<PM, true>(h, P, p);
        else launch_fused_f<PM, false>(h, P, p);
    }
    template <int PM, bool FACT>
    static void launch_fused_f(fehip_solver *h, Plan *P, const FusedParams<T> &p) {
        constexpr int WC = fused_wide_c<Acc, K, FACT>();
        if constexpr (WC > 1) {
            if (fused_cols(h, sizeof(T), K) == WC) return launch_fused_g<PM, FACT, WC>(h, P, p);
        }
        launch_fused_g<PM, FACT, 1>(h, P, p);
    }
    template <int PM, bool FACT, int C>
    static void launch_fused_g(fehip_solver *h, Plan *P, const FusedParams<T> &p) {
        const dim3 gr(P->fused_blocks), bl(kBlock);
        // the sums-only pass has no few-level sums (only the staged vector)
        const size_t lds = PM == 1 ? (size_t)p.nvl * K * sizeof(T) : P->fused_lds;
        // the packed ref stream read once and the few-level vector staged in
        // LDS (PK) in the piece forms and the projected sums-only pass
        constexpr bool PKF = PM == 1 || PM == 3 || PM == 4;
        if (h->G == 3 && PKF && p.packed && p.nvl > 0)
            hipLaunchKernelGGL((k_fused<T, Acc, K, 3, FACT, PM, C, PKF>), gr, bl, lds, h->stream, p);
        else if (h->G == 3) hipLaunchKernelGGL((k_fused<T, Acc, K, 3, FACT, PM, C>), gr, bl, lds, h->stream, p);
        else if (h->G == 2) hipLaunchKernelGGL((k_fused<T, Acc, K, 2, FACT, PM, C>), gr, bl, lds, h->stream, p);
        else if constexpr (C == 1 && PM < 3)  // any FE count (the piece forms have two or three)
            hipLaunchKernelGGL((k_fused<T, Acc, K, 0, FACT, PM>), gr, bl, lds, h->stream, p);
    }

    // u <- b*u + a*A v with the raw D'u sums of FE s and of the dense FEs.
    // pm (projected iteration): 1 = only the FE s sums of D*v over the
    // other FEs; 2 = u <- b*u + a*(A_o v - A_s zs), zs the FE s correction,
    // without FE s sums
    static void fused(fehip_solver *h, Plan *P, const ColOps *ops, const T *coef, T *target, int pm = 0) {
        FusedParams<T> p{};
        p.N = h->N;
        p.R = P->R;
        p.G = h->G;
        p.s = 0;
        p.ntiles = P->ntiles_s;
        // kernel FE order: FE s first; the first few-level FE of the pass
        // last (compile-time positions in k_fused); the rest between
        const bool dense_in = P->fused_dense && P->ndense > 0;
        const int dlast = dense_in ? P->dense_fe[0] : -1;
        int kord[kMaxG], kpos[kMaxG], nk = 0;
        kord[nk++] = h->s;
        for (int g = 0; g < h->G; g++)
            if (g != h->s && g != dlast) kord[nk++] = g;
        if (dlast >= 0) kord[nk++] = dlast;
        for (int k = 0; k < h->G; k++) {
            const int g = kord[k];
            kpos[g] = k;
            p.refs[k] = h->fe[g].ref_int;
            p.cache[k] = (const T *)h->fe[g].cache_int[pidx<T>()];
            p.coef_off[k] = h->off[g];
            p.gsrc[k] = (h->factored ? (const T *)P->vt : coef) + h->off[g] * K;
        }
        // one ref stream for the last two FEs when they are the packed pair
        // (the kernel reads it twice at the same addresses and unpacks)
        if (dense_in && h->G == 3 && h->packed && kord[1] == h->pack_g && kord[2] == h->pack_d) {
            p.refs[1] = p.refs[2] = h->packed;
            p.packed = 1;
            p.nvl = P->nvl;
        }
        if (dense_in) {
            p.ndense = P->ndense;
            p.nslot = P->nslot;
            for (int q = 0; q < P->ndense; q++) {
                p.dense_g[q] = kpos[P->dense_fe[q]];
                p.dense_slot0[q] = P->slot0[q];
                p.dense_row0[q] = P->drow0[q];
            }
        }
        const FeInfo &fs = h->fe[h->s];
        p.heads = fs.heads;
        p.wprefix = fs.wprefix;
        p.seg_level = fs.seg_level;
        p.nsegs_s = std::max<int32_t>(P->seg_nsegs[0], 1);
        p.coef = coef;
        p.target = target;
        p.wsum = P->wsum;
        p.tileA = P->tileA;
        p.tileB = P->tileB;
        p.dpart = P->dpart;
        p.ops = ops;
        p.partials = P->part_fused;
        p.all_done = P->all_done;
        p.sqrtw = (const T *)h->sqrtw_T[pidx<T>()];
        p.vt = (const T *)P->vt;
        if (pm == 1) {
            p.ndense = 0;
            p.nslot = 0;
            p.zsw = (T *)P->zs;
            p.gsrc[0] = (const T *)(h->factored ? h->scale2_T[pidx<T>()] : h->fe[h->s].scale_T[pidx<T>()]);
            launch_fused<1>(h, P, p);
        } else if (pm == 3 || pm == 4) {
            const PieceForm &pc = P->pc;
            p.heads = pc.heads;
            p.wprefix = pc.wprefix;
            p.seg_level = pc.piece_s;
            p.nsegs_s = (int32_t)pc.np;
            p.psum = (T *)pc.psum;
            if (pm == 4) {  // the projected u pass: FE s's correction at its level
                p.gsrc[0] = (const T *)P->zs;
                launch_fused<4>(h, P, p);
            } else {
                launch_fused<3>(h, P, p);
            }
        } else if (pm == 2) {
            p.gsrc[0] = (const T *)P->zs;
            launch_fused<2>(h, P, p);
        } else {
            launch_fused<0>(h, P, p);
        }
    }

    // y <- y - A_s A_s' y: the projection off FE s (the columns of A_s are
    // orthonormal: disjoint supports, unit norms by scale!), through zs.
    // mode MODE_INIT leaves the ||y||^2 partials of k_db for fin_db.  Uses
    // ops_one = (1, 0) and ops_sub = (-1, 1), set by start().
    static void proj_s(fehip_solver *h, Plan *P, T *y, int mode) {
        s_sums(h, P, y);
        s_apply(h, P, y, mode);
    }

    // zs = A_s' y (the FE s block of A'y; ops_one)
    static void s_sums(fehip_solver *h, Plan *P, const T *y) {
        const FeInfo &f = h->fe[h->s];
        DtxSegParams<T> q{};
        q.N = h->N;
        q.nseg = 1;
        q.first_tile = 0;
        q.total_tiles = P->ntiles_s;
        SegFE<T> &sf = q.fe[0];
        sf.perm = f.kind == KIND_SORTED ? h->ident : f.iperm;
        sf.cache = (const T *)f.cache_csr[pidx<T>()];
        sf.heads = f.heads;
        sf.wprefix = f.wprefix;
        sf.nsegs = std::max<int32_t>(P->seg_nsegs[0], 1);
        sf.seg_level = f.seg_level;
        sf.coef_off = 0;
        sf.tstart = P->tstart + P->seg_ts0[0];
        sf.tile0 = 0;
        sf.ntiles = P->ntiles_s;
        q.u = y;
        q.v = (T *)P->zs;
        q.tileA = P->tileA;
        q.tileB = P->tileB;
        q.ops = P->ops_one;
        q.partials = P->part_seg;
        q.all_done = nullptr;
        q.mode = MODE_NONE;
        hipLaunchKernelGGL((k_dtx_seg<T, Acc, K>),
                           dim3((unsigned)std::min<int64_t>(grid_for(P->ntiles_s, kWavesPerBlock), 2048)),
                           dim3(kBlock), 0, h->stream, q);
        FixParams<T> fx{};
        fx.nitems = P->nitems_s;
        fx.items = P->items_s;
        fx.tileA = P->tileA;
        fx.tileB = P->tileB;
        fx.dpart = P->dpart;
        fx.v = (T *)P->zs;
        fx.ops = P->ops_one;
        fx.partials = P->part_fix;
        fx.all_done = nullptr;
        fx.mode = MODE_NONE;
        hipLaunchKernelGGL((k_dtx_fixup<T, K>), dim3(P->fix_blocks), dim3(kBlock), 0, h->stream, fx);
    }

    // y <- y - A_s zs (ops_sub); mode MODE_INIT leaves ||y||^2 partials
    static void s_apply(fehip_solver *h, Plan *P, T *y, int mode) {
        const FeInfo &f = h->fe[h->s];
        DbParams<T> d{};
        d.N = h->N;
        d.G = 1;
        d.refs[0] = f.ref_int;
        d.cache[0] = (const T *)f.cache_int[pidx<T>()];
        d.coef_off[0] = 0;
        d.coef = (const T *)P->zs;
        d.target = y;
        d.ops = P->ops_sub;
        d.partials = P->part_db;
        d.all_done = nullptr;
        d.mode = mode;
        hipLaunchKernelGGL((k_db<T, K>), dim3(P->db_blocks), dim3(kBlock), 0, h->stream, d);
    }

    // projected iteration, after the sums-only fused pass (which wrote zs for
    // the FE s levels inside a tile): zs = -A_s' t, t = A_o v, for the levels
    // whose segment crosses tiles.  Under the factored form the raw sums of
    // sqrt(w)*t take scale_s^2: the FE s scale once for A_s' and once for the
    // A_s that the second fused pass applies to zs.
    static void fix_s(fehip_solver *h, Plan *P) {
        FixParams<T> f{};
        f.nitems = P->nitems_s;
        f.items = P->items_s;
        f.tileA = P->tileA;
        f.tileB = P->tileB;
        f.dpart = P->dpart;
        f.v = (T *)P->zs;
        f.ops = P->ops_neg;
        f.partials = P->part_fix;
        f.all_done = P->all_done;
        f.mode = MODE_NONE;
        f.s = h->s;
        if (h->factored) f.fscale[h->s] = (const T *)h->scale2_T[pidx<T>()];
        hipLaunchKernelGGL((k_dtx_fixup<T, K>), dim3(P->fix_blocks), dim3(kBlock), 0, h->stream, f);
    }

    static void dtx_dense(fehip_solver *h, Plan *P, const ColOps *ops, const T *y, const int32_t *early_exit,
                          bool raw = false) {
        if (!P->ndense) return;
        DtxDenseParams<T> p{};
        p.N = h->N;
        p.ndense = P->ndense;
        p.chunk = P->dense_chunk;
        for (int q = 0; q < P->ndense; q++) {
            const FeInfo &f = h->fe[P->dense_fe[q]];
            DenseFE<T> &d = p.fe[q];
            d.refs = f.ref_int;
            d.cache = (const T *)f.cache_int[pidx<T>()];
            d.nlev = (int32_t)f.n;
            d.blk0 = P->blk0[q];
            d.nblk = (q + 1 < P->ndense ? P->blk0[q + 1] : P->dense_blocks) - P->blk0[q];
            d.coef_off = h->off[P->dense_fe[q]];
            d.dpart_row0 = P->drow0[q];
        }
        p.u = y;
        p.dpart = P->dpart;
        p.ops = ops;
        p.all_done = early_exit;
        const size_t lds = (size_t)P->dense_lds_wave * P->dense_wpb;
        if (raw)
            hipLaunchKernelGGL((k_dtx_dense<T, Acc, K, true>), dim3(P->dense_blocks), dim3(kWave * P->dense_wpb), lds,
                               h->stream, p);
        else
            hipLaunchKernelGGL((k_dtx_dense<T, Acc, K>), dim3(P->dense_blocks), dim3(kWave * P->dense_wpb), lds,
                               h->stream, p);
    }

    // proj: the projected iteration's form -- the items of the FEs other
    // than s only, FE s untouched (its coefficients stay zero)
    // part (plain fused iteration only): 1 = FE s streaming and the FE s /
    // fused few-level items (on `st`, beside the CSR pass), 2 = the other
    // items into the second half of the partials
    static FixParams<T> fix_params(fehip_solver *h, Plan *P, int mode, const ColOps *ops, T *coef,
                                   const int32_t *early_exit, bool fused, bool proj, int part) {
        FixParams<T> f{};
        f.nitems = proj ? P->nitems_o : P->nitems;
        f.items = proj ? P->items_o : P->items;
        if (part == 1) {
            f.nitems = proj ? P->nitems_d : P->nitems_sd;
            f.items = proj ? P->items_d : P->items_sd;
        } else if (part == 2) {
            f.nitems = P->nitems_c;
            f.items = P->items_c;
        }
        f.tileA = P->tileA;
        f.tileB = P->tileB;
        f.dpart = P->dpart;
        f.v = coef;
        f.ops = ops;
        f.partials = P->part_fix + (part == 2 ? (size_t)P->fix_blocks * K : 0);
        f.all_done = early_exit;
        f.mode = mode;
        if (fused) {
            if (!proj && part != 2) {
                f.wsum = P->wsum;
                f.sdirect = P->sdirect;
                f.ns = h->n[h->s];
                f.s_off = h->off[h->s];
                f.s = h->s;
            }
            if (h->factored) {
                // raw sums of sqrt(w) .* u take their scale here: FE s's and the
                // fused few-level FEs' always, the CSR FEs' and the separate
                // few-level pass's when u is stored scaled
                for (int g = 0; g < h->G; g++) {
                    const bool dense = h->fe[g].kind == KIND_DENSE;
                    const bool raw = g == h->s ? !proj : (dense ? (P->fused_dense || ut(h)) : ut(h));
                    if (raw) f.fscale[g] = (const T *)h->fe[g].scale_T[pidx<T>()];
                }
            }
        }
        return f;
    }

    static void dtx_fixup(fehip_solver *h, Plan *P, int mode, const ColOps *ops, T *coef,
                          const int32_t *early_exit, bool fused = false, bool proj = false, int part = 0,
                          hipStream_t st = nullptr) {
        const FixParams<T> f = fix_params(h, P, mode, ops, coef, early_exit, fused, proj, part);
        hipLaunchKernelGGL((k_dtx_fixup<T, K>), dim3(P->fix_blocks), dim3(kBlock), 0, st ? st : h->stream, f);
    }

    // the iteration's CSR pass and the fix-up part that needs only the fused
    // pass (part 1: FE s, few-level FEs) in one launch, fix-up blocks first;
    // part 2 (segments crossing CSR tiles) is launched after it.  Norm
    // partials as in the split form (fin_dtx split).
    static void seg_fix(fehip_solver *h, Plan *P, const ColOps *ops, const T *y, T *coef,
                        const int32_t *early_exit, bool proj = false) {
        const DtxSegParams<T> p = seg_params(h, P, MODE_ITER, ops, y, coef, early_exit, true);
        const FixParams<T> f = fix_params(h, P, MODE_ITER, ops, coef, early_exit, true, proj, 1);
        const int nfix = (int)P->fix_blocks;
        const dim3 gr(P->fix_blocks + P->csr_blocks);
        if (ut(h)) hipLaunchKernelGGL((k_dtx_seg_fix<T, Acc, K, true>), gr, dim3(kBlock), 0, h->stream, p, f, nfix);
        else hipLaunchKernelGGL((k_dtx_seg_fix<T, Acc, K, false>), gr, dim3(kBlock), 0, h->stream, p, f, nfix);
    }

    // the iteration takes the merged launch (FEs beside FE s in CSR form)
    static bool use_seg_fix(const fehip_solver *h, const Plan *P) { return h->seg_fix && P->csr_blocks; }

    // Piece form: D'u of FE s (its contiguous pieces, identity order) and of
    // FE t (its pieces through pt_perm) as segmented sums of the piece sums,
    // each level's sum taking its FE's scale (RAW)
    static DtxSegParams<T> piece_seg_params(fehip_solver *h, Plan *P, const ColOps *ops, T *coef,
                                            const int32_t *early_exit, bool proj = false) {
        const PieceForm &pc = P->pc;
        DtxSegParams<T> p{};
        p.N = pc.np;
        p.nseg = 2;
        p.first_tile = proj ? pc.tile0_t : 0;  // projected: FE t's tiles only
        p.total_tiles = pc.total_tiles;
        for (int q = 0; q < 2; q++) {
            const int g = q == 0 ? h->s : h->t_sec;
            const FeInfo &fi = q == 0 ? pc.fs : pc.ft;
            SegFE<T> &e = p.fe[q];
            e.perm = q == 0 ? h->ident : pc.pt_perm;
            e.cache = nullptr;  // RAW: not read
            e.heads = fi.heads;
            e.wprefix = fi.wprefix;
            e.seg_level = fi.seg_level;
            e.plev = q == 0 ? pc.piece_s : pc.pt_lev;  // pieces by FE s level are in piece order
            e.nsegs = q == 0 ? pc.nsegs_s : pc.nsegs_t;
            e.lscale = (const T *)h->fe[g].scale_T[pidx<T>()];
            e.coef_off = h->off[g];
            e.tstart = pc.tstart + (q == 0 ? 0 : pc.ts0_t);
            e.tile0 = q == 0 ? 0 : pc.tile0_t;
            e.ntiles = q == 0 ? pc.tile0_t : pc.total_tiles - pc.tile0_t;
        }
        p.u = (const T *)pc.psum;
        p.v = coef;
        p.tileA = pc.tileA;
        p.tileB = pc.tileB;
        p.ops = ops;
        p.partials = pc.part + (size_t)(proj ? pc.fix1p : pc.fix1) * K;
        p.all_done = early_exit;
        p.mode = MODE_ITER;
        return p;
    }

    // part 1: few-level FEs' levels and levels without observations (first
    // blocks of the piece launch); part 2: levels whose pieces cross tiles
    static FixParams<T> piece_fix_params(fehip_solver *h, Plan *P, const ColOps *ops, T *coef,
                                         const int32_t *early_exit, int part, bool proj = false) {
        const PieceForm &pc = P->pc;
        FixParams<T> f{};
        f.nitems = part == 1 ? (proj ? pc.n1p : pc.n1) : (proj ? pc.n2p : pc.n2);
        f.items = part == 1 ? (proj ? pc.items1p : pc.items1) : (proj ? pc.items2p : pc.items2);
        f.tileA = pc.tileA;
        f.tileB = pc.tileB;
        f.dpart = P->dpart;
        f.v = coef;
        f.ops = ops;
        f.partials = pc.part + (part == 2 ? (size_t)((proj ? pc.fix1p : pc.fix1) + pc.blocks) * K : 0);
        f.all_done = early_exit;
        f.mode = MODE_ITER;
        // raw sums of sqrt(w) .* u take their FE's scale here: the pieces' FEs,
        // and the few-level FEs summed in a factored fused pass (the cache
        // form's few-level sums carry their caches)
        for (int g = 0; g < h->G; g++) {
            const bool raw = g == h->s || g == h->t_sec ||
                             (h->fe[g].kind == KIND_DENSE && P->fused_dense && h->factored);
            if (raw) f.fscale[g] = (const T *)h->fe[g].scale_T[pidx<T>()];
        }
        return f;
    }

    static void piece_dtx(fehip_solver *h, Plan *P, const ColOps *ops, T *coef, const int32_t *early_exit,
                          bool proj = false) {
        const DtxSegParams<T> p = piece_seg_params(h, P, ops, coef, early_exit, proj);
        const FixParams<T> f = piece_fix_params(h, P, ops, coef, early_exit, 1, proj);
        const unsigned nfix = proj ? P->pc.fix1p : P->pc.fix1;
        // the factored fused pass's width (the piece sums are sqrt(w) .* u sums
        // in both forms), whenever the wide lane map is on
        constexpr int WC = fused_wide_c<Acc, K, true>();
        if (WC > 1 && fused_cols(h, sizeof(T), K) > 1)
            hipLaunchKernelGGL((k_dtx_seg_fix<T, Acc, K, true, true, WC>), dim3(nfix + P->pc.blocks), dim3(kBlock),
                               0, h->stream, p, f, (int)nfix);
        else
            hipLaunchKernelGGL((k_dtx_seg_fix<T, Acc, K, true, true>), dim3(nfix + P->pc.blocks), dim3(kBlock), 0,
                               h->stream, p, f, (int)nfix);
    }

    static void piece_fix2(fehip_solver *h, Plan *P, const ColOps *ops, T *coef, const int32_t *early_exit,
                           bool proj = false) {
        const FixParams<T> f = piece_fix_params(h, P, ops, coef, early_exit, 2, proj);
        hipLaunchKernelGGL((k_dtx_fixup<T, K>), dim3(proj ? P->pc.fix2p : P->pc.fix2), dim3(kBlock), 0, h->stream, f);
    }

    static void piece_fin_dtx(fehip_solver *h, Plan *P, bool proj = false) {
        const PieceForm &pc = P->pc;
        const int64_t n = proj ? pc.fix1p + pc.blocks + (pc.n2p ? pc.fix2p : 0)
                               : pc.fix1 + pc.blocks + (pc.n2 ? pc.fix2 : 0);
        hipLaunchKernelGGL((k_fin_dtx<Tr, K>), dim3(1), dim3(kFinBlock), 0, h->stream, (const double *)pc.part, n,
                           (const double *)nullptr, (int64_t)0, MODE_ITER, P->state, P->ops_db, P->all_done,
                           (const int32_t *)P->all_done);
    }

    // one LSMR iteration in the piece form
    static void piece_iteration(fehip_solver *h, Plan *P) {
        fused(h, P, P->ops_db, (const T *)P->v, (T *)P->u, 3);
        fin_db(h, P, MODE_ITER, P->all_done, true);
        if (!P->fused_dense) dtx_dense(h, P, P->ops_dtx, (const T *)P->u, P->all_done, false);
        piece_dtx(h, P, P->ops_dtx, (T *)P->v, P->all_done);
        if (P->pc.n2) piece_fix2(h, P, P->ops_dtx, (T *)P->v, P->all_done);
        piece_fin_dtx(h, P);
        update(h, P);
    }

    static void dtx(fehip_solver *h, Plan *P, int mode, const ColOps *ops, const T *y, T *coef,
                    const int32_t *early_exit) {
        dtx_seg(h, P, mode, ops, y, coef, early_exit);
        dtx_dense(h, P, ops, y, early_exit);
        dtx_fixup(h, P, mode, ops, coef, early_exit);
    }

    // Average device time of each kernel of one iteration over `reps`
    // isolated launches (HIP events on the handle's stream).
    // Average device time of each launch of the LSMR iteration, measured in
    // place: `reps` iterations run eagerly (not from the graph), in their
    // order, with a HIP event between consecutive launches on the handle's
    // stream, so each kernel meets the cache state the real iteration leaves
    // it.  The iterations run a real LSMR start on filler data (b = 1, x = 0)
    // with zero tolerances, so no column stops.
    static int time_kernels(fehip_solver *h, int reps, double *out_ms) {
        Plan *P = nullptr;
        CHK((get_plan_for<T, Acc, K>(h, &P)));
        const int64_t NK = h->N * K, pK = h->p * K;
        hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(NK, kBlock)), dim3(kBlock), 0, h->stream, (T *)P->u, NK, (T)1);
        hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(pK, kBlock)), dim3(kBlock), 0, h->stream, (T *)P->x, pK, (T)0);
        hipLaunchKernelGGL((k_lsmr_setup<Tr>), dim3(1), dim3(64), 0, h->stream, (LsmrState<Tr> *)P->state, K, K, 0.0,
                           0.0, 0.0, (int64_t)1 << 40, P->all_done);
        hipLaunchKernelGGL(k_set_ops, dim3(1), dim3(64), 0, h->stream, P->ops_once, -1.0, 1.0, K, K);
        db(h, P, MODE_INIT, P->ops_once, (const T *)P->x, (T *)P->u, nullptr);
        fin_db(h, P, MODE_INIT, nullptr);
        dtx(h, P, MODE_INIT, P->ops_dtx, (const T *)P->u, (T *)P->v, nullptr);
        fin_dtx(h, P, MODE_INIT, nullptr);
        hipLaunchKernelGGL((k_init_vectors<T, Tr, K>), dim3(P->upd_blocks), dim3(kBlock), 0, h->stream, pK, (T *)P->v,
                           (T *)P->hh, (T *)P->hbar, (const void *)P->state, (const T *)h->scale_all[pidx<T>()],
                           (T *)P->vt);
        if (ut_now(h, P) && h->sqrtw_T[pidx<T>()])
            hipLaunchKernelGGL(k_scale_rows<T>, dim3(grid_for(NK, kBlock)), dim3(kBlock), 0, h->stream, (T *)P->u,
                               (const T *)h->sqrtw_T[pidx<T>()], h->N, K);
        HIPCHK(hipGetLastError());
        hipEvent_t ev[10];
        for (auto &e : ev) HIPCHK(hipEventCreate(&e));
        // the launches of iteration(), one after another on one stream;
        // out_ms slots: fused, fin_db (with the previous iteration's stopping
        // test), CSR D'u, separate few-level D'u, fix-up (both parts),
        // fin_dtx, update, and an empty event pair (the overhead of the
        // timing itself, included in every other slot)
        static const int slot_of_gap[9] = {0, 1, 4, 3, 2, 4, 5, 6, 7};
        auto one = [&](bool rec) {
            auto mark = [&](int q) {
                if (rec) (void)hipEventRecord(ev[q], h->stream);
            };
            if (use_pieces(h, P)) {
                // slots as below: the piece launch counts as the CSR slot,
                // its part 2 as the fix-up
                mark(0);
                fused(h, P, P->ops_db, (const T *)P->v, (T *)P->u, 3);
                mark(1);
                fin_db(h, P, MODE_ITER, P->all_done, true);
                mark(2);
                mark(3);
                if (!P->fused_dense) dtx_dense(h, P, P->ops_dtx, (const T *)P->u, P->all_done, false);
                mark(4);
                piece_dtx(h, P, P->ops_dtx, (T *)P->v, P->all_done);
                mark(5);
                if (P->pc.n2) piece_fix2(h, P, P->ops_dtx, (T *)P->v, P->all_done);
                mark(6);
                piece_fin_dtx(h, P);
                mark(7);
                update(h, P);
                mark(8);
                mark(9);
                return;
            }
            mark(0);
            fused(h, P, P->ops_db, (const T *)P->v, (T *)P->u);
            mark(1);
            fin_db(h, P, MODE_ITER, P->all_done, true);
            const bool sf = use_seg_fix(h, P);
            mark(2);
            if (K >= kSplitFixK && !sf) dtx_fixup(h, P, MODE_ITER, P->ops_dtx, (T *)P->v, P->all_done, true, false, 1);
            mark(3);
            if (!P->fused_dense) dtx_dense(h, P, P->ops_dtx, (const T *)P->u, P->all_done, ut(h));
            mark(4);
            // the merged launch counts as the CSR slot, its part 2 as the fix-up
            if (use_seg_fix(h, P)) seg_fix(h, P, P->ops_dtx, (const T *)P->u, (T *)P->v, P->all_done);
            else dtx_seg(h, P, MODE_ITER, P->ops_dtx, (const T *)P->u, (T *)P->v, P->all_done, true, ut(h));
            mark(5);
            if ((K >= kSplitFixK && !sf) || (sf && P->nitems_c))
                dtx_fixup(h, P, MODE_ITER, P->ops_dtx, (T *)P->v, P->all_done, true, false, 2);
            else if (!sf) dtx_fixup(h, P, MODE_ITER, P->ops_dtx, (T *)P->v, P->all_done, true);
            mark(6);
            fin_dtx(h, P, MODE_ITER, P->all_done, true, sf ? P->nitems_c > 0 : K >= kSplitFixK);
            mark(7);
            update(h, P);
            mark(8);
            mark(9);
        };
        one(false);
        one(false);
        double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int r = 0; r < reps; r++) {
            one(true);
            HIPCHK(hipEventSynchronize(ev[9]));
            for (int q = 0; q < 9; q++) {
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, ev[q], ev[q + 1]));
                acc[slot_of_gap[q]] += ms;
            }
        }
        for (auto &e : ev) (void)hipEventDestroy(e);
        int32_t stopped = 0;
        HIPCHK(hipMemcpy(&stopped, P->all_done, sizeof(int32_t), hipMemcpyDeviceToHost));
        if (stopped) return fail(FEHIP_EHIP, "kernel timing: the filler LSMR stopped early");
        for (int q = 0; q < 8; q++) out_ms[q] = acc[q] / std::max(reps, 1);
        if (!P->csr_blocks && !use_pieces(h, P)) out_ms[2] = 0.0;
        if (P->fused_dense || !P->ndense) out_ms[3] = 0.0;
        return 0;
    }

    // Time the launches of one LSMR iteration under candidate schedules and
    // keep the fastest: tiles of the fused pass and of the CSR launch around
    // the occupancy-sized defaults, and the few-level FEs in or out of the
    // fused pass.  out: tiles_s, tiles_c, fuse_dense, poll_iters, best ms,
    // default ms.
    static int autotune(fehip_solver *h, int reps, double *out) {
        h->sched_tiles_s = h->sched_tiles_c = 0;
        h->sched_fuse_dense = -1;
        destroy_all_plans(h);
        Plan *P = nullptr;
        CHK((get_plan_for<T, Acc, K>(h, &P)));
        const int64_t wf = P->waves_f, wc = P->waves_c;
        const bool dense_choice = P->nslot > 0 && fused_lds_bytes(h, sizeof(T), K) > 0;
        auto measure = [&](int64_t ts, int64_t tc, int fd, double &ms) -> int {
            h->sched_tiles_s = ts;
            h->sched_tiles_c = tc;
            h->sched_fuse_dense = fd;
            destroy_all_plans(h);
            double k[8];
            CHK(time_kernels(h, reps, k));
            ms = 0.0;
            for (double x : k) ms += x;
            return 0;
        };
        double def_ms = 0.0;
        CHK(measure(wf, wc, 1, def_ms));
        double best = def_ms;
        int64_t bs = wf, bc = wc;
        int bd = 1;
        const int64_t cs[] = {wf / 2, wf, 2 * wf};
        const int64_t cc[] = {wc / 2, wc, 2 * wc, 4 * wc};
        for (int fd = 1; fd >= (dense_choice ? 0 : 1); fd--)
            for (int64_t ts : cs)
                for (int64_t tc : cc) {
                    if (ts < 1 || tc < 1 || (ts == wf && tc == wc && fd == 1)) continue;
                    double ms = 0.0;
                    CHK(measure(ts, tc, fd, ms));
                    if (ms < best) {
                        best = ms;
                        bs = ts;
                        bc = tc;
                        bd = fd;
                    }
                }
        h->sched_tiles_s = bs;
        h->sched_tiles_c = bc;
        h->sched_fuse_dense = bd;
        destroy_all_plans(h);
        out[0] = (double)bs;
        out[1] = (double)bc;
        out[2] = (double)bd;
        out[3] = (double)h->poll_iters;
        out[4] = best;
        out[5] = def_ms;
        return 0;
    }

    static void update(fehip_solver *h, Plan *P) {
        UpdParams<T> p{};
        p.pK = h->p * K;
        p.v = (T *)P->v;
        p.h = (T *)P->hh;
        p.hbar = (T *)P->hbar;
        p.x = (T *)P->x;
        p.partials = P->part_upd;
        p.all_done = P->all_done;
        p.state = P->state;
        p.scale = (const T *)h->scale_all[pidx<T>()];
        p.vt = (T *)P->vt;
        if (h->fwl_now) {
            p.skip0 = h->off[h->s] * K;
            p.skipn = h->n[h->s] * K;
        }
        hipLaunchKernelGGL((k_update<T, Tr, K>), dim3(P->upd_blocks), dim3(kBlock), 0, h->stream, p);
    }

    // one LSMR iteration (src/utils/lsmr.jl:103-215) as stream-ordered
    // launches: the fused D*x pass (with D'u of FE s and the dense FEs),
    // beta, D'u of the CSR FEs, the fix-up, alpha and the rotations, the
    // vector update, the stopping rules
    static void iteration(fehip_solver *h, Plan *P) {
        if (h->fwl_now && use_pieces_proj(h, P)) {
            // projected, piece form: the u pass leaves the piece sums; FE t's
            // level sums come from them (FE s's coefficients stay zero)
            fused(h, P, P->ops_one, (const T *)P->v, (T *)P->u, 1);
            fix_s(h, P);
            fused(h, P, P->ops_db, (const T *)P->v, (T *)P->u, 4);
            fin_db(h, P, MODE_ITER, P->all_done, true);
            if (!P->fused_dense) dtx_dense(h, P, P->ops_dtx, (const T *)P->u, P->all_done, false);
            piece_dtx(h, P, P->ops_dtx, (T *)P->v, P->all_done, true);
            if (P->pc.n2p) piece_fix2(h, P, P->ops_dtx, (T *)P->v, P->all_done, true);
            piece_fin_dtx(h, P, true);
            update(h, P);
            return;
        }
        if (h->fwl_now) {
            // projected: A = (I - A_s A_s') A_o, coefficients of FE s zero
            fused(h, P, P->ops_one, (const T *)P->v, (T *)P->u, 1);
            fix_s(h, P);
            fused(h, P, P->ops_db, (const T *)P->v, (T *)P->u, 2);
            fin_db(h, P, MODE_ITER, P->all_done, true);
            if (!P->fused_dense) dtx_dense(h, P, P->ops_dtx, (const T *)P->u, P->all_done, ut(h));
            if (use_seg_fix(h, P)) {
                seg_fix(h, P, P->ops_dtx, (const T *)P->u, (T *)P->v, P->all_done, true);
                if (P->nitems_c) dtx_fixup(h, P, MODE_ITER, P->ops_dtx, (T *)P->v, P->all_done, true, true, 2);
                fin_dtx(h, P, MODE_ITER, P->all_done, true, P->nitems_c > 0);
            } else {
                dtx_seg(h, P, MODE_ITER, P->ops_dtx, (const T *)P->u, (T *)P->v, P->all_done, true, ut(h));
                dtx_fixup(h, P, MODE_ITER, P->ops_dtx, (T *)P->v, P->all_done, true, true);
                fin_dtx(h, P, MODE_ITER, P->all_done, true);
            }
            update(h, P);
            return;
        }
        if (use_pieces(h, P)) {
            piece_iteration(h, P);
            return;
        }
        fused(h, P, P->ops_db, (const T *)P->v, (T *)P->u);
        fin_db(h, P, MODE_ITER, P->all_done, true);
        if (use_seg_fix(h, P)) {
            if (!P->fused_dense) dtx_dense(h, P, P->ops_dtx, (const T *)P->u, P->all_done, ut(h));
            // part 2 only if some CSR segment crosses tiles (or a level is
            // empty, or the few-level FEs ran separately)
            seg_fix(h, P, P->ops_dtx, (const T *)P->u, (T *)P->v, P->all_done);
            if (P->nitems_c) dtx_fixup(h, P, MODE_ITER, P->ops_dtx, (T *)P->v, P->all_done, true, false, 2);
            fin_dtx(h, P, MODE_ITER, P->all_done, true, P->nitems_c > 0);
            update(h, P);
            return;
        }
        if (K < kSplitFixK) {
            if (!P->fused_dense) dtx_dense(h, P, P->ops_dtx, (const T *)P->u, P->all_done, ut(h));
            dtx_seg(h, P, MODE_ITER, P->ops_dtx, (const T *)P->u, (T *)P->v, P->all_done, true, ut(h));
            dtx_fixup(h, P, MODE_ITER, P->ops_dtx, (T *)P->v, P->all_done, true);
            fin_dtx(h, P, MODE_ITER, P->all_done, true);
            update(h, P);
            return;
        }
        // wide batches: the FE s fix-up needs only the fused pass and beta,
        // so it runs as a second branch beside the CSR pass
        (void)hipEventRecord(h->fork_ev, h->stream);
        (void)hipStreamWaitEvent(h->stream2, h->fork_ev, 0);
        dtx_fixup(h, P, MODE_ITER, P->ops_dtx, (T *)P->v, P->all_done, true, false, 1, h->stream2);
        if (!P->fused_dense) dtx_dense(h, P, P->ops_dtx, (const T *)P->u, P->all_done, ut(h));
        dtx_seg(h, P, MODE_ITER, P->ops_dtx, (const T *)P->u, (T *)P->v, P->all_done, true, ut(h));
        dtx_fixup(h, P, MODE_ITER, P->ops_dtx, (T *)P->v, P->all_done, true, false, 2);
        (void)hipEventRecord(h->join_ev, h->stream2);
        (void)hipStreamWaitEvent(h->stream, h->join_ev, 0);
        fin_dtx(h, P, MODE_ITER, P->all_done, true, true);
        update(h, P);
    }

    static int ensure_graph(fehip_solver *h, Plan *P) {
        const int M = h->poll_iters;
        if (P->graph && P->graph_iters == M && P->graph_fwl == h->fwl_now) return 0;
        // the split fix-up's second stream exists only on handles whose graph
        // uses it: every stream of the process shares a few hardware queues,
        // and idle extra streams cost concurrency to many handles (panels)
        if (K >= kSplitFixK && !use_seg_fix(h, P) && !h->fwl_now && !h->stream2) {
            HIPCHK(hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&h->fork_ev, hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&h->join_ev, hipEventDisableTiming));
        }
        if (P->graph) {
            (void)hipGraphExecDestroy(P->graph);
            P->graph = nullptr;
        }
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        for (int it = 0; it < M; it++) iteration(h, P);
        hipError_t e = hipStreamEndCapture(h->stream, &g);
        if (e != hipSuccess) return fail(FEHIP_EHIP, std::string("graph capture failed: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&P->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fail(FEHIP_EHIP, std::string("graph instantiate failed: ") + hipGetErrorString(e));
        P->graph_iters = M;
        P->graph_fwl = h->fwl_now;
        return 0;
    }

    // LSMR on the columns held in P->u (b) with x as the initial iterate
    // (src/utils/lsmr.jl:46-222); per-column results in the state array.
    // lsmr_start queues the set-up and the first two replays of the captured
    // iterations; lsmr_poll advances one replay when its stop flag has come
    // back (blocking or not).  Launches after the stop exit at entry.
    struct Run {
        int64_t launched = 0, maxiter = 0;
        int cur = 0;
    };

    static int launch_replay(fehip_solver *h, Plan *P, Run &run, int slot) {
        HIPCHK(hipGraphLaunch(P->graph, h->stream));
        HIPCHK(hipMemcpyAsync(h->h_flag + 1 + slot, P->all_done, sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipEventRecord(h->poll_ev[slot], h->stream));
        run.launched += P->graph_iters;
        return 0;
    }

    static int lsmr_start(fehip_solver *h, Plan *P, int Kreal, double tol, int64_t maxiter, Run &run) {
        hipLaunchKernelGGL((k_lsmr_setup<Tr>), dim3(1), dim3(64), 0, h->stream, (LsmrState<Tr> *)P->state, K, Kreal,
                           tol, tol, 1e8, maxiter, P->all_done);
        if (h->fwl_now) {
            // projected: x_s = 0 ; u = (I - A_s A_s')(b - A_o x) with b
            // already projected ; v = A_o'u with v_s = 0
            const int64_t so = h->off[h->s] * K, sn = h->n[h->s] * K;
            hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(sn, kBlock)), dim3(kBlock), 0, h->stream, (T *)P->x + so, sn,
                               (T)0);
            hipLaunchKernelGGL(k_set_ops, dim3(1), dim3(64), 0, h->stream, P->ops_once, -1.0, 1.0, K, Kreal);
            db(h, P, MODE_NONE, P->ops_once, (const T *)P->x, (T *)P->u, nullptr);
            proj_s(h, P, (T *)P->u, MODE_INIT);
            fin_db(h, P, MODE_INIT, nullptr);
            hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(sn, kBlock)), dim3(kBlock), 0, h->stream, (T *)P->v + so, sn,
                               (T)0);
            dtx_seg(h, P, MODE_INIT, P->ops_dtx, (const T *)P->u, (T *)P->v, nullptr, true);
            dtx_dense(h, P, P->ops_dtx, (const T *)P->u, nullptr);
            dtx_fixup(h, P, MODE_INIT, P->ops_dtx, (T *)P->v, nullptr, false, true);
            fin_dtx(h, P, MODE_INIT, nullptr, true);
        } else {
            // u = b - A x ; beta = ||u||  (:63-65)
            hipLaunchKernelGGL(k_set_ops, dim3(1), dim3(64), 0, h->stream, P->ops_once, -1.0, 1.0, K, Kreal);
            db(h, P, MODE_INIT, P->ops_once, (const T *)P->x, (T *)P->u, nullptr);
            fin_db(h, P, MODE_INIT, nullptr);
            // v = A'u ; alpha = ||v|| and the scalar set-up  (:66-98)
            dtx(h, P, MODE_INIT, P->ops_dtx, (const T *)P->u, (T *)P->v, nullptr);
            fin_dtx(h, P, MODE_INIT, nullptr);
        }
        hipLaunchKernelGGL((k_init_vectors<T, Tr, K>), dim3(P->upd_blocks), dim3(kBlock), 0, h->stream, h->p * K,
                           (T *)P->v, (T *)P->hh, (T *)P->hbar, (const void *)P->state,
                           (const T *)h->scale_all[pidx<T>()], (T *)P->vt);
        // the factored iteration with K < 8 keeps u as sqrt(w) .* u (not
        // the piece form)
        if (ut_now(h, P) && h->sqrtw_T[pidx<T>()])
            hipLaunchKernelGGL(k_scale_rows<T>, dim3(grid_for(h->N * K, kBlock)), dim3(kBlock), 0, h->stream,
                               (T *)P->u, (const T *)h->sqrtw_T[pidx<T>()], h->N, K);
        HIPCHK(hipGetLastError());
        CHK(ensure_graph(h, P));
        // Two replays in flight: while the host waits on replay i's copy of
        // the stop flag, replay i+1 is already queued, so the GPU never
        // drains.  A stop at set-up (b = 0, A'b = 0 or maxiter = 0) makes
        // both exit at entry.
        run = Run{};
        run.maxiter = maxiter;
        CHK(launch_replay(h, P, run, 0));
        CHK(launch_replay(h, P, run, 1));
        return 0;
    }

    // finished = true once the stop flag is seen (the stream is then idle)
    static int lsmr_poll(fehip_solver *h, Plan *P, Run &run, bool block, bool &finished) {
        finished = false;
        if (block) {
            HIPCHK(hipEventSynchronize(h->poll_ev[run.cur]));
        } else {
            const hipError_t q = hipEventQuery(h->poll_ev[run.cur]);
            if (q == hipErrorNotReady) return 0;
            if (q != hipSuccess) return fail(FEHIP_EHIP, std::string("event query: ") + hipGetErrorString(q));
        }
        if (h->h_flag[1 + run.cur]) {
            HIPCHK(hipStreamSynchronize(h->stream));
            finished = true;
            return 0;
        }
        if (run.launched > run.maxiter + 3 * P->graph_iters)
            return fail(FEHIP_EHIP, "LSMR did not reach its stopping rule within maxiter");
        CHK(launch_replay(h, P, run, run.cur));
        run.cur ^= 1;
        return 0;
    }

    static int lsmr(fehip_solver *h, Plan *P, int Kreal, double tol, int64_t maxiter) {
        Run run;
        CHK(lsmr_start(h, P, Kreal, tol, maxiter, run));
        for (bool fin = false; !fin;) CHK(lsmr_poll(h, P, run, true, fin));
        return 0;
    }

    // solve_residuals! of one panel as a resumable job (for the concurrent
    // driver of many panels): begin queues the load and the LSMR, step
    // advances it without blocking, end stores the residuals
    struct Job {
        fehip_solver *h = nullptr;
        Plan *P = nullptr;
        Run run;
        double *dY = nullptr;
        int Kreal = 0;
        int64_t ldY = 0;
        int64_t *iters = nullptr;
        uint8_t *conv = nullptr;
    };
    static int job_begin(Job &j, double tol, int64_t maxiter) {
        FwlScope fwl(j.h);
        CHK((get_plan_for<T, Acc, K>(j.h, &j.P)));
        CHK(start(j.h, j.P, j.dY, j.Kreal, j.ldY, true));
        if (j.h->G > 1) CHK(lsmr_start(j.h, j.P, j.Kreal, tol, maxiter - 1, j.run));
        return 0;
    }
    static int job_step(Job &j, bool &finished) {
        if (j.h->G == 1) {
            finished = true;
            return 0;
        }
        return lsmr_poll(j.h, j.P, j.run, false, finished);
    }
    static int job_end(Job &j) {
        FwlScope fwl(j.h);
        std::vector<int64_t> it(j.Kreal, 1);
        std::vector<int> cv(j.Kreal, 1);
        if (j.h->G > 1) {
            CHK(read_state(j.h, j.P, j.Kreal, it, cv));
            for (int c = 0; c < j.Kreal; c++) it[c] += 1;
        }
        CHK(finish(j.h, j.P, j.dY, j.Kreal, j.ldY));
        for (int c = 0; c < j.Kreal; c++) {
            if (j.iters) j.iters[c] = it[c];
            if (j.conv) j.conv[c] = (uint8_t)cv[c];
        }
        return 0;
    }

    static int read_state(fehip_solver *h, Plan *P, int Kreal, std::vector<int64_t> &iters,
                          std::vector<int> &conv) {
        std::vector<LsmrState<Tr>> st(K);
        HIPCHK(hipMemcpyAsync(st.data(), P->state, sizeof(LsmrState<Tr>) * K, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        h->last_info.assign((size_t)Kreal * 9, 0.0);
        for (int c = 0; c < Kreal; c++) {
            double *o = h->last_info.data() + (size_t)c * 9;
            o[0] = (double)st[c].iter;
            o[1] = (double)st[c].istop;
            o[2] = (double)st[c].normr;
            o[3] = (double)st[c].normAr;
            o[4] = (double)st[c].normA;
            o[5] = (double)st[c].condA;
            o[6] = (double)st[c].normx;
            o[7] = (double)st[c].alpha;
            o[8] = (double)st[c].beta;
        }
        iters.resize(Kreal);
        conv.resize(Kreal);
        for (int c = 0; c < Kreal; c++) {
            iters[c] = st[c].iter;
            const int is = st[c].istop;
            conv[c] = !(is == 3 || is == 6 || is == 7);  // src/utils/lsmr.jl:218
        }
        return 0;
    }

    // r = sqrt(w) .* y ; b = copy(r) (:51-55) and x = A'b (:56)
    static int start(fehip_solver *h, Plan *P, const double *dY, int Kreal, int64_t ldY, bool x_from_At) {
        const int64_t NK = h->N * K;
        if constexpr (K >= 8)
            hipLaunchKernelGGL(k_load_cols_t<T>, dim3(grid_for(h->N, kTileRows)), dim3(kBlock), 0, h->stream, dY, ldY,
                               K, Kreal, h->pos_s, h->sqrtw_int, h->N, (T *)P->r, (T *)P->u);
        else
            hipLaunchKernelGGL(k_load_cols<T>, dim3(grid_for(NK, kBlock)), dim3(kBlock), 0, h->stream, dY, ldY, K,
                               Kreal, h->perm_s, h->sqrtw_int, h->N, (T *)P->r, (T *)P->u);
        if (h->fwl_now) {
            // projected problem: b <- (I - A_s A_s') b
            CHK(fwl_setup(h, P, Kreal));
            proj_s(h, P, (T *)P->r, MODE_NONE);
            hipLaunchKernelGGL(k_copy<T>, dim3(grid_for(NK, kBlock)), dim3(kBlock), 0, h->stream, (const T *)P->r,
                               (T *)P->u, NK);
        }
        if (x_from_At) {
            hipLaunchKernelGGL(k_set_ops, dim3(1), dim3(64), 0, h->stream, P->ops_once, 1.0, 0.0, K, Kreal);
            dtx(h, P, MODE_NONE, P->ops_once, (const T *)P->r, (T *)P->x, nullptr);
        }
        HIPCHK(hipGetLastError());
        return 0;
    }

    // the projected iteration's constant operators and scale_s^2
    static int fwl_setup(fehip_solver *h, Plan *P, int Kreal) {
        hipLaunchKernelGGL(k_set_ops, dim3(1), dim3(64), 0, h->stream, P->ops_one, 1.0, 0.0, K, Kreal);
        hipLaunchKernelGGL(k_set_ops, dim3(1), dim3(64), 0, h->stream, P->ops_neg, -1.0, 0.0, K, Kreal);
        hipLaunchKernelGGL(k_set_ops, dim3(1), dim3(64), 0, h->stream, P->ops_sub, -1.0, 1.0, K, Kreal);
        if (h->factored) {
            constexpr int q = pidx<T>();
            const int64_t ns = h->n[h->s];
            if (!h->scale2_T[q]) CHK(dalloc_bytes(h, &h->scale2_T[q], sizeof(T) * ns));
            hipLaunchKernelGGL(k_square<T>, dim3(grid_for(ns, kBlock)), dim3(kBlock), 0, h->stream,
                               (const T *)h->fe[h->s].scale_T[q], ns, (T *)h->scale2_T[q]);
        }
        HIPCHK(hipGetLastError());
        return 0;
    }

    // r = r - A x ; r ./= sqrt(w) ; back to the caller's columns (:62-66)
    static int finish(fehip_solver *h, Plan *P, double *dY, int Kreal, int64_t ldY) {
        const int64_t NK = h->N * K;
        hipLaunchKernelGGL(k_set_ops, dim3(1), dim3(64), 0, h->stream, P->ops_once, -1.0, 1.0, K, Kreal);
        db(h, P, MODE_NONE, P->ops_once, (const T *)P->x, (T *)P->r, nullptr);
        if (h->fwl_now) proj_s(h, P, (T *)P->r, MODE_NONE);  // r = b - A x with A projected
        if constexpr (K >= 8)
            hipLaunchKernelGGL(k_store_cols_t<T>, dim3(grid_for(h->N, kTileRows)), dim3(kBlock), 0, h->stream, dY, ldY,
                               K, Kreal, h->pos_s, h->sqrtw_int, h->N, (const T *)P->r);
        else
            hipLaunchKernelGGL(k_store_cols<T>, dim3(grid_for(NK, kBlock)), dim3(kBlock), 0, h->stream, dY, ldY, K,
                               Kreal, h->perm_s, h->sqrtw_int, h->N, (const T *)P->r);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(h->stream));
        return 0;
    }

    // inner solve_residuals! (src/AbstractFixedEffectSolver.jl:49-68) on Kreal
    // device columns
    static int solve_residuals(fehip_solver *h, double *dY, int Kreal, int64_t ldY, double tol, int64_t maxiter,
                               int64_t *iters, uint8_t *converged) {
        Plan *P = nullptr;
        CHK((get_plan_for<T, Acc, K>(h, &P)));
        CHK(start(h, P, dY, Kreal, ldY, true));
        std::vector<int64_t> it(Kreal, 0);
        std::vector<int> cv(Kreal, 1);
        if (h->G > 1) {  // :58-61
            CHK(lsmr(h, P, Kreal, tol, maxiter - 1));
            CHK(read_state(h, P, Kreal, it, cv));
            for (int c = 0; c < Kreal; c++) it[c] += 1;
        } else {
            for (int c = 0; c < Kreal; c++) it[c] = 1;
        }
        CHK(finish(h, P, dY, Kreal, ldY));
        for (int c = 0; c < Kreal; c++) {
            if (iters) iters[c] = it[c];
            if (converged) converged[c] = (uint8_t)cv[c];
        }
        return 0;
    }

    static WsParams ws_params(fehip_solver *h) {
        WsParams w{};
        w.G = h->G;
        for (int g = 0; g < h->G; g++) {
            w.fe[g].scale = h->fe[g].scale_d;
            w.fe[g].off = h->fe[g].ooff;
            w.fe[g].off_p = h->off[g];
            w.fe[g].n = (int32_t)h->n[g];
            w.fe[g].interacted = h->fe[g].z_orig != nullptr;
        }
        w.degree = h->degree;
        return w;
    }

    static int ensure_ws(fehip_solver *h) {
        if (!h->degree) {
            CHK(dalloc(h, &h->degree, h->p));
            for (int g = 0; g < h->G; g++) {
                const FeInfo &f = h->fe[g];
                int32_t *d = h->degree + h->off[g];
                if (g == h->s)
                    hipLaunchKernelGGL(k_degree_sorted, dim3(grid_for(f.n, 64)), dim3(64), 0, h->stream, f.ioff,
                                       h->t_sec >= 0 ? (const int32_t *)h->fe[h->t_sec].ref_int : nullptr, f.n, d);
                else if (f.kind == KIND_CSR)
                    hipLaunchKernelGGL(k_degree_csr, dim3(grid_for(f.n, 64)), dim3(64), 0, h->stream, f.ioff, f.iperm,
                                       (const int32_t *)h->fe[h->s].ref_int, f.n, d);
                else
                    hipLaunchKernelGGL(k_degree_count, dim3(grid_for(f.n, 256)), dim3(256), 0, h->stream, f.ooff,
                                       f.n, d);
                HIPCHK(hipGetLastError());
            }
        }
        if (!h->ws_feats) CHK(dalloc(h, &h->ws_feats, h->p * 16));
        if (!h->ws_pred) CHK(dalloc(h, &h->ws_pred, h->p));
        if (!h->ws_sig) CHK(dalloc(h, &h->ws_sig, 1));
        if (!h->ws_coef) CHK(dalloc(h, &h->ws_coef, h->p));
        return 0;
    }

    // Features of one caller column dy (device, N doubles) on the K = 1 plan:
    // x = A'b (the reference start), one alternating-projection pass
    // (x_g = A_g' r, r -= A_g x_g for g = 1..G, left in P->hh), sigma = rms(b)
    static int features(fehip_solver *h, const double *dy, float *feats, double *coef_ap) {
        if (K != 1) return fail(FEHIP_EARG, "features run on single columns");
        Plan *P = nullptr;
        CHK((get_plan_for<T, Acc, 1>(h, &P)));
        CHK(ensure_ws(h));
        CHK(start(h, P, dy, 1, h->N, true));
        hipLaunchKernelGGL((k_col_rms<T>), dim3(1), dim3(kFinBlock), 0, h->stream, (const T *)P->r, h->N, h->ws_sig);
        T *xap = (T *)P->hh, *mask = (T *)P->hbar, *tmp = (T *)P->v, *rw = (T *)P->u;
        const unsigned gp = grid_for(h->p, kBlock);
        hipLaunchKernelGGL(k_fill<T>, dim3(gp), dim3(kBlock), 0, h->stream, xap, h->p, (T)0);
        for (int g = 0; g < h->G; g++) {
            hipLaunchKernelGGL(k_set_ops, dim3(1), dim3(64), 0, h->stream, P->ops_once, 1.0, 0.0, 1, 1);
            dtx(h, P, MODE_NONE, P->ops_once, rw, tmp, nullptr);
            hipLaunchKernelGGL(k_copy_range<T>, dim3(grid_for(h->n[g], kBlock)), dim3(kBlock), 0, h->stream,
                               tmp + h->off[g], xap + h->off[g], h->n[g]);
            hipLaunchKernelGGL(k_fill<T>, dim3(gp), dim3(kBlock), 0, h->stream, mask, h->p, (T)0);
            hipLaunchKernelGGL(k_copy_range<T>, dim3(grid_for(h->n[g], kBlock)), dim3(kBlock), 0, h->stream,
                               xap + h->off[g], mask + h->off[g], h->n[g]);
            hipLaunchKernelGGL(k_set_ops, dim3(1), dim3(64), 0, h->stream, P->ops_once, -1.0, 1.0, 1, 1);
            db(h, P, MODE_NONE, P->ops_once, mask, rw, nullptr);
        }
        hipLaunchKernelGGL((k_ws_features<T>), dim3(gp), dim3(kBlock), 0, h->stream, ws_params(h), h->p,
                           (const T *)P->x, (const T *)xap, (const double *)h->ws_sig, feats, coef_ap);
        HIPCHK(hipGetLastError());
        return 0;
    }

    // solve_residuals! from given initial coefficients (unscaled, p per
    // column; a NULL column keeps the reference start A'b), or from the
    // network's start when net is given
    static int solve_residuals_from(fehip_solver *h, double *dY, int Kreal, int64_t ldY, double tol, int64_t maxiter,
                                    const double *const *coef0, fehip_mlp *net, int64_t *iters, uint8_t *converged) {
        FwlScope fwl(h);
        Plan *P = nullptr;
        CHK((get_plan_for<T, Acc, K>(h, &P)));
        CHK(ensure_ws(h));
        const unsigned gp = grid_for(h->p, kBlock);
        // the network's starts first: the feature pass runs on the K = 1
        // buffers, which are this solve's own when K = 1
        if (net) {
            if (h->ws_x0_cols < Kreal) {
                dfree(h, h->ws_x0);
                h->ws_x0 = nullptr;
                CHK(dalloc(h, &h->ws_x0, (size_t)h->p * Kreal));
                h->ws_x0_cols = Kreal;
            }
            for (int c = 0; c < Kreal; c++) {
                CHK((Engine<T, Acc, Tr, 1>::features(h, dY + (int64_t)c * ldY, h->ws_feats, nullptr)));
                HIPCHK(hipStreamSynchronize(h->stream));
                CHK(mlp_predict_device(net, h->ws_feats, h->p, h->ws_pred));
                Plan *P1 = nullptr;
                CHK((get_plan_for<T, Acc, 1>(h, &P1)));
                hipLaunchKernelGGL((k_x0_from_pred<T>), dim3(gp), dim3(kBlock), 0, h->stream, (const T *)P1->hh,
                                   (const float *)h->ws_pred, (const double *)h->ws_sig, ws_params(h), h->p,
                                   h->ws_x0 + (size_t)c * h->p);
                HIPCHK(hipGetLastError());
            }
        }
        CHK(start(h, P, dY, Kreal, ldY, true));
        for (int c = 0; c < Kreal; c++) {
            if (net) {
                hipLaunchKernelGGL((k_put_col<T>), dim3(gp), dim3(kBlock), 0, h->stream,
                                   (const double *)h->ws_x0 + (size_t)c * h->p, h->p, (T *)P->x, K, c);
            } else if (coef0 && coef0[c]) {
                HIPCHK(hipMemcpyAsync(h->ws_coef, coef0[c], sizeof(double) * h->p, hipMemcpyHostToDevice, h->stream));
                CHK(s_block_to_internal(h, h->ws_coef));
                hipLaunchKernelGGL((k_x0_from_coef<T>), dim3(gp), dim3(kBlock), 0, h->stream,
                                   (const double *)h->ws_coef, ws_params(h), h->p, (T *)P->x, K, c);
            }
            HIPCHK(hipGetLastError());
        }
        std::vector<int64_t> it(Kreal, 1);
        std::vector<int> cv(Kreal, 1);
        CHK(lsmr(h, P, Kreal, tol, maxiter - 1));
        CHK(read_state(h, P, Kreal, it, cv));
        for (int c = 0; c < Kreal; c++) it[c] += 1;
        CHK(finish(h, P, dY, Kreal, ldY));
        for (int c = 0; c < Kreal; c++) {
            if (iters) iters[c] = it[c];
            if (converged) converged[c] = (uint8_t)cv[c];
        }
        return 0;
    }

    // inner solve_coefficients! up to the collect (:147-158)
    static int solve_coefficients(fehip_solver *h, double *dy, double tol, int64_t maxiter, double *dcoef,
                                  int64_t *iters, uint8_t *converged, const double *coef0 = nullptr) {
        Plan *P = nullptr;
        CHK((get_plan_for<T, Acc, K>(h, &P)));
        const int64_t NK = h->N * K;
        hipLaunchKernelGGL(k_load_cols<T>, dim3(grid_for(NK, kBlock)), dim3(kBlock), 0, h->stream, dy, h->N, K, 1,
                           h->perm_s, h->sqrtw_int, h->N, (T *)P->u, (T *)nullptr);
        hipLaunchKernelGGL(k_fill<T>, dim3(grid_for(h->p * K, kBlock)), dim3(kBlock), 0, h->stream, (T *)P->x,
                           h->p * K, (T)0);
        if (coef0) {
            CHK(ensure_ws(h));
            HIPCHK(hipMemcpyAsync(h->ws_coef, coef0, sizeof(double) * h->p, hipMemcpyHostToDevice, h->stream));
            CHK(s_block_to_internal(h, h->ws_coef));
            hipLaunchKernelGGL((k_x0_from_coef<T>), dim3(grid_for(h->p, kBlock)), dim3(kBlock), 0, h->stream,
                               (const double *)h->ws_coef, ws_params(h), h->p, (T *)P->x, K, 0);
        }
        if (h->fwl_now) {
            // projected: b kept in r; LSMR on (I - A_s A_s') b
            hipLaunchKernelGGL(k_copy<T>, dim3(grid_for(NK, kBlock)), dim3(kBlock), 0, h->stream, (const T *)P->u,
                               (T *)P->r, NK);
            CHK(fwl_setup(h, P, 1));
            proj_s(h, P, (T *)P->u, MODE_NONE);
        }
        CHK(lsmr(h, P, 1, tol, maxiter));
        std::vector<int64_t> it;
        std::vector<int> cv;
        CHK(read_state(h, P, 1, it, cv));
        if (h->fwl_now) {
            // the eliminated FE's coefficients: the least-squares fit of the
            // rest, x_s = A_s'(b - A_o x_o) (A_s has orthonormal columns)
            hipLaunchKernelGGL(k_set_ops, dim3(1), dim3(64), 0, h->stream, P->ops_once, -1.0, 1.0, K, 1);
            db(h, P, MODE_NONE, P->ops_once, (const T *)P->x, (T *)P->r, nullptr);
            s_sums(h, P, (const T *)P->r);
            hipLaunchKernelGGL(k_copy<T>, dim3(grid_for(h->n[h->s] * K, kBlock)), dim3(kBlock), 0, h->stream,
                               (const T *)P->zs, (T *)P->x + h->off[h->s] * K, h->n[h->s] * K);
        }
        for (int g = 0; g < h->G; g++)
            hipLaunchKernelGGL(k_scale_coef<T>, dim3(grid_for(h->n[g], kBlock)), dim3(kBlock), 0, h->stream,
                               (const T *)P->x + h->off[g], (const T *)h->fe[g].scale_T[pidx<T>()], h->n[g],
                               dcoef + h->off[g]);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(h->stream));
        if (iters) *iters = it[0] / 2;  // :159
        if (converged) *converged = (uint8_t)cv[0];
        return 0;
    }

    // test hooks: one mul! on K = 1 (caller's order in and out)
    static int mul_A(fehip_solver *h, double *dy, const double *dx, double a, double b) {
        Plan *P = nullptr;
        CHK((get_plan_for<T, Acc, K>(h, &P)));
        hipLaunchKernelGGL(k_load_cols<T>, dim3(grid_for(h->N, kBlock)), dim3(kBlock), 0, h->stream, dy, h->N, 1, 1,
                           h->perm_s, (const double *)nullptr, h->N, (T *)P->r, (T *)nullptr);
        hipLaunchKernelGGL(k_to_T<T>, dim3(grid_for(h->p, kBlock)), dim3(kBlock), 0, h->stream, dx, (T *)P->x, h->p);
        hipLaunchKernelGGL(k_set_ops, dim3(1), dim3(64), 0, h->stream, P->ops_once, a, b, K, 1);
        db(h, P, MODE_NONE, P->ops_once, (const T *)P->x, (T *)P->r, nullptr);
        hipLaunchKernelGGL(k_store_cols<T>, dim3(grid_for(h->N, kBlock)), dim3(kBlock), 0, h->stream, dy, h->N, 1, 1,
                           h->perm_s, (const double *)nullptr, h->N, (const T *)P->r);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(h->stream));
        return 0;
    }

    static int mul_At(fehip_solver *h, double *dx, const double *dy, double a, double b) {
        Plan *P = nullptr;
        CHK((get_plan_for<T, Acc, K>(h, &P)));
        hipLaunchKernelGGL(k_load_cols<T>, dim3(grid_for(h->N, kBlock)), dim3(kBlock), 0, h->stream, dy, h->N, 1, 1,
                           h->perm_s, (const double *)nullptr, h->N, (T *)P->r, (T *)nullptr);
        hipLaunchKernelGGL(k_to_T<T>, dim3(grid_for(h->p, kBlock)), dim3(kBlock), 0, h->stream, (const double *)dx,
                           (T *)P->x, h->p);
        hipLaunchKernelGGL(k_set_ops, dim3(1), dim3(64), 0, h->stream, P->ops_once, a, b, K, 1);
        dtx(h, P, MODE_NONE, P->ops_once, (const T *)P->r, (T *)P->x, nullptr);
        hipLaunchKernelGGL(k_to_double<T>, dim3(grid_for(h->p, kBlock)), dim3(kBlock), 0, h->stream, (const T *)P->x,
                           dx, h->p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(h->stream));
        return 0;
    }
};

template <typename T, typename Acc, typename Tr, typename F>
static int for_K(int K, F &&f) {
    switch (K) {
        case 1: return f(Engine<T, Acc, Tr, 1>{});
        case 2: return f(Engine<T, Acc, Tr, 2>{});
        case 4: return f(Engine<T, Acc, Tr, 4>{});
        case 8: return f(Engine<T, Acc, Tr, 8>{});
        case 16: return f(Engine<T, Acc, Tr, 16>{});
        case 32: return f(Engine<T, Acc, Tr, 32>{});
        default: return fail(FEHIP_EARG, "unsupported batch width");
    }
}

// runtime dispatch over (precision, K): f receives an Engine instance
template <typename F>
static int dispatch(int precision, int K, F &&f) {
    if (precision == FEHIP_FP64) return for_K<double, double, double>(K, f);
    if (precision == FEHIP_FP32) return for_K<float, float, float>(K, f);
    return for_K<float, F32x2, double>(K, f);
}

// FEHIP_MIXED: LSMR with FP32 streams (compensated segment sums, FP64
// scalars) to a loose tolerance, then FP64 LSMR iterations started from that
// solution (x handed over in FP64) under the caller's stopping rule; the
// residual is formed in FP64.  The FP32 phase alone cannot reach 1e-8: the
// caches rounded to FP32 perturb the operator and bias the least-squares
// fixed point by about eps32 * cond(A).
template <int K>
struct Mixed {
    using E32 = Engine<float, F32x2, double, K>;
    using E64 = Engine<double, double, double, K>;

    // phase 1 from x0 (A'b for residuals, 0 for coefficients); returns the
    // largest phase-1 iteration count
    static int phase1(fehip_solver *h, Plan *P32, Plan *P64, const double *dY, int Kreal, int64_t ldY, bool x_from_At,
                      double tol, int64_t maxiter, std::vector<int64_t> &it1) {
        CHK(E32::start(h, P32, dY, Kreal, ldY, x_from_At));
        if (!x_from_At)
            hipLaunchKernelGGL(k_fill<float>, dim3(grid_for(h->p * K, kBlock)), dim3(kBlock), 0, h->stream,
                               (float *)P32->x, h->p * K, 0.f);
        CHK(E32::lsmr(h, P32, Kreal, std::max(tol, h->mixed_tol), maxiter));
        std::vector<int> cv;
        CHK(E32::read_state(h, P32, Kreal, it1, cv));
        hipLaunchKernelGGL(k_to_double<float>, dim3(grid_for(h->p * K, kBlock)), dim3(kBlock), 0, h->stream,
                           (const float *)P32->x, (double *)P64->x, h->p * K);
        HIPCHK(hipGetLastError());
        return 0;
    }

    static int solve_residuals(fehip_solver *h, double *dY, int Kreal, int64_t ldY, double tol, int64_t maxiter,
                               int64_t *iters, uint8_t *converged) {
        if (h->G == 1) return E64::solve_residuals(h, dY, Kreal, ldY, tol, maxiter, iters, converged);
        Plan *P32 = nullptr, *P64 = nullptr;
        CHK((get_plan_for<float, F32x2, K>(h, &P32)));
        CHK((get_plan_for<double, double, K>(h, &P64)));
        std::vector<int64_t> it1, it2;
        std::vector<int> cv;
        CHK(phase1(h, P32, P64, dY, Kreal, ldY, true, tol, maxiter - 1, it1));
        const int64_t used = *std::max_element(it1.begin(), it1.end());
        CHK(E64::start(h, P64, dY, Kreal, ldY, false));
        CHK(E64::lsmr(h, P64, Kreal, tol, std::max<int64_t>(1, maxiter - 1 - used)));
        CHK(E64::read_state(h, P64, Kreal, it2, cv));
        CHK(E64::finish(h, P64, dY, Kreal, ldY));
        for (int c = 0; c < Kreal; c++) {
            h->last_info[(size_t)c * 9] = (double)(it1[c] + it2[c]);
            if (iters) iters[c] = it1[c] + it2[c] + 1;
            if (converged) converged[c] = (uint8_t)cv[c];
        }
        return 0;
    }

    static int solve_coefficients(fehip_solver *h, double *dy, double tol, int64_t maxiter, double *dcoef,
                                  int64_t *iters, uint8_t *converged) {
        Plan *P32 = nullptr, *P64 = nullptr;
        CHK((get_plan_for<float, F32x2, K>(h, &P32)));
        CHK((get_plan_for<double, double, K>(h, &P64)));
        std::vector<int64_t> it1, it2;
        std::vector<int> cv;
        CHK(phase1(h, P32, P64, dy, 1, h->N, false, tol, maxiter, it1));
        CHK(E64::start(h, P64, dy, 1, h->N, false));
        CHK(E64::lsmr(h, P64, 1, tol, std::max<int64_t>(1, maxiter - it1[0])));
        CHK(E64::read_state(h, P64, 1, it2, cv));
        for (int g = 0; g < h->G; g++)
            hipLaunchKernelGGL(k_scale_coef<double>, dim3(grid_for(h->n[g], kBlock)), dim3(kBlock), 0, h->stream,
                               (const double *)P64->x + h->off[g], (const double *)h->fe[g].scale_T[0], h->n[g],
                               dcoef + h->off[g]);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(h->stream));
        const int64_t tot = it1[0] + it2[0];
        h->last_info[0] = (double)tot;
        if (iters) *iters = tot / 2;
        if (converged) *converged = (uint8_t)cv[0];
        return 0;
    }
};

template <typename F>
static int for_K_mixed(int K, F &&f) {
    switch (K) {
        case 1: return f(Mixed<1>{});
        case 2: return f(Mixed<2>{});
        case 4: return f(Mixed<4>{});
        case 8: return f(Mixed<8>{});
        case 16: return f(Mixed<16>{});
        case 32: return f(Mixed<32>{});
        default: return fail(FEHIP_EARG, "unsupported batch width");
    }
}

// widest batch (a power of two <= 32) whose coefficient vectors p*K*8 stay
// below 2^32 bytes (the fused pass's 32-bit offsets); 0 if not even one
static int max_batch(const fehip_solver *h) {
    int k = kMaxK;
    while (k > 0 && (double)h->p * k * sizeof(double) >= 4294967296.0) k >>= 1;
    return k;
}

static int padded_K(int K) {
    int k = 1;
    while (k < K) k <<= 1;
    return k;
}

static int solve_residuals_dev(fehip_solver *h, double *dY, int Kreal, int64_t ldY, double tol, int64_t maxiter,
                               int64_t *iters, uint8_t *converged) {
    FwlScope fwl(h);
    if (h->precision == FEHIP_MIXED)
        return for_K_mixed(padded_K(Kreal), [&](auto e) {
            return decltype(e)::solve_residuals(h, dY, Kreal, ldY, tol, maxiter, iters, converged);
        });
    return dispatch(h->precision, padded_K(Kreal), [&](auto e) {
        return decltype(e)::solve_residuals(h, dY, Kreal, ldY, tol, maxiter, iters, converged);
    });
}

// LSMR from zero, coefficients unscaled into dc (p doubles, device).
static int solve_coefficients_dev(fehip_solver *h, double *dy, double tol, int64_t maxiter, double *dc,
                                  int64_t *iters, uint8_t *converged) {
    FwlScope fwl(h);
    if (h->precision == FEHIP_MIXED) h->fwl_now = false;  // its two phases keep the reference LSMR
    if (h->precision == FEHIP_MIXED)
        return for_K_mixed(1, [&](auto e) {
            return decltype(e)::solve_coefficients(h, dy, tol, maxiter, dc, iters, converged);
        });
    return dispatch(h->precision, 1, [&](auto e) {
        return decltype(e)::solve_coefficients(h, dy, tol, maxiter, dc, iters, converged);
    });
}

// ---------------------------------------------------------------------------
// normalize! / full on the device (src/FixedEffect.jl:106-140)
// ---------------------------------------------------------------------------

// Connected components of the FEs without interaction (normalize! only
// rescales those, :107-109), built once per handle: they depend on the refs
// alone.  Leaves cc_nodes (levels sorted by component, unobserved last),
// cc_start (component offsets into cc_nodes) and cc_ncomp (count, device).
static int ensure_components(fehip_solver *h) {
    if (h->cc_ready) return 0;
    std::vector<int> idx;
    for (int g = 0; g < h->G; g++)
        if (!h->fe[g].z_orig) idx.push_back(g);
    if (idx.size() < 2) {
        h->cc_nu = (int)idx.size();
        h->cc_ready = true;
        return 0;
    }
    const int64_t p = h->p;
    if (p >= (int64_t)INT32_MAX - 2) return fail(FEHIP_EARG, "too many levels for the component labels");
    int32_t *parent = nullptr, *key = nullptr, *skeys = nullptr;
    uint8_t *seen = nullptr;
    uint32_t *flag = nullptr, *rid = nullptr, *sums = nullptr;
    CHK(dalloc(h, &parent, p));
    CHK(dalloc(h, &seen, p));
    CHK(dalloc(h, &flag, p));
    CHK(dalloc(h, &rid, p + 1));
    CHK(dalloc(h, &sums, (p + kSortTile - 1) / kSortTile + 2));
    CHK(dalloc(h, &key, p));
    CHK(dalloc(h, &skeys, p));
    CHK(dalloc(h, &h->cc_nodes, p));
    CHK(dalloc(h, &h->cc_start, p + 2));
    CHK(dalloc(h, &h->cc_ncomp, 1));
    hipLaunchKernelGGL(k_cc_init, dim3(grid_for(p, kBlock)), dim3(kBlock), 0, h->stream, parent, seen, p);
    CcParams prm{};
    for (size_t j = 0; j < idx.size(); j++) {
        prm.refs[j] = h->fe[idx[j]].ref_int;
        prm.node0[j] = h->off[idx[j]];
    }
    prm.nu = (int32_t)idx.size();
    hipLaunchKernelGGL(k_cc_hook, dim3(grid_for(h->N, kBlock)), dim3(kBlock), 0, h->stream, prm, h->N, parent,
                       seen);
    hipLaunchKernelGGL(k_cc_roots, dim3(grid_for(p, kBlock)), dim3(kBlock), 0, h->stream, parent,
                       (const uint8_t *)seen, p, flag);
    HIPCHK(hipGetLastError());
    CHK(exclusive_scan_u32(h, flag, p, rid, true, sums));
    hipLaunchKernelGGL(k_cc_label, dim3(grid_for(p, kBlock)), dim3(kBlock), 0, h->stream, parent,
                       (const uint8_t *)seen, (const uint32_t *)rid, p, key);
    HIPCHK(hipGetLastError());
    CHK(radix_sort(h, key, p, p + 1, skeys, h->cc_nodes));
    hipLaunchKernelGGL(k_cc_starts, dim3(grid_for(p + 1, kBlock)), dim3(kBlock), 0, h->stream,
                       (const int32_t *)skeys, p, h->cc_start);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(h->cc_ncomp, rid + p, sizeof(uint32_t), hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    dfree(h, parent);
    dfree(h, seen);
    dfree(h, flag);
    dfree(h, rid);
    dfree(h, sums);
    dfree(h, key);
    dfree(h, skeys);
    h->cc_nu = (int)idx.size();
    h->cc_role.assign(h->G, -1);
    for (size_t j = 0; j < idx.size(); j++) h->cc_role[idx[j]] = (int)j;
    h->cc_ready = true;
    return 0;
}

// normalize! (:106-111) of the concatenated coefficients dc (p doubles, device).
static int normalize_dev(fehip_solver *h, double *dc) {
    CHK(ensure_components(h));
    if (h->cc_nu < 2) return 0;
    RescaleParams prm{};
    for (int g = 0; g <= h->G; g++) prm.off[g] = h->off[g];
    for (int g = 0; g < h->G; g++) prm.role[g] = h->cc_role[g];
    prm.G = h->G;
    prm.nu = h->cc_nu;
    const unsigned grid = (unsigned)std::min<int64_t>(h->p, 1024);
    hipLaunchKernelGGL(k_rescale, dim3(grid), dim3(kBlock), 0, h->stream, prm, (const int32_t *)h->cc_nodes,
                       (const int32_t *)h->cc_start, (const uint32_t *)h->cc_ncomp, dc);
    HIPCHK(hipGetLastError());
    return 0;
}

// full (:137-140) of FE g into out (N doubles, device, caller's order).
static int full_dev(fehip_solver *h, const double *dc, int g, double *out) {
    hipLaunchKernelGGL(k_full, dim3(grid_for(h->N, kBlock)), dim3(kBlock), 0, h->stream, dc + h->off[g],
                       (const int32_t *)h->fe[g].ref_int, (const int32_t *)h->perm_s, h->N, out);
    HIPCHK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// grouping of raw keys (src/FixedEffect.jl:18-21)
// ---------------------------------------------------------------------------

// Number the distinct values of u (n unsigned keys < 2^bits) densely in
// ascending order into dense (n int64); returns the count in *ngroups.
// kb/vb: ping-pong key/value buffers of n entries; hist/sums/flag/ex: scratch.
// Device buffers of one grouping call of N observations (group_keys).
struct GroupScratch {
    uint64_t *kb[2] = {nullptr, nullptr};  // N each: radix sort key buffers
    int32_t *vb[2] = {nullptr, nullptr};   // N each: radix sort index buffers
    uint32_t *hist = nullptr;              // 256 * ntiles + 1
    uint32_t *sums = nullptr;              // (256 * ntiles + kSortTile - 1) / kSortTile + 2
    uint32_t *flag = nullptr;              // N
    uint32_t *ex = nullptr;                // N + 1
    uint32_t *sums2 = nullptr;             // (N + kSortTile - 1) / kSortTile + 2
    int64_t *dense = nullptr, *dcol = nullptr;  // N each: numberings
    uint64_t *u = nullptr;                 // N: keys as unsigned offsets
    long long *mm = nullptr;               // 2: min, max
    int64_t ntiles = 0;
};

// Carves the scratch of an N-observation grouping call out of one block of
// memory, each piece 256-byte aligned; returns the bytes used (base may be
// nullptr to only measure).
static size_t carve_group_scratch(unsigned char *base, int64_t N, GroupScratch &gs) {
    size_t off = 0;
    auto take = [&](size_t bytes) -> unsigned char * {
        unsigned char *q = base ? base + off : nullptr;
        off += (std::max<size_t>(bytes, 1) + 255) / 256 * 256;
        return q;
    };
    gs.ntiles = (N + kSortTile - 1) / kSortTile;
    for (int b = 0; b < 2; b++) gs.kb[b] = (uint64_t *)take(sizeof(uint64_t) * N);
    for (int b = 0; b < 2; b++) gs.vb[b] = (int32_t *)take(sizeof(int32_t) * N);
    gs.hist = (uint32_t *)take(sizeof(uint32_t) * (256 * gs.ntiles + 1));
    gs.sums = (uint32_t *)take(sizeof(uint32_t) * ((256 * gs.ntiles + kSortTile - 1) / kSortTile + 2));
    gs.flag = (uint32_t *)take(sizeof(uint32_t) * N);
    gs.ex = (uint32_t *)take(sizeof(uint32_t) * (N + 1));
    gs.sums2 = (uint32_t *)take(sizeof(uint32_t) * ((N + kSortTile - 1) / kSortTile + 2));
    gs.dense = (int64_t *)take(sizeof(int64_t) * N);
    gs.dcol = (int64_t *)take(sizeof(int64_t) * N);
    gs.u = (uint64_t *)take(sizeof(uint64_t) * N);
    gs.mm = (long long *)take(sizeof(long long) * 2);
    return off;
}

__global__ void k_minmax_init(long long *mm) {
    if (threadIdx.x == 0) {
        mm[0] = INT64_MAX;
        mm[1] = INT64_MIN;
    }
}

static int number_keys(hipStream_t st, GroupScratch &gs, const uint64_t *u, int64_t n, int bits, int64_t *dense,
                       int64_t *ngroups) {
    const int passes = std::max(1, (bits + 7) / 8);
    const uint64_t *kin = u;
    const int32_t *vin = nullptr;
    int buf = 0;
    for (int ps = 0; ps < passes; ps++) {
        uint64_t *kout = gs.kb[buf];
        int32_t *vout = gs.vb[buf];
        hipLaunchKernelGGL(k_radix_hist<uint64_t>, dim3((unsigned)gs.ntiles), dim3(kBlock), 0, st, kin, n, 8 * ps,
                           gs.hist, gs.ntiles);
        CHK(exclusive_scan_u32(st, gs.hist, 256 * gs.ntiles, gs.hist, false, gs.sums));
        hipLaunchKernelGGL(k_radix_scatter<uint64_t>, dim3((unsigned)gs.ntiles), dim3(kBlock), 0, st, kin, vin, kout,
                           vout, n, 8 * ps, gs.hist, gs.ntiles);
        HIPCHK(hipGetLastError());
        kin = kout;
        vin = vout;
        buf ^= 1;
    }
    hipLaunchKernelGGL(k_group_heads, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, kin, n, gs.flag);
    CHK(exclusive_scan_u32(st, gs.flag, n, gs.ex, true, gs.sums2));
    hipLaunchKernelGGL(k_group_assign, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, st, (const uint32_t *)gs.ex,
                       (const uint32_t *)gs.flag, vin, n, dense);
    HIPCHK(hipGetLastError());
    uint32_t total = 0;
    HIPCHK(hipMemcpyAsync(&total, gs.ex + n, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    *ngroups = total;
    return 0;
}

static int bits_of(uint64_t range) {
    int b = 0;
    while (b < 64 && (range >> b) != 0) b++;
    return b;
}

// The grouping on device-resident keys (dkeys[c]: N int64 each) into drefs
// (N int64, 1-based), with every buffer taken from gs; on stream st.
static int group_keys_device(hipStream_t st, GroupScratch &gs, int64_t N, int32_t ncols,
                             const int64_t *const *dkeys, int64_t *drefs, int64_t *ngroups) {
    int64_t n_prev = 0;
    for (int c = 0; c < ncols; c++) {
        hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(64), 0, st, gs.mm);
        hipLaunchKernelGGL(k_minmax_i64, dim3((unsigned)std::min<int64_t>(grid_for(N, kBlock), 4096)), dim3(kBlock),
                           0, st, dkeys[c], N, gs.mm);
        long long hm[2];
        HIPCHK(hipMemcpyAsync(hm, gs.mm, sizeof(hm), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        hipLaunchKernelGGL(k_key_offset, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, st, dkeys[c], N, (int64_t)hm[0],
                           gs.u);
        int64_t nc = 0;
        CHK(number_keys(st, gs, gs.u, N, bits_of((uint64_t)hm[1] - (uint64_t)hm[0]), c == 0 ? gs.dense : gs.dcol,
                        &nc));
        if (c == 0) {
            n_prev = nc;
            continue;
        }
        // combinations of distinct values: n_prev * nc <= N * N < 2^63
        hipLaunchKernelGGL(k_group_combine, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, st,
                           (const int64_t *)gs.dense, (const int64_t *)gs.dcol, N, nc, gs.u);
        CHK(number_keys(st, gs, gs.u, N, bits_of((uint64_t)(n_prev * nc - 1)), gs.dense, &n_prev));
    }
    hipLaunchKernelGGL(k_plus_one, dim3(grid_for(N, kBlock)), dim3(kBlock), 0, st, (const int64_t *)gs.dense, N,
                       drefs);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    *ngroups = n_prev;
    return 0;
}

// fehip_group: host keys in, host refs out, device buffers allocated per call
static int group_keys(fehip_solver *h, int64_t N, int32_t ncols, const int64_t *const *keys, int64_t *refs,
                      int64_t *ngroups) {
    GroupScratch gs;
    const size_t bytes = carve_group_scratch(nullptr, N, gs);
    unsigned char *block = nullptr;
    int64_t *dkeys = nullptr, *drefs = nullptr;
    CHK(dalloc(h, &block, bytes));
    carve_group_scratch(block, N, gs);
    CHK(dalloc(h, &dkeys, (size_t)N * ncols));
    CHK(dalloc(h, &drefs, N));
    std::vector<const int64_t *> dk(ncols);
    for (int c = 0; c < ncols; c++) {
        dk[c] = dkeys + (size_t)c * N;
        HIPCHK(hipMemcpyAsync(dkeys + (size_t)c * N, keys[c], sizeof(int64_t) * N, hipMemcpyHostToDevice, h->stream));
    }
    CHK(group_keys_device(h->stream, gs, N, ncols, dk.data(), drefs, ngroups));
    HIPCHK(hipMemcpyAsync(refs, drefs, sizeof(int64_t) * N, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

// The grouping without allocation (fehip_group_device_noalloc): its scratch
// is one file-scope device array, each GPU holding its own copy for the life
// of the process.  Host code never passes the array's name to a kernel (on
// the host it is only a stand-in); the device address comes from
// hipGetSymbolAddress, once per device with that device current.
constexpr int64_t kGroupNoAllocMaxN = int64_t(1) << 20;
constexpr size_t kGroupScratchBytes = size_t(60) << 20;  // carve_group_scratch(2^20) is 56.3 MiB
static __device__ __align__(256) unsigned char scratch[kGroupScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[kMaxDevices] = {};

// device address of this GPU's scratch array
static int scratch_on_device(unsigned char **out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(FEHIP_EARG, "device index beyond " + std::to_string(kMaxDevices));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        unsigned char *ptr = nullptr;
        HIPCHK(hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

int fehip_create(fehip_t **out, int64_t N, int32_t G, const int64_t *const *refs, const int64_t *nlevels,
                 const double *const *interaction, const double *weights, int32_t precision) {
    g_err.clear();
    return create_impl<int64_t>(out, N, G, refs, nlevels, interaction, weights, precision);
}

int fehip_create_i32(fehip_t **out, int64_t N, int32_t G, const int32_t *const *refs, const int64_t *nlevels,
                     const double *const *interaction, const double *weights, int32_t precision) {
    g_err.clear();
    return create_impl<int32_t>(out, N, G, refs, nlevels, interaction, weights, precision);
}

int fehip_update_weights(fehip_t *h, const double *weights) {
    g_err.clear();
    if (!h) return fail(FEHIP_EARG, "handle is NULL");
    return do_update_weights(h, weights);
}

int fehip_solve_residuals_device(fehip_t *h, double *dY, int32_t K, int64_t ldY, double tol, int64_t maxiter,
                                 int64_t *iters, uint8_t *converged) {
    g_err.clear();
    if (!h) return fail(FEHIP_EARG, "handle is NULL");
    if (K < 1) return fail(FEHIP_EARG, "K must be >= 1");
    if (!dY) return fail(FEHIP_EARG, "dY is NULL");
    if (ldY < h->N) return fail(FEHIP_EDIM, "ldY < N");
    if (maxiter < 1) return fail(FEHIP_EARG, "maxiter must be >= 1");
    const int bw = max_batch(h);
    if (bw < 1) return fail(FEHIP_EDIM, "p * 8 bytes exceeds 2^32");
    for (int32_t c0 = 0; c0 < K; c0 += bw) {
        const int kb = std::min<int32_t>(bw, K - c0);
        CHK(solve_residuals_dev(h, dY + (int64_t)c0 * ldY, kb, ldY, tol, maxiter, iters ? iters + c0 : nullptr,
                                converged ? converged + c0 : nullptr));
    }
    return 0;
}

int fehip_solve_residuals_panels(fehip_t *const *hs, int32_t npanels, double *const *dY, const int32_t *K,
                                 const int64_t *ldY, double tol, int64_t maxiter, int64_t *const *iters,
                                 uint8_t *const *converged) {
    g_err.clear();
    if (!hs || !dY || !K || !ldY || npanels < 0) return fail(FEHIP_EARG, "bad arguments");
    if (maxiter < 1) return fail(FEHIP_EARG, "maxiter must be >= 1");
    for (int32_t i = 0; i < npanels; i++) {
        if (!hs[i] || !dY[i]) return fail(FEHIP_EARG, "panel " + std::to_string(i) + ": NULL handle or column block");
        if (K[i] < 1 || K[i] > kMaxK) return fail(FEHIP_EARG, "panel " + std::to_string(i) + ": K must be in 1..32");
        if (ldY[i] < hs[i]->N) return fail(FEHIP_EDIM, "panel " + std::to_string(i) + ": ldY < N");
        for (int32_t j = 0; j < i; j++)
            if (hs[j] == hs[i]) return fail(FEHIP_EARG, "a handle appears twice (one stream per panel)");
    }
    // one resumable job per panel, each on its handle's own stream; mixed
    // precision panels (two LSMR phases) run as one blocking step
    struct AnyJob {
        std::function<int()> begin, end;
        std::function<int(bool &)> step;
        bool done = false;
    };
    std::vector<AnyJob> jobs(npanels);
    for (int32_t i = 0; i < npanels; i++) {
        fehip_solver *h = hs[i];
        int64_t *it = iters ? iters[i] : nullptr;
        uint8_t *cv = converged ? converged[i] : nullptr;
        if (h->precision == FEHIP_MIXED) {
            double *d = dY[i];
            const int32_t k = K[i];
            const int64_t ld = ldY[i];
            jobs[i].begin = [=] { return solve_residuals_dev(h, d, k, ld, tol, maxiter, it, cv); };
            jobs[i].step = [](bool &f) {
                f = true;
                return 0;
            };
            jobs[i].end = [] { return 0; };
            continue;
        }
        CHK(dispatch(h->precision, padded_K(K[i]), [&](auto e) {
            using E = decltype(e);
            auto j = std::make_shared<typename E::Job>();
            j->h = h;
            j->dY = dY[i];
            j->Kreal = K[i];
            j->ldY = ldY[i];
            j->iters = it;
            j->conv = cv;
            jobs[i].begin = [j, tol, maxiter] { return E::job_begin(*j, tol, maxiter); };
            jobs[i].step = [j](bool &f) { return E::job_step(*j, f); };
            jobs[i].end = [j] { return E::job_end(*j); };
            return 0;
        }));
    }
    int rc = 0;
    for (int32_t i = 0; i < npanels && rc == 0; i++) rc = jobs[i].begin();
    for (int32_t left = npanels; rc == 0 && left > 0;) {
        for (int32_t i = 0; i < npanels && rc == 0; i++) {
            if (jobs[i].done) continue;
            bool fin = false;
            rc = jobs[i].step(fin);
            if (rc == 0 && fin) {
                rc = jobs[i].end();
                jobs[i].done = true;
                left--;
            }
        }
    }
    if (rc != 0)
        for (int32_t i = 0; i < npanels; i++) (void)hipStreamSynchronize(hs[i]->stream);
    return rc;
}

int fehip_solve_residuals_panels_host(fehip_t *const *hs, int32_t npanels, double *const *const *cols,
                                      const int32_t *K, double tol, int64_t maxiter, int64_t *const *iters,
                                      uint8_t *const *converged) {
    g_err.clear();
    if (!hs || !cols || !K || npanels < 0) return fail(FEHIP_EARG, "bad arguments");
    for (int32_t i = 0; i < npanels; i++) {
        if (!hs[i] || !cols[i]) return fail(FEHIP_EARG, "panel " + std::to_string(i) + ": NULL handle or columns");
        if (K[i] < 1 || K[i] > kMaxK) return fail(FEHIP_EARG, "panel " + std::to_string(i) + ": K must be in 1..32");
        for (int32_t c = 0; c < K[i]; c++)
            if (!cols[i][c]) return fail(FEHIP_EARG, "panel " + std::to_string(i) + ": NULL column");
    }
    // each panel's columns are staged on its own handle (and stream)
    std::vector<double *> dY(npanels, nullptr);
    std::vector<int64_t> ld(npanels);
    int rc = 0;
    for (int32_t i = 0; i < npanels && rc == 0; i++) {
        fehip_solver *h = hs[i];
        ld[i] = h->N;
        rc = dalloc(h, &dY[i], (size_t)h->N * K[i]);
        for (int32_t c = 0; c < K[i] && rc == 0; c++) {
            const hipError_t e = hipMemcpyAsync(dY[i] + (int64_t)c * h->N, cols[i][c], sizeof(double) * h->N,
                                                hipMemcpyHostToDevice, h->stream);
            if (e != hipSuccess) rc = fail(FEHIP_EHIP, hipGetErrorString(e));
        }
    }
    if (rc == 0) rc = fehip_solve_residuals_panels(hs, npanels, dY.data(), K, ld.data(), tol, maxiter, iters, converged);
    for (int32_t i = 0; i < npanels && rc == 0; i++) {
        fehip_solver *h = hs[i];
        for (int32_t c = 0; c < K[i] && rc == 0; c++) {
            const hipError_t e = hipMemcpyAsync(cols[i][c], dY[i] + (int64_t)c * h->N, sizeof(double) * h->N,
                                                hipMemcpyDeviceToHost, h->stream);
            if (e != hipSuccess) rc = fail(FEHIP_EHIP, hipGetErrorString(e));
        }
    }
    for (int32_t i = 0; i < npanels; i++) {
        if (hs[i]) (void)hipStreamSynchronize(hs[i]->stream);
        if (dY[i]) dfree(hs[i], dY[i]);
    }
    return rc;
}

int fehip_solve_residuals(fehip_t *h, double *const *cols, int32_t K, double tol, int64_t maxiter, int64_t *iters,
                          uint8_t *converged) {
    g_err.clear();
    if (!h) return fail(FEHIP_EARG, "handle is NULL");
    if (K < 1 || !cols) return fail(FEHIP_EARG, "need K >= 1 columns");
    if (maxiter < 1) return fail(FEHIP_EARG, "maxiter must be >= 1");
    const int bw = max_batch(h);
    if (bw < 1) return fail(FEHIP_EDIM, "p * 8 bytes exceeds 2^32");
    const int32_t kb = std::min<int32_t>(bw, K);
    double *dY = nullptr;
    CHK(dalloc(h, &dY, (size_t)h->N * kb));
    int rc = 0;
    for (int32_t c0 = 0; c0 < K && rc == 0; c0 += bw) {
        const int k = std::min<int32_t>(bw, K - c0);
        for (int c = 0; c < k && rc == 0; c++) {
            hipError_t e = hipMemcpyAsync(dY + (int64_t)c * h->N, cols[c0 + c], sizeof(double) * h->N,
                                          hipMemcpyHostToDevice, h->stream);
            if (e != hipSuccess) rc = fail(FEHIP_EHIP, hipGetErrorString(e));
        }
        if (rc == 0)
            rc = solve_residuals_dev(h, dY, k, h->N, tol, maxiter, iters ? iters + c0 : nullptr,
                                     converged ? converged + c0 : nullptr);
        for (int c = 0; c < k && rc == 0; c++) {
            hipError_t e = hipMemcpyAsync(cols[c0 + c], dY + (int64_t)c * h->N, sizeof(double) * h->N,
                                          hipMemcpyDeviceToHost, h->stream);
            if (e != hipSuccess) rc = fail(FEHIP_EHIP, hipGetErrorString(e));
        }
        if (rc == 0) {
            hipError_t e = hipStreamSynchronize(h->stream);
            if (e != hipSuccess) rc = fail(FEHIP_EHIP, hipGetErrorString(e));
        }
    }
    dfree(h, dY);
    return rc;
}

int fehip_solve_coefficients(fehip_t *h, const double *y, double tol, int64_t maxiter, double *const *coef,
                             int64_t *iters, uint8_t *converged) {
    g_err.clear();
    if (!h) return fail(FEHIP_EARG, "handle is NULL");
    if (!y || !coef) return fail(FEHIP_EARG, "y and coef are required");
    if (maxiter < 0) return fail(FEHIP_EARG, "maxiter must be >= 0");
    double *dy = nullptr, *dc = nullptr;
    CHK(dalloc(h, &dy, h->N));
    CHK(dalloc(h, &dc, h->p));
    int rc = 0;
    {
        hipError_t e = hipMemcpyAsync(dy, y, sizeof(double) * h->N, hipMemcpyHostToDevice, h->stream);
        if (e != hipSuccess) rc = fail(FEHIP_EHIP, hipGetErrorString(e));
    }
    if (rc == 0) rc = solve_coefficients_dev(h, dy, tol, maxiter, dc, iters, converged);
    if (rc == 0) rc = s_block_to_caller(h, dc);
    for (int g = 0; g < h->G && rc == 0; g++) {
        hipError_t e = hipMemcpy(coef[g], dc + h->off[g], sizeof(double) * h->n[g], hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = fail(FEHIP_EHIP, hipGetErrorString(e));
    }
    dfree(h, dy);
    dfree(h, dc);
    return rc;
}

int fehip_solve_coefficients_full(fehip_t *h, const double *y, double tol, int64_t maxiter, double *const *coef,
                                  double *const *obs, int64_t *iters, uint8_t *converged) {
    g_err.clear();
    if (!h) return fail(FEHIP_EARG, "handle is NULL");
    if (!y) return fail(FEHIP_EARG, "y is required");
    if (maxiter < 0) return fail(FEHIP_EARG, "maxiter must be >= 0");
    double *dy = nullptr, *dc = nullptr;
    CHK(dalloc(h, &dy, h->N));
    CHK(dalloc(h, &dc, h->p));
    int rc = 0;
    {
        hipError_t e = hipMemcpyAsync(dy, y, sizeof(double) * h->N, hipMemcpyHostToDevice, h->stream);
        if (e != hipSuccess) rc = fail(FEHIP_EHIP, hipGetErrorString(e));
    }
    if (rc == 0) rc = solve_coefficients_dev(h, dy, tol, maxiter, dc, iters, converged);
    if (rc == 0) rc = normalize_dev(h, dc);
    // dy is free once the solve has read it: it holds one FE's expansion at a time
    for (int g = 0; obs && g < h->G && rc == 0; g++) {
        if (!obs[g]) continue;
        rc = full_dev(h, dc, g, dy);
        if (rc) break;
        hipError_t e = hipMemcpyAsync(obs[g], dy, sizeof(double) * h->N, hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) rc = fail(FEHIP_EHIP, hipGetErrorString(e));
    }
    if (rc == 0 && coef) rc = s_block_to_caller(h, dc);
    for (int g = 0; coef && g < h->G && rc == 0; g++) {
        if (!coef[g]) continue;
        hipError_t e = hipMemcpyAsync(coef[g], dc + h->off[g], sizeof(double) * h->n[g], hipMemcpyDeviceToHost,
                                      h->stream);
        if (e != hipSuccess) rc = fail(FEHIP_EHIP, hipGetErrorString(e));
    }
    if (rc == 0) {
        hipError_t e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) rc = fail(FEHIP_EHIP, hipGetErrorString(e));
    }
    dfree(h, dy);
    dfree(h, dc);
    return rc;
}

int fehip_group(int64_t N, int32_t ncols, const int64_t *const *keys, int64_t *refs, int64_t *ngroups) {
    g_err.clear();
    if (N < 0 || ncols < 1 || !keys || !refs || !ngroups) return fail(FEHIP_EARG, "bad arguments");
    for (int c = 0; c < ncols; c++)
        if (!keys[c]) return fail(FEHIP_EARG, "keys[c] is required for every column");
    if (N == 0) {
        *ngroups = 0;
        return 0;
    }
    if (N >= (int64_t)INT32_MAX) return fail(FEHIP_EARG, "too many observations for one grouping call");
    fehip_solver *h = new fehip_solver();
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    int rc = e == hipSuccess ? group_keys(h, N, ncols, keys, refs, ngroups) : fail(FEHIP_EHIP, hipGetErrorString(e));
    fehip_destroy(h);
    return rc;
}

int fehip_group_device_noalloc(int64_t N, int32_t ncols, const int64_t *const *dkeys, int64_t *drefs,
                               int64_t *ngroups) {
    g_err.clear();
    if (N < 0 || ncols < 1 || !dkeys || !drefs || !ngroups) return fail(FEHIP_EARG, "bad arguments");
    for (int c = 0; c < ncols; c++)
        if (!dkeys[c]) return fail(FEHIP_EARG, "dkeys[c] is required for every column");
    if (N > kGroupNoAllocMaxN)
        return fail(FEHIP_EDIM, "N = " + std::to_string(N) + " exceeds the no-allocation grouping's maximum of " +
                                    std::to_string(kGroupNoAllocMaxN) +
                                    " observations (its device scratch has a fixed size); use fehip_group");
    if (N == 0) {
        *ngroups = 0;
        return 0;
    }
    unsigned char *base = nullptr;
    CHK(scratch_on_device(&base));
    GroupScratch gs;
    if (carve_group_scratch(nullptr, N, gs) > kGroupScratchBytes)
        return fail(FEHIP_EHIP, "grouping scratch layout exceeds the device array");
    carve_group_scratch(base, N, gs);
    return group_keys_device(nullptr, gs, N, ncols, dkeys, drefs, ngroups);
}

int fehip_normalize(fehip_t *h, double *const *coef) {
    g_err.clear();
    if (!h || !coef) return fail(FEHIP_EARG, "bad arguments");
    for (int g = 0; g < h->G; g++)
        if (!coef[g]) return fail(FEHIP_EARG, "coef[g] is required for every fixed effect");
    double *dc = nullptr;
    CHK(dalloc(h, &dc, h->p));
    int rc = 0;
    for (int g = 0; g < h->G && rc == 0; g++) {
        hipError_t e = hipMemcpyAsync(dc + h->off[g], coef[g], sizeof(double) * h->n[g], hipMemcpyHostToDevice,
                                      h->stream);
        if (e != hipSuccess) rc = fail(FEHIP_EHIP, hipGetErrorString(e));
    }
    if (rc == 0) rc = s_block_to_internal(h, dc);
    if (rc == 0) rc = normalize_dev(h, dc);
    if (rc == 0) rc = s_block_to_caller(h, dc);
    for (int g = 0; g < h->G && rc == 0; g++) {
        hipError_t e = hipMemcpyAsync(coef[g], dc + h->off[g], sizeof(double) * h->n[g], hipMemcpyDeviceToHost,
                                      h->stream);
        if (e != hipSuccess) rc = fail(FEHIP_EHIP, hipGetErrorString(e));
    }
    if (rc == 0) {
        hipError_t e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) rc = fail(FEHIP_EHIP, hipGetErrorString(e));
    }
    dfree(h, dc);
    return rc;
}

int fehip_features(fehip_t *h, const double *y, float *feats, double *coef_ap, double *sigma) {
    g_err.clear();
    if (!h || !y || !feats) return fail(FEHIP_EARG, "bad arguments");
    double *dy = nullptr, *dca = nullptr;
    CHK(dalloc(h, &dy, h->N));
    CHK(dalloc(h, &dca, h->p));
    int rc = 0;
    hipError_t e = hipMemcpyAsync(dy, y, sizeof(double) * h->N, hipMemcpyHostToDevice, h->stream);
    if (e != hipSuccess) rc = fail(FEHIP_EHIP, hipGetErrorString(e));
    const int prec = h->precision == FEHIP_MIXED ? FEHIP_FP64 : h->precision;
    if (rc == 0)
        rc = dispatch(prec, 1, [&](auto eng) -> int {
            using E1 = Engine<typename decltype(eng)::TT, typename decltype(eng)::AccT, typename decltype(eng)::TrT, 1>;
            CHK(E1::ensure_ws(h));
            return E1::features(h, dy, h->ws_feats, dca);
        });
    if (rc == 0 && h->s_int_of) {
        // FE s's rows and coefficients in the caller's numbering
        const int64_t ns = h->n[h->s], o = h->off[h->s];
        float *rows = nullptr;
        CHK(dalloc(h, &rows, ns * 16));
        hipLaunchKernelGGL(k_gather_rows, dim3(grid_for(ns * 16, kBlock)), dim3(kBlock), 0, h->stream,
                           (const float *)h->ws_feats + o * 16, (const int32_t *)h->s_int_of, ns, 16, rows);
        HIPCHK(hipMemcpyAsync(h->ws_feats + o * 16, rows, sizeof(float) * ns * 16, hipMemcpyDeviceToDevice,
                              h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        dfree(h, rows);
        rc = s_block_to_caller(h, dca);
    }
    if (rc == 0) {
        HIPCHK(hipMemcpyAsync(feats, h->ws_feats, sizeof(float) * h->p * 16, hipMemcpyDeviceToHost, h->stream));
        if (coef_ap) HIPCHK(hipMemcpyAsync(coef_ap, dca, sizeof(double) * h->p, hipMemcpyDeviceToHost, h->stream));
        if (sigma) HIPCHK(hipMemcpyAsync(sigma, h->ws_sig, sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
    }
    dfree(h, dy);
    dfree(h, dca);
    return rc;
}

static int solve_from_common(fehip_t *h, double *const *cols, int32_t K, double tol, int64_t maxiter,
                             const double *const *coef0, fehip_mlp_t *net, int64_t *iters, uint8_t *converged) {
    if (!h) return fail(FEHIP_EARG, "handle is NULL");
    if (K < 1 || K > kMaxK || !cols) return fail(FEHIP_EARG, "need 1..32 columns");
    if (maxiter < 1) return fail(FEHIP_EARG, "maxiter must be >= 1");
    double *dY = nullptr;
    CHK(dalloc(h, &dY, (size_t)h->N * K));
    int rc = 0;
    for (int c = 0; c < K && rc == 0; c++) {
        hipError_t e = hipMemcpyAsync(dY + (int64_t)c * h->N, cols[c], sizeof(double) * h->N, hipMemcpyHostToDevice,
                                      h->stream);
        if (e != hipSuccess) rc = fail(FEHIP_EHIP, hipGetErrorString(e));
    }
    const int prec = h->precision == FEHIP_MIXED ? FEHIP_FP64 : h->precision;
    if (rc == 0)
        rc = dispatch(prec, padded_K(K), [&](auto eng) {
            return decltype(eng)::solve_residuals_from(h, dY, K, h->N, tol, maxiter, coef0, net, iters, converged);
        });
    for (int c = 0; c < K && rc == 0; c++) {
        hipError_t e = hipMemcpyAsync(cols[c], dY + (int64_t)c * h->N, sizeof(double) * h->N, hipMemcpyDeviceToHost,
                                      h->stream);
        if (e != hipSuccess) rc = fail(FEHIP_EHIP, hipGetErrorString(e));
    }
    if (rc == 0) {
        hipError_t e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) rc = fail(FEHIP_EHIP, hipGetErrorString(e));
    }
    dfree(h, dY);
    return rc;
}

int fehip_solve_residuals_x0(fehip_t *h, double *const *cols, int32_t K, double tol, int64_t maxiter,
                             const double *const *coef0, int64_t *iters, uint8_t *converged) {
    g_err.clear();
    return solve_from_common(h, cols, K, tol, maxiter, coef0, nullptr, iters, converged);
}

int fehip_solve_residuals_learned(fehip_t *h, double *const *cols, int32_t K, double tol, int64_t maxiter,
                                  fehip_mlp_t *net, int64_t *iters, uint8_t *converged) {
    g_err.clear();
    if (!net) return fail(FEHIP_EARG, "network is NULL");
    return solve_from_common(h, cols, K, tol, maxiter, nullptr, net, iters, converged);
}

int fehip_solve_residuals_learned_device(fehip_t *h, double *dY, int32_t K, int64_t ldY, double tol,
                                         int64_t maxiter, fehip_mlp_t *net, int64_t *iters, uint8_t *converged) {
    g_err.clear();
    if (!h || !net || !dY) return fail(FEHIP_EARG, "bad arguments");
    if (K < 1 || K > kMaxK) return fail(FEHIP_EARG, "need 1..32 columns");
    if (ldY < h->N) return fail(FEHIP_EDIM, "ldY < N");
    if (maxiter < 1) return fail(FEHIP_EARG, "maxiter must be >= 1");
    const int prec = h->precision == FEHIP_MIXED ? FEHIP_FP64 : h->precision;
    return dispatch(prec, padded_K(K), [&](auto eng) {
        return decltype(eng)::solve_residuals_from(h, dY, K, ldY, tol, maxiter, nullptr, net, iters, converged);
    });
}

int fehip_solve_coefficients_x0(fehip_t *h, const double *y, double tol, int64_t maxiter, const double *coef0,
                                double *const *coef, int64_t *iters, uint8_t *converged) {
    g_err.clear();
    if (!h || !y || !coef) return fail(FEHIP_EARG, "bad arguments");
    if (maxiter < 0) return fail(FEHIP_EARG, "maxiter must be >= 0");
    double *dy = nullptr, *dc = nullptr;
    CHK(dalloc(h, &dy, h->N));
    CHK(dalloc(h, &dc, h->p));
    int rc = 0;
    hipError_t e = hipMemcpyAsync(dy, y, sizeof(double) * h->N, hipMemcpyHostToDevice, h->stream);
    if (e != hipSuccess) rc = fail(FEHIP_EHIP, hipGetErrorString(e));
    const int prec = h->precision == FEHIP_MIXED ? FEHIP_FP64 : h->precision;
    FwlScope fwl(h);
    if (rc == 0)
        rc = dispatch(prec, 1, [&](auto eng) {
            return decltype(eng)::solve_coefficients(h, dy, tol, maxiter, dc, iters, converged, coef0);
        });
    if (rc == 0) rc = s_block_to_caller(h, dc);
    for (int g = 0; g < h->G && rc == 0; g++) {
        hipError_t e2 = hipMemcpyAsync(coef[g], dc + h->off[g], sizeof(double) * h->n[g], hipMemcpyDeviceToHost,
                                       h->stream);
        if (e2 != hipSuccess) rc = fail(FEHIP_EHIP, hipGetErrorString(e2));
    }
    if (rc == 0) {
        hipError_t e2 = hipStreamSynchronize(h->stream);
        if (e2 != hipSuccess) rc = fail(FEHIP_EHIP, hipGetErrorString(e2));
    }
    dfree(h, dy);
    dfree(h, dc);
    return rc;
}

int fehip_get_csr(fehip_t *h, int32_t g, int32_t *perm, int32_t *offsets) {
    g_err.clear();
    if (!h) return fail(FEHIP_EARG, "handle is NULL");
    if (g < 0 || g >= h->G) return fail(FEHIP_EARG, "no such fixed effect");
    const FeInfo &f = h->fe[g];
    if (g == h->s && h->s_int_of) {
        // the blocks of the internal numbering, in caller level order
        std::vector<int32_t> pi(h->N), oi(f.n + 1);
        HIPCHK(hipMemcpy(pi.data(), f.operm, sizeof(int32_t) * h->N, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(oi.data(), f.ooff, sizeof(int32_t) * (f.n + 1), hipMemcpyDeviceToHost));
        int32_t at = 0;
        for (int64_t k = 0; k < f.n; k++) {
            const int32_t j = h->h_s_int_of[k];
            if (offsets) offsets[k] = at;
            for (int32_t q = oi[j]; q < oi[j + 1]; q++, at++)
                if (perm) perm[at] = pi[q];
        }
        if (offsets) offsets[f.n] = at;
        return 0;
    }
    if (perm) HIPCHK(hipMemcpy(perm, f.operm, sizeof(int32_t) * h->N, hipMemcpyDeviceToHost));
    if (offsets) HIPCHK(hipMemcpy(offsets, f.ooff, sizeof(int32_t) * (f.n + 1), hipMemcpyDeviceToHost));
    return 0;
}

int fehip_get_scale(fehip_t *h, int32_t g, double *scale) {
    g_err.clear();
    if (!h) return fail(FEHIP_EARG, "handle is NULL");
    if (g < 0 || g >= h->G || !scale) return fail(FEHIP_EARG, "bad arguments");
    HIPCHK(hipMemcpy(scale, h->fe[g].scale_d, sizeof(double) * h->fe[g].n, hipMemcpyDeviceToHost));
    if (g == h->s && h->s_int_of) {
        const std::vector<double> si(scale, scale + h->fe[g].n);
        for (int64_t k = 0; k < h->fe[g].n; k++) scale[k] = si[h->h_s_int_of[k]];
    }
    return 0;
}

int fehip_get_cache(fehip_t *h, int32_t g, double *cache) {
    g_err.clear();
    if (!h) return fail(FEHIP_EARG, "handle is NULL");
    if (g < 0 || g >= h->G || !cache) return fail(FEHIP_EARG, "bad arguments");
    double *tmp = nullptr, *tmp2 = nullptr;
    CHK(dalloc(h, &tmp, h->N));
    CHK(dalloc(h, &tmp2, h->N));
    if (h->has64)
        hipLaunchKernelGGL(k_to_double<double>, dim3(grid_for(h->N, kBlock)), dim3(kBlock), 0, h->stream,
                           (const double *)h->fe[g].cache_int[0], tmp, h->N);
    else
        hipLaunchKernelGGL(k_to_double<float>, dim3(grid_for(h->N, kBlock)), dim3(kBlock), 0, h->stream,
                           (const float *)h->fe[g].cache_int[1], tmp, h->N);
    hipLaunchKernelGGL(k_scatter_perm<double>, dim3(grid_for(h->N, kBlock)), dim3(kBlock), 0, h->stream, tmp,
                       h->perm_s, h->N, tmp2);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(cache, tmp2, sizeof(double) * h->N, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    dfree(h, tmp);
    dfree(h, tmp2);
    return 0;
}

static int mul_common(fehip_t *h, double *io, const double *other, double a, double b, bool adjoint) {
    g_err.clear();
    if (!h) return fail(FEHIP_EARG, "handle is NULL");
    if (!io || !other) return fail(FEHIP_EARG, "bad arguments");
    double *dy = nullptr, *dx = nullptr;
    CHK(dalloc(h, &dy, h->N));
    CHK(dalloc(h, &dx, h->p));
    const double *y_host = adjoint ? other : io;
    double *x_host = adjoint ? io : (double *)other;
    HIPCHK(hipMemcpy(dy, y_host, sizeof(double) * h->N, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dx, x_host, sizeof(double) * h->p, hipMemcpyHostToDevice));
    CHK(s_block_to_internal(h, dx));
    int rc = dispatch(h->precision == FEHIP_MIXED ? FEHIP_FP64 : h->precision, 1, [&](auto e) {
        using E = decltype(e);
        return adjoint ? E::mul_At(h, dx, dy, a, b) : E::mul_A(h, dy, dx, a, b);
    });
    if (rc == 0 && adjoint) rc = s_block_to_caller(h, dx);
    if (rc == 0) {
        if (adjoint) HIPCHK(hipMemcpy(io, dx, sizeof(double) * h->p, hipMemcpyDeviceToHost));
        else HIPCHK(hipMemcpy(io, dy, sizeof(double) * h->N, hipMemcpyDeviceToHost));
    }
    dfree(h, dy);
    dfree(h, dx);
    return rc;
}

int fehip_mul_A(fehip_t *h, double *y, const double *x, double a, double b) {
    return mul_common(h, y, x, a, b, false);
}

int fehip_mul_At(fehip_t *h, double *x, const double *y, double a, double b) {
    return mul_common(h, x, y, a, b, true);
}

int fehip_lsmr_info(fehip_t *h, int32_t col, double *out) {
    g_err.clear();
    if (!h || !out) return fail(FEHIP_EARG, "bad arguments");
    if (col < 0 || (size_t)(col + 1) * 9 > h->last_info.size())
        return fail(FEHIP_EARG, "no LSMR result for that column");
    std::memcpy(out, h->last_info.data() + (size_t)col * 9, sizeof(double) * 9);
    return 0;
}

int fehip_time_kernels(fehip_t *h, int32_t K, int32_t reps, double *out_ms) {
    g_err.clear();
    if (!h || !out_ms) return fail(FEHIP_EARG, "bad arguments");
    if (K < 1 || K > kMaxK || reps < 1) return fail(FEHIP_EARG, "K must be in 1..32 and reps >= 1");
    return dispatch(h->precision, padded_K(K), [&](auto e) { return decltype(e)::time_kernels(h, reps, out_ms); });
}

int fehip_plan_info(fehip_t *h, int32_t K, int32_t fp32_storage, int64_t *out) {
    g_err.clear();
    if (!h || !out) return fail(FEHIP_EARG, "bad arguments");
    if (K < 1 || K > kMaxK) return fail(FEHIP_EARG, "K must be in 1..32");
    if (fp32_storage ? !h->has32 : !h->has64) return fail(FEHIP_EARG, "the handle has no such storage precision");
    auto fill = [&](Plan *P, int wc) {
        out[0] = P->fused_dense ? 1 : 0;
        out[1] = P->R;
        out[2] = P->ntiles_s;
        out[3] = P->total_tiles;
        out[4] = P->nitems;
        out[5] = P->nslot;
        out[6] = P->pc.np;
        out[7] = P->csr_blocks;
        out[8] = wc;
        return 0;
    };
    const int Kp = padded_K(K);
    if (fp32_storage) {
        if (h->precision == FEHIP_FP32)
            return for_K<float, float, float>(Kp, [&](auto e) {
                using E = decltype(e);
                Plan *P = nullptr;
                CHK((get_plan_for<typename E::TT, typename E::AccT, E::kK>(h, &P)));
                return fill(P, fused_cols(h, sizeof(typename E::TT), E::kK));
            });
        return for_K<float, F32x2, double>(Kp, [&](auto e) {
            using E = decltype(e);
            Plan *P = nullptr;
            CHK((get_plan_for<typename E::TT, typename E::AccT, E::kK>(h, &P)));
            return fill(P, fused_cols(h, sizeof(typename E::TT), E::kK));
        });
    }
    return for_K<double, double, double>(Kp, [&](auto e) {
        using E = decltype(e);
        Plan *P = nullptr;
        CHK((get_plan_for<typename E::TT, typename E::AccT, E::kK>(h, &P)));
        return fill(P, fused_cols(h, sizeof(typename E::TT), E::kK));
    });
}

int fehip_set_algorithm(fehip_t *h, int32_t algorithm) {
    g_err.clear();
    if (!h) return fail(FEHIP_EARG, "handle is NULL");
    if (algorithm != FEHIP_LSMR && algorithm != FEHIP_LSMR_PROJECTED) return fail(FEHIP_EARG, "unknown algorithm");
    h->fwl_req = algorithm == FEHIP_LSMR_PROJECTED;
    return 0;
}

int fehip_set_schedule(fehip_t *h, int64_t tiles_s, int64_t tiles_c, int32_t fuse_dense, int32_t poll_iters) {
    g_err.clear();
    if (!h) return fail(FEHIP_EARG, "handle is NULL");
    if (tiles_s < 0 || tiles_c < 0 || tiles_s > (1 << 20) || tiles_c > (1 << 20))
        return fail(FEHIP_EARG, "tile counts must be in 0..2^20 (0 = default)");
    if (fuse_dense < -1 || fuse_dense > 1) return fail(FEHIP_EARG, "fuse_dense must be -1, 0 or 1");
    if (poll_iters < 0 || poll_iters > 1024) return fail(FEHIP_EARG, "poll_iters must be in 0..1024 (0 = keep)");
    h->sched_tiles_s = tiles_s;
    h->sched_tiles_c = tiles_c;
    h->sched_fuse_dense = fuse_dense;
    if (poll_iters > 0) h->poll_iters = poll_iters;
    destroy_all_plans(h);
    return 0;
}

int fehip_autotune(fehip_t *h, int32_t K, int32_t reps, double *out) {
    g_err.clear();
    if (!h || !out) return fail(FEHIP_EARG, "bad arguments");
    if (K < 1 || K > kMaxK || reps < 1) return fail(FEHIP_EARG, "K must be in 1..32 and reps >= 1");
    return dispatch(h->precision, padded_K(K), [&](auto e) { return decltype(e)::autotune(h, reps, out); });
}

int fehip_dims(const fehip_t *h, int64_t *N, int32_t *G, int64_t *p) {
    if (!h) return fail(FEHIP_EARG, "handle is NULL");
    if (N) *N = h->N;
    if (G) *G = h->G;
    if (p) *p = h->p;
    return 0;
}

const char *fehip_last_error(void) { return g_err.c_str(); }

void fehip_destroy(fehip_t *h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (auto &kv : h->plans) destroy_plan(h, kv.second);
    h->plans.clear();
    for (void *q : h->allocs) (void)hipFree(q);
    h->allocs.clear();
    if (h->h_flag) (void)hipHostFree(h->h_flag);
    for (int e = 0; e < 2; e++)
        if (h->poll_ev[e]) (void)hipEventDestroy(h->poll_ev[e]);
    if (h->fork_ev) (void)hipEventDestroy(h->fork_ev);
    if (h->join_ev) (void)hipEventDestroy(h->join_ev);
    if (h->stream2) {
        (void)hipStreamSynchronize(h->stream2);
        (void)hipStreamDestroy(h->stream2);
    }
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

}  // extern "C"
