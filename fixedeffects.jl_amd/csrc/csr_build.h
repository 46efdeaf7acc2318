// csr_build.h -- device construction of the per-fixed-effect CSR bucketing.
//
// The reference builds it on the host with a counting sort
// (bucketize_refs, ext/MetalExt.jl:40-64): observations grouped by level,
// ascending observation index within a level, plus level offsets.  Here it is
// a least-significant-digit radix sort of (level, observation) pairs with
// 8-bit digits; every pass is a stable scatter, so the result is the same
// stable order, bit for bit.  Offsets come from level boundaries in the sorted
// keys (no atomics).  Segment head bits and their word prefix counts drive
// the segmented D'x kernels.
#pragma once

#include "fehip_common.h"

namespace fehip {

constexpr int kSortItems = 16;                    // items per thread per tile
constexpr int kSortTile = kBlock * kSortItems;    // 4096 keys per workgroup

// Per-tile digit histogram, stored digit-major: hist[d * ntiles + tile].
// KT: int32_t (non-negative level keys) or uint64_t (grouping keys).
template <typename KT>
__global__ __launch_bounds__(kBlock) void k_radix_hist(const KT *__restrict__ keys, int64_t n,
                                                       int shift, uint32_t *__restrict__ hist,
                                                       int64_t ntiles) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
    for (int r = 0; r < kSortItems; r++) {
        int64_t i = base + (int64_t)r * kBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[(int)((keys[i] >> shift) & 255)], 1u);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter of one digit pass.  Tile elements are visited in index order
// (round r, thread t -> element r*256 + t); the rank of an element among the
// earlier elements of its digit is: the tile's global base for that digit +
// the digit count of earlier rounds + the count in lower waves of this round
// + its rank among same-digit lanes of its wave (ballot peer mask).
// vals_in == nullptr means the identity (element index).
template <typename KT>
__global__ __launch_bounds__(kBlock) void k_radix_scatter(
    const KT *__restrict__ keys_in, const int32_t *__restrict__ vals_in,
    KT *__restrict__ keys_out, int32_t *__restrict__ vals_out, int64_t n, int shift,
    const uint32_t *__restrict__ scanned, int64_t ntiles) {
    __shared__ uint32_t running[256];
    __shared__ uint32_t wcnt[kWavesPerBlock][256];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    running[tid] = scanned[(int64_t)tid * ntiles + blockIdx.x];
#pragma unroll
    for (int w = 0; w < kWavesPerBlock; w++) wcnt[w][tid] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
    const uint64_t lt = lanes_lt(lane);
    for (int r = 0; r < kSortItems; r++) {
        const int64_t i = base + (int64_t)r * kBlock + tid;
        const bool valid = i < n;
        const KT key = valid ? keys_in[i] : (KT)0;
        const int32_t val = valid ? (vals_in ? vals_in[i] : (int32_t)i) : 0;
        const int d = (int)((key >> shift) & 255);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const int bit = (d >> b) & 1;
            const uint64_t m = __ballot(valid && bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        if (valid && rank == 0) wcnt[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = running[d] + rank;
            for (int w = 0; w < wave; w++) pos += wcnt[w][d];
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int w = 0; w < kWavesPerBlock; w++) {
            add += wcnt[w][tid];
            wcnt[w][tid] = 0;
        }
        running[tid] += add;
        __syncthreads();
    }
}

// ---- exclusive scan of uint32 (three phases) ----------------------------

// Phase 1: sum of each 4096-element chunk.
__global__ __launch_bounds__(kBlock) void k_scan_chunk_sums(const uint32_t *__restrict__ in, int64_t n,
                                                            uint32_t *__restrict__ sums) {
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
    uint32_t s = 0;
    for (int r = 0; r < kSortItems; r++) {
        int64_t i = base + (int64_t)r * kBlock + threadIdx.x;
        if (i < n) s += in[i];
    }
    // integer block sum (exact, order-free)
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, kWave);
    __shared__ uint32_t ws[kWavesPerBlock];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kWavesPerBlock; w++) t += ws[w];
        sums[blockIdx.x] = t;
    }
}

// Block-wide exclusive scan of one uint32 per thread; returns the prefix and
// writes the block total to *total.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *total) {
    __shared__ uint32_t ws[kWavesPerBlock];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        uint32_t y = __shfl_up(x, d, kWave);
        if (lane >= d) x += y;
    }
    if (lane == 63) ws[wave] = x;
    __syncthreads();
    uint32_t woff = 0, tot = 0;
    for (int w = 0; w < kWavesPerBlock; w++) {
        if (w < wave) woff += ws[w];
        tot += ws[w];
    }
    __syncthreads();
    *total = tot;
    return woff + x - v;
}

// Phase 2: in-place exclusive scan of the chunk sums by one workgroup.
__global__ __launch_bounds__(kBlock) void k_scan_sums_single(uint32_t *__restrict__ sums, int64_t nb) {
    uint32_t carry = 0;
    for (int64_t b0 = 0; b0 < nb; b0 += kBlock) {
        int64_t i = b0 + threadIdx.x;
        uint32_t v = i < nb ? sums[i] : 0;
        uint32_t tot;
        uint32_t ex = block_exclusive_scan(v, &tot);
        if (i < nb) sums[i] = carry + ex;
        carry += tot;
    }
}

// Phase 3: exclusive scan inside each chunk plus the chunk's offset.  Each
// thread owns kSortItems consecutive elements.  Writes out[n] = total when
// write_total is set (out must then hold n + 1 entries).
__global__ __launch_bounds__(kBlock) void k_scan_apply(const uint32_t *__restrict__ in, int64_t n,
                                                       const uint32_t *__restrict__ sums,
                                                       uint32_t *__restrict__ out, int write_total) {
    const int64_t base = (int64_t)blockIdx.x * kSortTile + (int64_t)threadIdx.x * kSortItems;
    uint32_t v[kSortItems];
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        int64_t i = base + j;
        v[j] = i < n ? in[i] : 0;
        s += v[j];
    }
    uint32_t tot;
    uint32_t run = sums[blockIdx.x] + block_exclusive_scan(s, &tot);
#pragma unroll
    for (int j = 0; j < kSortItems; j++) {
        int64_t i = base + j;
        if (i < n) out[i] = run;
        run += v[j];
    }
    if (write_total && blockIdx.x == gridDim.x - 1 && threadIdx.x == kBlock - 1) out[n] = run;
}

// ---- CSR products of a sorted key array ----------------------------------

// offsets[k] = first position whose key is >= k, for k in 0..nlev (offsets[nlev] = n).
__global__ void k_offsets_from_sorted(const int32_t *__restrict__ skeys, int64_t n, int32_t nlev,
                                      int32_t *__restrict__ offsets) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j > n) return;
    const int32_t prev = j == 0 ? -1 : skeys[j - 1];
    const int32_t cur = j == n ? nlev : skeys[j];
    for (int32_t k = prev + 1; k <= cur; k++) offsets[k] = (int32_t)j;
}

// Head bits: bit j of word j/64 is set when position j starts a segment
// (j == 0 or key changes) or lies beyond n (sentinel, so the last segment
// ends at n-1).  One wave per word; nwords covers n plus one spare word.
__global__ __launch_bounds__(kBlock) void k_head_bits(const int32_t *__restrict__ skeys, int64_t n,
                                                      uint64_t *__restrict__ heads, int64_t nwords,
                                                      uint32_t *__restrict__ wcount) {
    const int64_t w = (int64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (w >= nwords) return;
    const int64_t j = w * 64 + lane;
    bool h;
    if (j >= n) h = true;
    else if (j == 0) h = true;
    else h = skeys[j] != skeys[j - 1];
    const uint64_t word = __ballot(h);
    if (lane == 0) {
        heads[w] = word;
        // heads inside [0, n) only, for the ordinal prefix
        uint64_t real = word;
        const int64_t lo = w * 64;
        if (lo >= n) real = 0;
        else if (n - lo < 64) real &= (~0ull >> (64 - (n - lo)));
        wcount[w] = (uint32_t)__popcll(real);
    }
}

// seg_level[ordinal] = level of each non-empty segment, ordinal counted in
// position order (ordinal = heads before j, using the word prefix).
__global__ void k_seg_level(const int32_t *__restrict__ skeys, int64_t n,
                            const uint64_t *__restrict__ heads, const uint32_t *__restrict__ wprefix,
                            int32_t *__restrict__ seg_level) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint64_t word = heads[j >> 6];
    const int b = (int)(j & 63);
    if (!((word >> b) & 1)) return;
    const uint32_t ord = wprefix[j >> 6] + (uint32_t)__popcll(word & lanes_le(b)) - 1u;
    seg_level[ord] = skeys[j];
}

// ---- pieces: runs of one (FE s level, FE t level) pair inside one tile ----
//
// The internal order sorts observations by FE s, then FE t, so each (s, t)
// cell is one run of positions.  A piece is such a run cut at the starts of
// the fused pass's tiles (every R positions), so no piece crosses a tile.
// Bit j of the piece head words is set when position j starts a piece (or
// lies at or beyond n, as sentinel); wcount counts the heads inside [0, n).
__global__ __launch_bounds__(kBlock) void k_piece_heads(const int32_t *__restrict__ ref_s,
                                                        const int32_t *__restrict__ ref_t, int64_t n, int64_t R,
                                                        uint64_t *__restrict__ heads, int64_t nwords,
                                                        uint32_t *__restrict__ wcount) {
    const int64_t w = (int64_t)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (w >= nwords) return;
    const int64_t j = w * 64 + lane;
    bool h;
    if (j >= n) h = true;
    else if (j == 0 || j % R == 0) h = true;
    else h = ref_s[j] != ref_s[j - 1] || ref_t[j] != ref_t[j - 1];
    const uint64_t word = __ballot(h);
    if (lane == 0) {
        heads[w] = word;
        uint64_t real = word;
        const int64_t lo = w * 64;
        if (lo >= n) real = 0;
        else if (n - lo < 64) real &= (~0ull >> (64 - (n - lo)));
        wcount[w] = (uint32_t)__popcll(real);
    }
}

// piece_s[ordinal], piece_t[ordinal]: the FE s and FE t levels of each piece
// (ordinal = piece heads before its first position)
__global__ void k_piece_levels(const int32_t *__restrict__ ref_s, const int32_t *__restrict__ ref_t, int64_t n,
                               const uint64_t *__restrict__ heads, const uint32_t *__restrict__ wprefix,
                               int32_t *__restrict__ piece_s, int32_t *__restrict__ piece_t) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint64_t word = heads[j >> 6];
    const int b = (int)(j & 63);
    if (!((word >> b) & 1)) return;
    const uint32_t ord = wprefix[j >> 6] + (uint32_t)__popcll(word & lanes_le(b)) - 1u;
    piece_s[ord] = ref_s[j];
    piece_t[ord] = ref_t[j];
}

// ---- internal numbering of the FE with most levels ------------------------
//
// Runs of one (s, t) pair in the order sorted by s, then t: run_s, run_t and
// the first position of each (ordinal from the head words, as pieces).
__global__ void k_run_starts(const int32_t *__restrict__ ss, const int32_t *__restrict__ tt, int64_t n,
                             const uint64_t *__restrict__ heads, const uint32_t *__restrict__ wprefix,
                             int32_t *__restrict__ run_s, int32_t *__restrict__ run_t,
                             int32_t *__restrict__ run_start) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const uint64_t word = heads[j >> 6];
    const int b = (int)(j & 63);
    if (!((word >> b) & 1)) return;
    const uint32_t ord = wprefix[j >> 6] + (uint32_t)__popcll(word & lanes_le(b)) - 1u;
    run_s[ord] = ss[j];
    run_t[ord] = tt[j];
    run_start[ord] = (int32_t)j;
}

__global__ void k_set_i32(int32_t *p, int32_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// main[k] = the t level of the longest run of s level k (runs of a level are
// in ascending t order, so ties go to the smaller t); nt when the level has
// no observation
__global__ void k_main_level(const int32_t *__restrict__ roff, const int32_t *__restrict__ run_start,
                             const int32_t *__restrict__ run_t, int64_t ns, int32_t nt, int32_t *__restrict__ main) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ns) return;
    int32_t best = nt, blen = 0;
    for (int32_t r = roff[k]; r < roff[k + 1]; r++) {
        const int32_t len = run_start[r + 1] - run_start[r];
        if (len > blen) {
            blen = len;
            best = run_t[r];
        }
    }
    main[k] = best;
}

// inv[perm[j]] = j
__global__ void k_invert_perm(const int32_t *__restrict__ perm, int64_t n, int32_t *__restrict__ inv) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) inv[perm[j]] = (int32_t)j;
}

// refs[i] = map[refs[i]] (in place)
__global__ void k_relabel(int32_t *__restrict__ refs, int64_t n, const int32_t *__restrict__ map) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) refs[i] = map[refs[i]];
}

// ---- reference refs -> device int32 (0-based) with range check -----------

template <typename R>
__global__ void k_refs_to_i32(const R *__restrict__ refs, int64_t n, int64_t nlev,
                              int32_t *__restrict__ out, int32_t *__restrict__ bad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t r = (int64_t)refs[i];
    if (r < 1 || r > nlev) {
        atomicOr(bad, 1);
        out[i] = 0;
    } else {
        out[i] = (int32_t)(r - 1);
    }
}

// out[j] = in[perm[j]]  (gather through a permutation)
template <typename T>
__global__ void k_gather_perm(const T *__restrict__ in, const int32_t *__restrict__ perm, int64_t n,
                              T *__restrict__ out) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) out[j] = in[perm[j]];
}

// out[perm[j]] = in[j]  (scatter through a permutation)
template <typename T>
__global__ void k_scatter_perm(const T *__restrict__ in, const int32_t *__restrict__ perm, int64_t n,
                               T *__restrict__ out) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) out[perm[j]] = in[j];
}

}  // namespace fehip
