// grouping.h -- dense level numbering of raw integer keys on the device
// (FixedEffect(keys...), src/FixedEffect.jl:18-21, which calls
// GroupedArray(keys..., sort = nothing)).
//
// One key column: the keys, shifted by their minimum to unsigned, are radix
// sorted with their observation index; a level starts wherever the sorted key
// changes, a scan numbers the levels, and the numbers are scattered back to
// the observations.  Levels are therefore numbered in ascending key order.
// Several columns: each is numbered alone, then the numbers are combined as
// prev * n_next + next (lexicographic order) and the combination is numbered
// again.
#pragma once

#include "fehip_common.h"

namespace fehip {

// mm[0] = min, mm[1] = max of keys (mm preset to INT64_MAX, INT64_MIN).
__global__ __launch_bounds__(kBlock) void k_minmax_i64(const int64_t *__restrict__ keys, int64_t n,
                                                       long long *__restrict__ mm) {
    long long lo = INT64_MAX, hi = INT64_MIN;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const long long v = keys[i];
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
    }
    for (int d = kWave / 2; d >= 1; d >>= 1) {
        const long long a = __shfl_xor(lo, d), b = __shfl_xor(hi, d);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    if ((threadIdx.x & (kWave - 1)) == 0) {
        atomicMin(&mm[0], lo);
        atomicMax(&mm[1], hi);
    }
}

// u[i] = key[i] - lo as an unsigned offset (order preserving).
__global__ void k_key_offset(const int64_t *__restrict__ keys, int64_t n, int64_t lo, uint64_t *__restrict__ u) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u[i] = (uint64_t)keys[i] - (uint64_t)lo;
}

// flag[j] = 1 where the sorted key changes (a level starts).
__global__ void k_group_heads(const uint64_t *__restrict__ skeys, int64_t n, uint32_t *__restrict__ flag) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) flag[j] = (j == 0 || skeys[j] != skeys[j - 1]) ? 1u : 0u;
}

// dense[observation] = 0-based level of sorted position j (ex = exclusive
// scan of flag, so ex + flag - 1 counts the level starts up to j).
__global__ void k_group_assign(const uint32_t *__restrict__ ex, const uint32_t *__restrict__ flag,
                               const int32_t *__restrict__ obs, int64_t n, int64_t *__restrict__ dense) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) dense[obs[j]] = (int64_t)ex[j] + (int64_t)flag[j] - 1;
}

// u[i] = prev[i] * n_next + next[i]: lexicographic combination of two numberings.
__global__ void k_group_combine(const int64_t *__restrict__ prev, const int64_t *__restrict__ next, int64_t n,
                                int64_t n_next, uint64_t *__restrict__ u) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u[i] = (uint64_t)prev[i] * (uint64_t)n_next + (uint64_t)next[i];
}

__global__ void k_plus_one(const int64_t *__restrict__ dense, int64_t n, int64_t *__restrict__ refs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) refs[i] = dense[i] + 1;
}

}  // namespace fehip
