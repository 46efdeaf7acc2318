// mlp.h -- the warm-start network: a 4-layer perceptron of width 128 on
// per-level features, trained and evaluated on the GPU with bf16 MFMA
// (v_mfma_f32_16x16x16_bf16, FP32 accumulate).
//
// Layers l = 0..3, row-major activations over n rows (fixed-effect levels):
//   Z_l = H_l W_l + b_l,  H_{l+1} = relu(Z_l) for l < 3,  y = Z_3[:, 0]
//   H_0 = X [n x 16] (features padded to 16); W_0 [16 x 128],
//   W_1, W_2 [128 x 128], W_3 [128 x 16] with only column 0 live.
// Master weights and Adam state are FP32; each step rounds them to bf16
// operand copies (and transposed copies for the backward products).
// Activations H_l are kept in bf16 (the next product's operand), Z_l in FP32.
// Every product is one of two MFMA kernels:
//   k_mlp_gemm_rows : C = A B        (forward, and dH = dZ W^T)
//   k_mlp_gemm_tn   : C = A^T D      (dW = H^T dZ, and db = 1^T dZ)
// so gradients come out of the same bf16/FP32 arithmetic as the forward.
#pragma once

#include "fehip_common.h"

namespace fehip {

constexpr int kMlpIn = 16;   // padded feature count
constexpr int kMlpW = 128;   // hidden width
constexpr int kMlpOut = 16;  // output layer width (column 0 is the prediction)

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
    // round to nearest even
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// C[M x N] = A[M x Kd] B[Kd x N] (+ bias), row-major bf16 operands, FP32
// accumulation.  M, Kd multiples of 16; N = 16 or 128.  One wave per 16-row
// strip of C covering all N columns: the A fragment of each 16-deep k step
// feeds N/16 MFMAs.  Operand layout of v_mfma_f32_16x16x16_bf16 (lane l,
// q = 0..3): A[l%16][4*(l/16)+q], B[4*(l/16)+q][l%16] -> D[4*(l/16)+q][l%16].
// Epilogue writes the FP32 value (zf) and/or the bf16 value after an
// optional ReLU (hb).
template <int N>
__global__ __launch_bounds__(kBlock) void k_mlp_gemm_rows(const uint16_t *__restrict__ A, int M, int Kd,
                                                          const uint16_t *__restrict__ B,
                                                          const float *__restrict__ bias, int relu,
                                                          float *__restrict__ zf, uint16_t *__restrict__ hb) {
    const int lane = threadIdx.x & 63;
    const int row0 = (blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6)) * 16;
    if (row0 >= M) return;
    constexpr int NT = N / 16;
    f32x4 acc[NT];
#pragma unroll
    for (int j = 0; j < NT; j++) acc[j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int li = lane & 15, lk = 4 * (lane >> 4);
    for (int k0 = 0; k0 < Kd; k0 += 16) {
        bf16x4 a;
        const uint16_t *ap = A + (int64_t)(row0 + li) * Kd + k0 + lk;
#pragma unroll
        for (int q = 0; q < 4; q++) a[q] = (short)ap[q];
#pragma unroll
        for (int j = 0; j < NT; j++) {
            bf16x4 b;
#pragma unroll
            for (int q = 0; q < 4; q++) b[q] = (short)B[(int64_t)(k0 + lk + q) * N + j * 16 + li];
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[j], 0, 0, 0);
        }
    }
#pragma unroll
    for (int j = 0; j < NT; j++) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int r = row0 + lk + q, col = j * 16 + li;
            float v = acc[j][q] + (bias ? bias[col] : 0.f);
            if (zf) zf[(int64_t)r * N + col] = v;
            if (relu) v = v > 0.f ? v : 0.f;
            if (hb) hb[(int64_t)r * N + col] = f32_to_bf16(v);
        }
    }
}

// part[chunk][Kd x N] = A^T D over the rows of one chunk (A [M x Kd],
// D [M x N], row-major bf16).  One wave per 16 x 16 tile of the product;
// rows accumulate in fixed order.  grid = (tiles/4, nchunks).
__global__ __launch_bounds__(kBlock) void k_mlp_gemm_tn(const uint16_t *__restrict__ A, int M, int Kd,
                                                        const uint16_t *__restrict__ D, int N, int chunk,
                                                        float *__restrict__ part) {
    const int lane = threadIdx.x & 63;
    const int tiles_n = N / 16;
    const int tile = blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
    if (tile >= (Kd / 16) * tiles_n) return;
    const int tk = tile / tiles_n, tn = tile % tiles_n;
    const int li = lane & 15, lk = 4 * (lane >> 4);
    f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
    const int m0 = blockIdx.y * chunk, m1 = m0 + chunk < M ? m0 + chunk : M;
    for (int m = m0; m < m1; m += 16) {
        // A operand = tile of A^T: (i = column of A, k = row of A)
        bf16x4 a, b;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            a[q] = (short)A[(int64_t)(m + lk + q) * Kd + tk * 16 + li];
            b[q] = (short)D[(int64_t)(m + lk + q) * N + tn * 16 + li];
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    float *out = part + (int64_t)blockIdx.y * Kd * N;
#pragma unroll
    for (int q = 0; q < 4; q++) out[(int64_t)(tk * 16 + lk + q) * N + tn * 16 + li] = acc[q];
}

// out[e] = scale * sum_c part[c][e]  (chunk order)
__global__ void k_mlp_sum_parts(const float *__restrict__ part, int nchunk, int64_t size, float scale,
                                float *__restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= size) return;
    float s = 0.f;
    for (int c = 0; c < nchunk; c++) s += part[(int64_t)c * size + e];
    out[e] = s * scale;
}

// Loss 0.5 * mean_{r < nvalid} (y_r - t_r)^2 with y_r = Z3[r][0]:
// dY[r][0] = (y_r - t_r) / nvalid, other columns and padding rows 0 (bf16);
// sq[r] = (y_r - t_r)^2 for the host's loss report.
__global__ void k_mlp_loss_grad(const float *__restrict__ Z3, const float *__restrict__ t, int M, int nvalid,
                                uint16_t *__restrict__ dY, float *__restrict__ sq) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= M) return;
    const float d = r < nvalid ? Z3[(int64_t)r * kMlpOut] - t[r] : 0.f;
    for (int j = 0; j < kMlpOut; j++) dY[(int64_t)r * kMlpOut + j] = j == 0 ? f32_to_bf16(d / (float)nvalid) : 0;
    sq[r] = d * d;
}

// dZ = dH * [Z > 0] in bf16 (backward through ReLU)
__global__ void k_mlp_relu_bwd(const float *__restrict__ dH, const float *__restrict__ Z, int64_t size,
                               uint16_t *__restrict__ dZ) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < size) dZ[e] = f32_to_bf16(Z[e] > 0.f ? dH[e] : 0.f);
}

// FP32 master weights [Kd x N] -> bf16 copy and transposed bf16 copy [N x Kd]
__global__ void k_mlp_round_w(const float *__restrict__ W, int Kd, int N, uint16_t *__restrict__ Wb,
                              uint16_t *__restrict__ WTb) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= Kd * N) return;
    const int k = e / N, j = e % N;
    const uint16_t b = f32_to_bf16(W[e]);
    Wb[e] = b;
    WTb[(int64_t)j * Kd + k] = b;
}

// Adam on FP32 parameters; c1 = 1 - b1^t, c2 = 1 - b2^t
__global__ void k_mlp_adam(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                           const float *__restrict__ g, int64_t size, float lr, float b1, float b2, float eps,
                           float c1, float c2) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= size) return;
    const float gi = g[e];
    const float mi = b1 * m[e] + (1.f - b1) * gi;
    const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    w[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void k_mlp_to_bf16(const float *__restrict__ a, int64_t size, uint16_t *__restrict__ b) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < size) b[e] = f32_to_bf16(a[e]);
}

}  // namespace fehip
