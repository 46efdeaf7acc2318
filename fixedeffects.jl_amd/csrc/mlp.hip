// mlp.hip -- host side of the warm-start network (fehip_mlp_* in include/fehip.h).
// All arithmetic runs in the MFMA / elementwise kernels of mlp.h.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/fehip.h"
#include "mlp.h"

using namespace fehip;

namespace fehip {
int set_error(int code, const std::string &msg);  // fehip.hip
}

#define MCHK(expr)                                                                                 \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return set_error(FEHIP_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

constexpr int kL = 4;
const int kDin[kL] = {kMlpIn, kMlpW, kMlpW, kMlpW};
const int kDout[kL] = {kMlpW, kMlpW, kMlpW, kMlpOut};

inline unsigned blocks(int64_t n, int per) { return (unsigned)std::max<int64_t>(1, (n + per - 1) / per); }

}  // namespace

struct fehip_mlp {
    hipStream_t stream = nullptr;
    // parameters: W_l [Din x Dout] then b_l [Dout], FP32, one flat buffer
    int64_t off_w[kL], off_b[kL], nparam = 0;
    float *param = nullptr, *grad = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    uint16_t *Wb[kL] = {}, *WTb[kL] = {};
    int64_t step = 0;
    // batch buffers (capacity rows)
    int64_t cap = 0;
    uint16_t *H[kL + 1] = {};  // bf16 activations, H[0] = X
    float *Z[kL] = {};         // FP32 pre-activations
    uint16_t *dZ = nullptr, *dZ2 = nullptr, *ones = nullptr;
    float *dH = nullptr, *t = nullptr, *sq = nullptr, *part = nullptr, *xin = nullptr, *bg = nullptr;
    int64_t part_cap = 0;
    std::vector<void *> allocs;
};

static int mlp_alloc(fehip_mlp *m, void **p, size_t bytes) {
    hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 16));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return set_error(FEHIP_EOOM, "hipMalloc failed for the network");
    }
    m->allocs.push_back(*p);
    return 0;
}

static void mlp_free(fehip_mlp *m, void *p) {
    if (!p) return;
    auto it = std::find(m->allocs.begin(), m->allocs.end(), p);
    if (it != m->allocs.end()) m->allocs.erase(it);
    (void)hipFree(p);
}

static int ensure_parts(fehip_mlp *m, int64_t floats) {
    if (floats <= m->part_cap) return 0;
    (void)hipStreamSynchronize(m->stream);  // no kernel may still use the old block
    mlp_free(m, m->part);
    m->part = nullptr;
    int rc = mlp_alloc(m, (void **)&m->part, sizeof(float) * floats);
    if (rc) return rc;
    m->part_cap = floats;
    return 0;
}

static int ensure_batch(fehip_mlp *m, int64_t rows) {
    if (rows <= m->cap) return 0;
    (void)hipStreamSynchronize(m->stream);
    for (int l = 0; l <= kL; l++) mlp_free(m, m->H[l]);
    for (int l = 0; l < kL; l++) mlp_free(m, m->Z[l]);
    void *q[] = {m->dZ, m->dZ2, m->ones, m->dH, m->t, m->sq, m->xin, m->bg};
    for (void *p : q) mlp_free(m, p);
    int rc;
    for (int l = 0; l <= kL; l++) {
        const int w = l == 0 ? kMlpIn : (l == kL ? kMlpOut : kMlpW);
        if ((rc = mlp_alloc(m, (void **)&m->H[l], sizeof(uint16_t) * rows * w))) return rc;
    }
    for (int l = 0; l < kL; l++)
        if ((rc = mlp_alloc(m, (void **)&m->Z[l], sizeof(float) * rows * kDout[l]))) return rc;
    if ((rc = mlp_alloc(m, (void **)&m->dZ, sizeof(uint16_t) * rows * kMlpW))) return rc;
    if ((rc = mlp_alloc(m, (void **)&m->dZ2, sizeof(uint16_t) * rows * kMlpW))) return rc;
    if ((rc = mlp_alloc(m, (void **)&m->dH, sizeof(float) * rows * kMlpW))) return rc;
    if ((rc = mlp_alloc(m, (void **)&m->t, sizeof(float) * rows))) return rc;
    if ((rc = mlp_alloc(m, (void **)&m->sq, sizeof(float) * rows))) return rc;
    if ((rc = mlp_alloc(m, (void **)&m->xin, sizeof(float) * rows * kMlpIn))) return rc;
    if ((rc = mlp_alloc(m, (void **)&m->bg, sizeof(float) * 16 * kMlpW))) return rc;
    // partial products of the transposed GEMMs: Kd x N <= 128 x 128 per chunk of 1024 rows
    if ((rc = ensure_parts(m, (int64_t)kMlpW * kMlpW * ((rows + 1023) / 1024)))) return rc;
    // ones column (bias gradients as 1^T dZ through the same MFMA product)
    if ((rc = mlp_alloc(m, (void **)&m->ones, sizeof(uint16_t) * rows * 16))) return rc;
    std::vector<uint16_t> h((size_t)rows * 16, 0);
    for (int64_t r = 0; r < rows; r++) h[(size_t)r * 16] = 0x3f80;  // bf16(1.0)
    MCHK(hipMemcpyAsync(m->ones, h.data(), sizeof(uint16_t) * h.size(), hipMemcpyHostToDevice, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    m->cap = rows;
    return 0;
}

static void round_weights(fehip_mlp *m) {
    for (int l = 0; l < kL; l++) {
        const int n = kDin[l] * kDout[l];
        hipLaunchKernelGGL(k_mlp_round_w, dim3(blocks(n, 256)), dim3(256), 0, m->stream, m->param + m->off_w[l],
                           kDin[l], kDout[l], m->Wb[l], m->WTb[l]);
    }
}

// forward over `rows` rows of H[0] (rows multiple of 16)
static void forward(fehip_mlp *m, int rows, bool keep_z) {
    for (int l = 0; l < kL; l++) {
        const bool last = l == kL - 1;
        float *zf = (keep_z || last) ? m->Z[l] : nullptr;
        const unsigned g = blocks(rows / 16, kWavesPerBlock);
        if (kDout[l] == kMlpW)
            hipLaunchKernelGGL(k_mlp_gemm_rows<kMlpW>, dim3(g), dim3(kBlock), 0, m->stream, m->H[l], rows, kDin[l],
                               m->Wb[l], m->param + m->off_b[l], last ? 0 : 1, zf, m->H[l + 1]);
        else
            hipLaunchKernelGGL(k_mlp_gemm_rows<kMlpOut>, dim3(g), dim3(kBlock), 0, m->stream, m->H[l], rows,
                               kDin[l], m->Wb[l], m->param + m->off_b[l], 0, zf, m->H[l + 1]);
    }
}

// A^T D summed over all rows (chunks of 1024 rows, summed in chunk order)
static int gemm_tn_full(fehip_mlp *m, const uint16_t *A, int rows, int Kd, const uint16_t *D, int N, float *out,
                        float scale) {
    const int chunk = 1024;
    const int nchunk = (rows + chunk - 1) / chunk;
    const int64_t size = (int64_t)Kd * N;
    int rc = ensure_parts(m, size * nchunk);
    if (rc) return rc;
    const int tiles = (Kd / 16) * (N / 16);
    hipLaunchKernelGGL(k_mlp_gemm_tn, dim3(blocks(tiles, kWavesPerBlock), nchunk), dim3(kBlock), 0, m->stream, A,
                       rows, Kd, D, N, chunk, m->part);
    hipLaunchKernelGGL(k_mlp_sum_parts, dim3(blocks(size, 256)), dim3(256), 0, m->stream, m->part, nchunk, size,
                       scale, out);
    return 0;
}

// backward of the mean-squared loss on `rows` rows (nvalid real): grads in m->grad
static int backward(fehip_mlp *m, int rows, int nvalid) {
    int rc;
    hipLaunchKernelGGL(k_mlp_loss_grad, dim3(blocks(rows, 256)), dim3(256), 0, m->stream, m->Z[kL - 1], m->t, rows,
                       nvalid, m->dZ, m->sq);
    // dZ holds dZ_l (width kDout[l]); walk layers down
    uint16_t *dz = m->dZ, *dz_next = m->dZ2;
    for (int l = kL - 1; l >= 0; l--) {
        if ((rc = gemm_tn_full(m, m->H[l], rows, kDin[l], dz, kDout[l], m->grad + m->off_w[l], 1.f))) return rc;
        // bias: 1^T dZ = row 0 of ones^T dZ
        if ((rc = gemm_tn_full(m, m->ones, rows, 16, dz, kDout[l], m->bg, 1.f))) return rc;
        MCHK(hipMemcpyAsync(m->grad + m->off_b[l], m->bg, sizeof(float) * kDout[l], hipMemcpyDeviceToDevice,
                            m->stream));
        if (l > 0) {
            // dH_l = dZ_l W_l^T (every kDin[l] here is the hidden width), then
            // dZ_{l-1} = dH_l * [Z_{l-1} > 0]
            const unsigned g = blocks(rows / 16, kWavesPerBlock);
            hipLaunchKernelGGL(k_mlp_gemm_rows<kMlpW>, dim3(g), dim3(kBlock), 0, m->stream, dz, rows, kDout[l],
                               m->WTb[l], (const float *)nullptr, 0, m->dH, (uint16_t *)nullptr);
            const int64_t size = (int64_t)rows * kDin[l];
            hipLaunchKernelGGL(k_mlp_relu_bwd, dim3(blocks(size, 256)), dim3(256), 0, m->stream, m->dH, m->Z[l - 1],
                               size, dz_next);
            std::swap(dz, dz_next);
        }
    }
    MCHK(hipGetLastError());
    return 0;
}

static int load_batch(fehip_mlp *m, const float *X, const float *t, int64_t n, int rows) {
    // X: n x kMlpIn (host FP32), rows >= n padded with zeros
    MCHK(hipMemsetAsync(m->xin, 0, sizeof(float) * rows * kMlpIn, m->stream));
    MCHK(hipMemcpyAsync(m->xin, X, sizeof(float) * n * kMlpIn, hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(k_mlp_to_bf16, dim3(blocks((int64_t)rows * kMlpIn, 256)), dim3(256), 0, m->stream, m->xin,
                       (int64_t)rows * kMlpIn, m->H[0]);
    if (t) {
        MCHK(hipMemsetAsync(m->t, 0, sizeof(float) * rows, m->stream));
        MCHK(hipMemcpyAsync(m->t, t, sizeof(float) * n, hipMemcpyHostToDevice, m->stream));
    }
    return 0;
}

namespace fehip {
// y_dev[r] = network(X_dev[r]) for n rows of FP32 features already on the
// device (used by the learned-start solve; runs on the network's stream and
// synchronizes it)
int mlp_predict_device(fehip_mlp *m, const float *X_dev, int64_t n, float *y_dev) {
    const int64_t chunk = 1 << 16;
    int rc;
    if ((rc = ensure_batch(m, chunk))) return rc;
    round_weights(m);
    for (int64_t s0 = 0; s0 < n; s0 += chunk) {
        const int nb = (int)std::min<int64_t>(chunk, n - s0);
        const int rb = (nb + 63) / 64 * 64;
        MCHK(hipMemsetAsync(m->xin, 0, sizeof(float) * rb * kMlpIn, m->stream));
        MCHK(hipMemcpyAsync(m->xin, X_dev + s0 * kMlpIn, sizeof(float) * nb * kMlpIn, hipMemcpyDeviceToDevice,
                            m->stream));
        hipLaunchKernelGGL(k_mlp_to_bf16, dim3(blocks((int64_t)rb * kMlpIn, 256)), dim3(256), 0, m->stream, m->xin,
                           (int64_t)rb * kMlpIn, m->H[0]);
        forward(m, rb, false);
        MCHK(hipMemcpy2DAsync(y_dev + s0, sizeof(float), m->Z[kL - 1], sizeof(float) * kMlpOut, sizeof(float), nb,
                              hipMemcpyDeviceToDevice, m->stream));
    }
    MCHK(hipStreamSynchronize(m->stream));
    return 0;
}
}  // namespace fehip

extern "C" {

int fehip_mlp_create(fehip_mlp_t **out, uint64_t seed) {
    if (!out) return set_error(FEHIP_EARG, "out is NULL");
    fehip_mlp *m = new fehip_mlp();
    int rc = 0;
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
        delete m;
        return set_error(FEHIP_EHIP, "hipStreamCreate failed");
    }
    int64_t o = 0;
    for (int l = 0; l < kL; l++) {
        m->off_w[l] = o;
        o += (int64_t)kDin[l] * kDout[l];
        m->off_b[l] = o;
        o += kDout[l];
    }
    m->nparam = o;
    // He initialization for the ReLU layers, small output layer; only the live
    // output column is non-zero
    std::vector<float> h(o, 0.f);
    std::mt19937_64 rng(seed);
    for (int l = 0; l < kL; l++) {
        std::normal_distribution<float> nd(0.f, std::sqrt(2.f / kDin[l]) * (l == kL - 1 ? 0.1f : 1.f));
        for (int k = 0; k < kDin[l]; k++)
            for (int j = 0; j < kDout[l]; j++)
                if (l < kL - 1 || j == 0) h[m->off_w[l] + (int64_t)k * kDout[l] + j] = nd(rng);
    }
    if ((rc = mlp_alloc(m, (void **)&m->param, sizeof(float) * o)) ||
        (rc = mlp_alloc(m, (void **)&m->grad, sizeof(float) * o)) ||
        (rc = mlp_alloc(m, (void **)&m->adam_m, sizeof(float) * o)) ||
        (rc = mlp_alloc(m, (void **)&m->adam_v, sizeof(float) * o))) {
        fehip_mlp_destroy(m);
        return rc;
    }
    for (int l = 0; l < kL; l++) {
        if ((rc = mlp_alloc(m, (void **)&m->Wb[l], sizeof(uint16_t) * kDin[l] * kDout[l])) ||
            (rc = mlp_alloc(m, (void **)&m->WTb[l], sizeof(uint16_t) * kDin[l] * kDout[l]))) {
            fehip_mlp_destroy(m);
            return rc;
        }
    }
    hipError_t e = hipMemcpyAsync(m->param, h.data(), sizeof(float) * o, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) e = hipMemsetAsync(m->adam_m, 0, sizeof(float) * o, m->stream);
    if (e == hipSuccess) e = hipMemsetAsync(m->adam_v, 0, sizeof(float) * o, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    if (e != hipSuccess) {
        fehip_mlp_destroy(m);
        return set_error(FEHIP_EHIP, hipGetErrorString(e));
    }
    *out = m;
    return 0;
}

int fehip_mlp_nparams(const fehip_mlp_t *m, int64_t *n) {
    if (!m || !n) return set_error(FEHIP_EARG, "bad arguments");
    *n = m->nparam;
    return 0;
}

int fehip_mlp_get_params(fehip_mlp_t *m, float *out) {
    if (!m || !out) return set_error(FEHIP_EARG, "bad arguments");
    MCHK(hipMemcpyAsync(out, m->param, sizeof(float) * m->nparam, hipMemcpyDeviceToHost, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    return 0;
}

int fehip_mlp_set_params(fehip_mlp_t *m, const float *in) {
    if (!m || !in) return set_error(FEHIP_EARG, "bad arguments");
    MCHK(hipMemcpyAsync(m->param, in, sizeof(float) * m->nparam, hipMemcpyHostToDevice, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    return 0;
}

int fehip_mlp_gradient(fehip_mlp_t *m, const float *X, const float *t, int64_t n, float *grad, double *loss) {
    if (!m || !X || !t || n < 1) return set_error(FEHIP_EARG, "bad arguments");
    if (n > (1 << 24)) return set_error(FEHIP_EARG, "batch too large");
    const int rows = (int)((n + 63) / 64 * 64);
    int rc;
    if ((rc = ensure_batch(m, rows))) return rc;
    if ((rc = load_batch(m, X, t, n, rows))) return rc;
    round_weights(m);
    forward(m, rows, true);
    if ((rc = backward(m, rows, (int)n))) return rc;
    if (grad) MCHK(hipMemcpyAsync(grad, m->grad, sizeof(float) * m->nparam, hipMemcpyDeviceToHost, m->stream));
    std::vector<float> sq(n);
    MCHK(hipMemcpyAsync(sq.data(), m->sq, sizeof(float) * n, hipMemcpyDeviceToHost, m->stream));
    MCHK(hipStreamSynchronize(m->stream));
    if (loss) {
        double s = 0;
        for (int64_t r = 0; r < n; r++) s += sq[r];
        *loss = 0.5 * s / (double)n;
    }
    return 0;
}

int fehip_mlp_fit(fehip_mlp_t *m, const float *X, const float *t, int64_t n, int32_t epochs, int32_t batch,
                  float lr, uint64_t seed, double *loss_per_epoch) {
    if (!m || !X || !t || n < 1 || epochs < 1 || batch < 64) return set_error(FEHIP_EARG, "bad arguments");
    const int rows = (batch + 63) / 64 * 64;
    int rc;
    if ((rc = ensure_batch(m, rows))) return rc;
    std::vector<int64_t> idx(n);
    for (int64_t i = 0; i < n; i++) idx[i] = i;
    std::mt19937_64 rng(seed);
    std::vector<float> bx((size_t)rows * kMlpIn), bt(rows), sq(rows);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int ep = 0; ep < epochs; ep++) {
        std::shuffle(idx.begin(), idx.end(), rng);
        double ssum = 0;
        for (int64_t s0 = 0; s0 < n; s0 += batch) {
            const int nb = (int)std::min<int64_t>(batch, n - s0);
            for (int r = 0; r < nb; r++) {
                std::memcpy(&bx[(size_t)r * kMlpIn], X + idx[s0 + r] * kMlpIn, sizeof(float) * kMlpIn);
                bt[r] = t[idx[s0 + r]];
            }
            const int rb = (nb + 63) / 64 * 64;
            if ((rc = load_batch(m, bx.data(), bt.data(), nb, rb))) return rc;
            round_weights(m);
            forward(m, rb, true);
            if ((rc = backward(m, rb, nb))) return rc;
            m->step++;
            const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
            hipLaunchKernelGGL(k_mlp_adam, dim3(blocks(m->nparam, 256)), dim3(256), 0, m->stream, m->param,
                               m->adam_m, m->adam_v, m->grad, m->nparam, lr, b1, b2, eps, c1, c2);
            MCHK(hipMemcpyAsync(sq.data(), m->sq, sizeof(float) * nb, hipMemcpyDeviceToHost, m->stream));
            MCHK(hipStreamSynchronize(m->stream));
            for (int r = 0; r < nb; r++) ssum += sq[r];
        }
        if (loss_per_epoch) loss_per_epoch[ep] = 0.5 * ssum / (double)n;
    }
    return 0;
}

int fehip_mlp_predict(fehip_mlp_t *m, const float *X, int64_t n, float *y) {
    if (!m || !X || !y || n < 1) return set_error(FEHIP_EARG, "bad arguments");
    const int64_t chunk = 1 << 16;
    int rc;
    if ((rc = ensure_batch(m, chunk))) return rc;
    round_weights(m);
    std::vector<float> z((size_t)chunk * kMlpOut);
    for (int64_t s0 = 0; s0 < n; s0 += chunk) {
        const int nb = (int)std::min<int64_t>(chunk, n - s0);
        const int rb = (nb + 63) / 64 * 64;
        if ((rc = load_batch(m, X + s0 * kMlpIn, nullptr, nb, rb))) return rc;
        forward(m, rb, false);
        MCHK(hipMemcpyAsync(z.data(), m->Z[kL - 1], sizeof(float) * nb * kMlpOut, hipMemcpyDeviceToHost, m->stream));
        MCHK(hipStreamSynchronize(m->stream));
        for (int r = 0; r < nb; r++) y[s0 + r] = z[(size_t)r * kMlpOut];
    }
    return 0;
}

void fehip_mlp_destroy(fehip_mlp_t *m) {
    if (!m) return;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (void *p : m->allocs) (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

}  // extern "C"
