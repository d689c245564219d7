// psr_mlp.hip -- learned chemistry: training and inference of the MLP
// surrogate of step_evolve (src/solve.c:170-224) on gfx950.  No counterpart
// exists in the reference.
//
// Training runs three kernels per batch:
//   k_mlp_fwdbwd  one wavefront per 16 samples: forward pass on BF16 MFMA
//                 (FP32 accumulation, pre-activations kept in registers),
//                 the squared-error loss, and the backward pass of the
//                 deltas; layer inputs and deltas go to HBM as BF16,
//                 feature-major (row f holds every sample of the batch);
//   k_mlp_wgrad   the weight gradients dW^T = delta x a^T as MFMA tiles with
//                 the batch as the reduction dimension (bias gradients use a
//                 column of ones), added into an FP32 gradient;
//   k_mlp_adam    Adam on the FP32 master parameters.
// Inference (k_mlp_forward here, k_mlp_update in psr_chem.hip for ticks)
// keeps every weight in registers as MFMA A operands.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <vector>

#include "psr_internal.h"
#include "psr_mlp.h"

struct MlpState {
    float *d_params = nullptr;  // MLP_NPARAM fp32 master copy
    float *d_m = nullptr, *d_v = nullptr, *d_grad = nullptr;
    float *d_norm = nullptr;  // MLP_NORM: mean, 1/sigma, lo, hi
    bool have_params = false, have_norm = false;
    int64_t adam_t = 0;
    // tuples captured from exact chemistry, MLP_ROW floats each
    float *d_cap = nullptr;
    unsigned long long *d_cap_n = nullptr;
    int64_t cap = 0;
    // training buffers
    short *d_act = nullptr;
    int act_bp = 0;
    double *d_loss = nullptr;
    float *d_rows = nullptr;  // uploaded rows (MLP_ROW floats each)
    int64_t rows_cap = 0;
    uint32_t *d_perm = nullptr;
    int64_t perm_cap = 0;
    uint32_t *d_rowof = nullptr;  // capture row per particle
    int64_t rowof_cap = 0;
};

namespace {

constexpr int MLP_BATCH_ALIGN = 256;  // batch positions per wgrad chunk
constexpr int MLP_NTILE = 57;         // dW / db tiles of the four layers

// BF16 activations and deltas of one batch, feature-major [f][bp]
struct MlpBuf {
    short *a0, *a1, *a2, *a3, *d1, *d2, *d3, *d4;
    int bp;
};

MlpBuf mlp_buf(short *base, int bp) {
    MlpBuf b;
    b.bp = bp;
    short *p = base;
    b.a0 = p;
    p += (size_t)MLP_INP * bp;
    b.a1 = p;
    p += (size_t)MLP_HID * bp;
    b.a2 = p;
    p += (size_t)MLP_HID * bp;
    b.a3 = p;
    p += (size_t)MLP_HID * bp;
    b.d1 = p;
    p += (size_t)MLP_HID * bp;
    b.d2 = p;
    p += (size_t)MLP_HID * bp;
    b.d3 = p;
    p += (size_t)MLP_HID * bp;
    b.d4 = p;
    return b;
}
constexpr size_t MLP_BUF_ROWS = MLP_INP + 6 * MLP_HID + 16;

__device__ __forceinline__ float silu_grad(float z) {
    const float s = __builtin_amdgcn_rcpf(1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

// A operand of the backward pass: W as (in x out), lane l holds W[k][o],
// k = 16 mt + l%16, o = 16 kb + 4(l/16) + j
__device__ __forceinline__ mlp_bf16x4 btile(const float *__restrict__ W, int nin, int nout, int mt, int kb,
                                            int lane) {
    const int k = 16 * mt + (lane & 15);
    const int o0 = 16 * kb + 4 * (lane >> 4);
    mlp_bf16x4 t;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int o = o0 + j;
        t[j] = (k < nin && o < nout) ? mlp_bf16(W[k * nout + o]) : (short)0;
    }
    return t;
}

// store a tile in D / B layout (rows 16 t + 4(l/16) + j, column q) as BF16
__device__ __forceinline__ void put_tile(short *buf, int bp, int t, int q, mlp_bf16x4 v, int lane) {
#pragma unroll
    for (int j = 0; j < 4; j++) buf[(size_t)(16 * t + 4 * (lane >> 4) + j) * bp + q] = v[j];
}

__device__ __forceinline__ mlp_bf16x4 to_bf16(mlp_f32x4 v) {
    mlp_bf16x4 h;
#pragma unroll
    for (int j = 0; j < 4; j++) h[j] = mlp_bf16(v[j]);
    return h;
}

__device__ __forceinline__ void load_bias(const float *__restrict__ P, float *s_bias) {
    for (int k = threadIdx.x; k < 3 * MLP_HID + MLP_OUT; k += blockDim.x) {
        const int l = k / MLP_HID, o = k % MLP_HID;
        const int off = l == 0 ? MLP_B1 : l == 1 ? MLP_B2 : l == 2 ? MLP_B3 : MLP_B4;
        s_bias[k] = P[off + o];
    }
}

// forward, loss and deltas of 16 batch positions per wavefront.  Batch
// position q holds row perm[b0 + q] (or b0 + q) when q < nb, else nothing.
// The loss is the mean over the nb samples of the weighted squared errors
// of the four outputs against mlp_target.
__global__ __launch_bounds__(256) void k_mlp_fwdbwd(const float *__restrict__ rows, const uint32_t *__restrict__ perm,
                                                    int64_t b0, int nb, const float *__restrict__ P,
                                                    const float *__restrict__ norm, MlpBuf Bf,
                                                    double *__restrict__ loss) {
    __shared__ float s_bias[3 * MLP_HID + MLP_OUT];
    load_bias(P, s_bias);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int q0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    if (q0 >= Bf.bp) return;
    const int q = q0 + (lane & 15);
    const bool valid = q < nb;
    const int64_t row = valid ? (perm ? (int64_t)perm[b0 + q] : b0 + q) : 0;
    const float *r = rows + row * MLP_ROW;
    const int r0 = 4 * (lane >> 4);
    mlp_bf16x4 x[2];
#pragma unroll
    for (int kb = 0; kb < 2; kb++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int i = 16 * kb + r0 + j;
            const float v = (valid && i < MLP_IN) ? (r[i] - norm[i]) * norm[MLP_IN + i] : 0.0f;
            x[kb][j] = mlp_bf16(v);
        }
    put_tile(Bf.a0, Bf.bp, 0, q, x[0], lane);
    put_tile(Bf.a0, Bf.bp, 1, q, x[1], lane);
    MlpW w;
    mlp_load_w(P, w, lane);
    mlp_f32x4 z1[4], z2[4], z3[4];
    mlp_bf16x4 h1[4], h2[4], h3[4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        mlp_f32x4 acc = mlp_bias(s_bias, MLP_HID, mt, lane);
        acc = mlp_mfma(w.w1[mt][0], x[0], acc);
        acc = mlp_mfma(w.w1[mt][1], x[1], acc);
        z1[mt] = acc;
        h1[mt] = mlp_act(acc);
        put_tile(Bf.a1, Bf.bp, mt, q, h1[mt], lane);
    }
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        mlp_f32x4 acc = mlp_bias(s_bias + MLP_HID, MLP_HID, mt, lane);
#pragma unroll
        for (int kb = 0; kb < 4; kb++) acc = mlp_mfma(w.w2[mt][kb], h1[kb], acc);
        z2[mt] = acc;
        h2[mt] = mlp_act(acc);
        put_tile(Bf.a2, Bf.bp, mt, q, h2[mt], lane);
    }
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        mlp_f32x4 acc = mlp_bias(s_bias + 2 * MLP_HID, MLP_HID, mt, lane);
#pragma unroll
        for (int kb = 0; kb < 4; kb++) acc = mlp_mfma(w.w3[mt][kb], h2[kb], acc);
        z3[mt] = acc;
        h3[mt] = mlp_act(acc);
        put_tile(Bf.a3, Bf.bp, mt, q, h3[mt], lane);
    }
    mlp_f32x4 out = mlp_bias(s_bias + 3 * MLP_HID, MLP_OUT, 0, lane);
#pragma unroll
    for (int kb = 0; kb < 4; kb++) out = mlp_mfma(w.w4[kb], h3[kb], out);
    // loss and output delta: lanes 0..15 hold the four outputs of sample q
    mlp_f32x4 d4 = {0.0f, 0.0f, 0.0f, 0.0f};
    float lsum = 0.0f;
    if (lane < 16 && valid) {
        float t[MLP_OUT], wt[MLP_OUT];
        mlp_target(r, r + MLP_IN, t, wt);
#pragma unroll
        for (int j = 0; j < MLP_OUT; j++) {
            const float d = out[j] - t[j];
            const float w2 = wt[j] * wt[j];
            lsum += w2 * d * d;
            d4[j] = 2.0f * w2 * d / (float)nb;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o, 64);
    if (lane == 0 && lsum != 0.0f) atomicAdd(loss, (double)lsum);
    const mlp_bf16x4 e4 = to_bf16(d4);
    put_tile(Bf.d4, Bf.bp, 0, q, e4, lane);
    // deltas of layer 3, 2, 1
    mlp_bf16x4 e3[4], e2[4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        mlp_f32x4 g = {0.0f, 0.0f, 0.0f, 0.0f};
        g = mlp_mfma(btile(P + MLP_W4, MLP_HID, MLP_OUT, mt, 0, lane), e4, g);
#pragma unroll
        for (int j = 0; j < 4; j++) g[j] *= silu_grad(z3[mt][j]);
        e3[mt] = to_bf16(g);
        put_tile(Bf.d3, Bf.bp, mt, q, e3[mt], lane);
    }
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        mlp_f32x4 g = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kb = 0; kb < 4; kb++) g = mlp_mfma(btile(P + MLP_W3, MLP_HID, MLP_HID, mt, kb, lane), e3[kb], g);
#pragma unroll
        for (int j = 0; j < 4; j++) g[j] *= silu_grad(z2[mt][j]);
        e2[mt] = to_bf16(g);
        put_tile(Bf.d2, Bf.bp, mt, q, e2[mt], lane);
    }
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        mlp_f32x4 g = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kb = 0; kb < 4; kb++) g = mlp_mfma(btile(P + MLP_W2, MLP_HID, MLP_HID, mt, kb, lane), e2[kb], g);
#pragma unroll
        for (int j = 0; j < 4; j++) g[j] *= silu_grad(z1[mt][j]);
        put_tile(Bf.d1, Bf.bp, mt, q, to_bf16(g), lane);
    }
}

// one wavefront per (tile, chunk of 256 batch positions): a 16x16 tile of
// dW^T (out x in), or of the bias column (B = a column of ones), summed over
// the chunk, then added to the FP32 gradient
__global__ __launch_bounds__(256) void k_mlp_wgrad(MlpBuf Bf, float *__restrict__ grad, int nchunk) {
    const int lane = threadIdx.x & 63;
    const int wid = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wid >= MLP_NTILE * nchunk) return;
    int tile = wid % MLP_NTILE;
    const int chunk = wid / MLP_NTILE;
    // tiles per layer: m-tiles x (bias + in-tiles)
    const short *d, *a;
    int nin, nout, woff, boff, nit;
    if (tile < 12) {
        d = Bf.d1, a = Bf.a0, nin = MLP_IN, nout = MLP_HID, woff = MLP_W1, boff = MLP_B1, nit = 2;
    } else if (tile < 32) {
        tile -= 12;
        d = Bf.d2, a = Bf.a1, nin = MLP_HID, nout = MLP_HID, woff = MLP_W2, boff = MLP_B2, nit = 4;
    } else if (tile < 52) {
        tile -= 32;
        d = Bf.d3, a = Bf.a2, nin = MLP_HID, nout = MLP_HID, woff = MLP_W3, boff = MLP_B3, nit = 4;
    } else {
        tile -= 52;
        d = Bf.d4, a = Bf.a3, nin = MLP_HID, nout = MLP_OUT, woff = MLP_W4, boff = MLP_B4, nit = 4;
    }
    const int mt = tile / (nit + 1);
    const int it = tile % (nit + 1) - 1;  // -1: the bias column
    const int r0 = 4 * (lane >> 4);
    const size_t arow = (size_t)(16 * mt + (lane & 15)) * Bf.bp;
    const size_t brow = (size_t)(16 * (it < 0 ? 0 : it) + (lane & 15)) * Bf.bp;
    const short one = mlp_bf16(1.0f);
    const mlp_bf16x4 ones = {(lane & 15) == 0 ? one : (short)0, (lane & 15) == 0 ? one : (short)0,
                             (lane & 15) == 0 ? one : (short)0, (lane & 15) == 0 ? one : (short)0};
    mlp_f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const int k0 = chunk * MLP_BATCH_ALIGN;
    for (int kk = k0; kk < k0 + MLP_BATCH_ALIGN; kk += 16) {
        const mlp_bf16x4 A = *(const mlp_bf16x4 *)(d + arow + kk + r0);
        const mlp_bf16x4 B = it < 0 ? ones : *(const mlp_bf16x4 *)(a + brow + kk + r0);
        acc = mlp_mfma(A, B, acc);
    }
    const int i = 16 * (it < 0 ? 0 : it) + (lane & 15);
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int o = 16 * mt + r0 + j;
        if (o >= nout) continue;
        if (it < 0) {
            if ((lane & 15) == 0) atomicAdd(&grad[boff + o], acc[j]);
        } else if (i < nin) {
            atomicAdd(&grad[woff + i * nout + o], acc[j]);
        }
    }
}

__global__ void k_mlp_adam(float *__restrict__ P, float *__restrict__ m, float *__restrict__ v,
                           float *__restrict__ g, int n, float lr, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    P[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
    g[i] = 0.0f;
}

// per-block sum, sum of squares, min and max of each raw feature
__global__ __launch_bounds__(256) void k_mlp_stats(const float *__restrict__ rows, int64_t n,
                                                   double *__restrict__ part) {
    __shared__ double s[4][MLP_IN][4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int f = 0; f < MLP_IN; f++) {
        double sm = 0, sq = 0, lo = 1e300, hi = -1e300;
        for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
            const double v = rows[k * MLP_ROW + f];
            sm += v;
            sq += v * v;
            lo = fmin(lo, v);
            hi = fmax(hi, v);
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            sm += __shfl_xor(sm, o, 64);
            sq += __shfl_xor(sq, o, 64);
            lo = fmin(lo, __shfl_xor(lo, o, 64));
            hi = fmax(hi, __shfl_xor(hi, o, 64));
        }
        if (lane == 0) {
            s[wv][f][0] = sm;
            s[wv][f][1] = sq;
            s[wv][f][2] = lo;
            s[wv][f][3] = hi;
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < MLP_IN * 4; k += blockDim.x) {
        const int f = k / 4, c = k % 4;
        double a = s[0][f][c];
        for (int w2 = 1; w2 < 4; w2++) {
            const double b = s[w2][f][c];
            a = c < 2 ? a + b : c == 2 ? fmin(a, b) : fmax(a, b);
        }
        part[(size_t)blockIdx.x * MLP_IN * 4 + k] = a;
    }
}

// predicted step results (mlp_predict) of n raw feature rows
// (stride MLP_IN); 64 rows per wavefront, features through LDS
__global__ __launch_bounds__(256) void k_mlp_forward(const float *__restrict__ x, int64_t n,
                                                     const float *__restrict__ P, const float *__restrict__ norm,
                                                     float *__restrict__ y) {
    __shared__ float s_bias[3 * MLP_HID + MLP_OUT];
    __shared__ float s_feat[4][64][MLP_IN + 1];
    __shared__ float s_out[4][64][MLP_OUT];
    load_bias(P, s_bias);
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    MlpW w;
    mlp_load_w(P, w, lane);
    for (int64_t base = ((int64_t)blockIdx.x * 4 + wv) * 64; base < n; base += (int64_t)gridDim.x * 256) {
        const int64_t i = base + lane;
        const bool ok = i < n;
        for (int f = 0; f < MLP_IN; f++)
            s_feat[wv][lane][f] = ok ? (x[i * MLP_IN + f] - norm[f]) * norm[MLP_IN + f] : 0.0f;
#pragma unroll
        for (int sub = 0; sub < 4; sub++) {
            mlp_bf16x4 xin[2];
            mlp_inputs(&s_feat[wv][16 * sub][0], MLP_IN + 1, lane, xin);
            const mlp_f32x4 o = mlp_forward16(w, s_bias, xin, lane);
            if (lane < 16)
#pragma unroll
                for (int j = 0; j < MLP_OUT; j++) s_out[wv][16 * sub + lane][j] = o[j];
        }
        if (ok) {
            double pr[MLP_OUT];
            mlp_predict(x + i * MLP_IN, s_out[wv][lane], pr);
#pragma unroll
            for (int j = 0; j < MLP_OUT; j++) y[i * MLP_OUT + j] = (float)pr[j];
        }
    }
}

}  // namespace

// ------------------------------------------------------------------ host
static int mlp_state(psr_ctx *c) {
    if (c->mlp) return PSR_OK;
    MlpState *m = new MlpState();
    c->mlp = m;
    PSR_HIP(c, hipMalloc(&m->d_params, sizeof(float) * 4 * MLP_NPARAM + sizeof(float) * MLP_NORM));
    m->d_m = m->d_params + MLP_NPARAM;
    m->d_v = m->d_m + MLP_NPARAM;
    m->d_grad = m->d_v + MLP_NPARAM;
    m->d_norm = m->d_grad + MLP_NPARAM;
    PSR_HIP(c, hipMemset(m->d_params, 0, sizeof(float) * 4 * MLP_NPARAM + sizeof(float) * MLP_NORM));
    PSR_HIP(c, hipMalloc(&m->d_loss, sizeof(double)));
    return PSR_OK;
}

void psr_mlp_free(psr_ctx *c) {
    MlpState *m = c->mlp;
    if (!m) return;
    if (m->d_params) (void)hipFree(m->d_params);
    if (m->d_cap) (void)hipFree(m->d_cap);
    if (m->d_cap_n) (void)hipFree(m->d_cap_n);
    if (m->d_act) (void)hipFree(m->d_act);
    if (m->d_loss) (void)hipFree(m->d_loss);
    if (m->d_rows) (void)hipFree(m->d_rows);
    if (m->d_perm) (void)hipFree(m->d_perm);
    if (m->d_rowof) (void)hipFree(m->d_rowof);
    delete m;
    c->mlp = nullptr;
}

MlpDev psr_mlp_dev(const psr_ctx *c) {
    MlpDev d{};
    if (c->mlp) {
        d.params = c->mlp->d_params;
        d.norm = c->mlp->d_norm;
        d.cap = c->mlp->cap > 0 ? c->mlp->d_cap : nullptr;
        d.cap_n = c->mlp->d_cap_n;
        d.cap_max = (unsigned long long)c->mlp->cap;
    }
    return d;
}

uint32_t *psr_mlp_rowof(psr_ctx *c, int64_t n) {
    MlpState *m = c->mlp;
    if (!m) return nullptr;
    if (m->rowof_cap < n) {
        if (psr_join(c) != PSR_OK || hipStreamSynchronize(c->stream) != hipSuccess) return nullptr;
        if (m->d_rowof) (void)hipFree(m->d_rowof);
        m->d_rowof = nullptr;
        m->rowof_cap = 0;
        if (hipMalloc(&m->d_rowof, sizeof(uint32_t) * n) != hipSuccess) {
            c->err = "psr_mlp: out of device memory for the capture rows";
            return nullptr;
        }
        m->rowof_cap = n;
    }
    return m->d_rowof;
}

bool psr_mlp_ready(const psr_ctx *c) { return c->mlp && c->mlp->have_params && c->mlp->have_norm; }

static int ensure_rows(psr_ctx *c, int64_t n) {
    MlpState *m = c->mlp;
    if (m->rows_cap >= n) return PSR_OK;
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    if (m->d_rows) PSR_HIP(c, hipFree(m->d_rows));
    m->d_rows = nullptr;
    m->rows_cap = 0;
    PSR_HIP(c, hipMalloc(&m->d_rows, sizeof(float) * MLP_ROW * std::max<int64_t>(n, 1)));
    m->rows_cap = n;
    return PSR_OK;
}

static int ensure_act(psr_ctx *c, int bp) {
    MlpState *m = c->mlp;
    if (m->act_bp >= bp) return PSR_OK;
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    if (m->d_act) PSR_HIP(c, hipFree(m->d_act));
    m->d_act = nullptr;
    m->act_bp = 0;
    PSR_HIP(c, hipMalloc(&m->d_act, sizeof(short) * MLP_BUF_ROWS * bp));
    m->act_bp = bp;
    return PSR_OK;
}

// rows (features + targets) from host arrays into the device row buffer
static int upload_rows(psr_ctx *c, const float *x, const float *y, int64_t n) {
    int rc = ensure_rows(c, n);
    if (rc) return rc;
    std::vector<float> h((size_t)n * MLP_ROW);
    for (int64_t k = 0; k < n; k++) {
        std::copy(x + k * MLP_IN, x + (k + 1) * MLP_IN, h.begin() + k * MLP_ROW);
        for (int j = 0; j < MLP_OUT; j++) h[k * MLP_ROW + MLP_IN + j] = y ? y[k * MLP_OUT + j] : 0.0f;
    }
    PSR_HIP(c, hipMemcpy(c->mlp->d_rows, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice));
    return PSR_OK;
}

// mean, 1/sigma and the domain of the raw features of n device rows
static int fit_norm(psr_ctx *c, const float *rows, int64_t n) {
    constexpr int NB = 256;
    double *d_part = nullptr;
    PSR_HIP(c, hipMalloc(&d_part, sizeof(double) * NB * MLP_IN * 4));
    k_mlp_stats<<<NB, 256, 0, c->stream>>>(rows, n, d_part);
    std::vector<double> part((size_t)NB * MLP_IN * 4);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(part.data(), d_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d_part);
    if (e != hipSuccess) PSR_FAIL(c, PSR_ERR_HIP, std::string("psr_mlp_train: ") + hipGetErrorString(e));
    float h[MLP_NORM];
    for (int f = 0; f < MLP_IN; f++) {
        double sm = 0, sq = 0, lo = 1e300, hi = -1e300;
        for (int b = 0; b < NB; b++) {
            const double *p = &part[((size_t)b * MLP_IN + f) * 4];
            sm += p[0];
            sq += p[1];
            lo = std::min(lo, p[2]);
            hi = std::max(hi, p[3]);
        }
        const double mu = sm / n;
        const double sd = std::sqrt(std::max(0.0, sq / n - mu * mu));
        // a feature constant in the data (an unused slot) is passed through
        const double sig = sd > 1e-6 * std::max(1.0, std::fabs(mu)) ? sd : 1.0;
        const double pad = 0.02 * (hi - lo);
        h[f] = (float)mu;
        h[MLP_IN + f] = (float)(1.0 / sig);
        h[2 * MLP_IN + f] = (float)(lo - pad);
        h[3 * MLP_IN + f] = (float)(hi + pad);
    }
    PSR_HIP(c, hipMemcpy(c->mlp->d_norm, h, sizeof(h), hipMemcpyHostToDevice));
    c->mlp->have_norm = true;
    return PSR_OK;
}

static int init_params(psr_ctx *c, uint64_t seed) {
    std::vector<float> p(MLP_NPARAM, 0.0f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int nin, int nout) {
        const double a = std::sqrt(6.0 / (nin + nout));
        std::uniform_real_distribution<double> u(-a, a);
        for (int k = 0; k < nin * nout; k++) p[off + k] = (float)u(rng);
    };
    fill(MLP_W1, MLP_IN, MLP_HID);
    fill(MLP_W2, MLP_HID, MLP_HID);
    fill(MLP_W3, MLP_HID, MLP_HID);
    fill(MLP_W4, MLP_HID, MLP_OUT);
    MlpState *m = c->mlp;
    PSR_HIP(c, hipMemcpy(m->d_params, p.data(), sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice));
    PSR_HIP(c, hipMemset(m->d_m, 0, sizeof(float) * 3 * MLP_NPARAM));  // m, v, grad
    m->adam_t = 0;
    m->have_params = true;
    return PSR_OK;
}

// forward + backward of rows [b0, b0 + nb) (through perm when given) and
// the gradient, added into d_grad
static int batch_grad(psr_ctx *c, const float *rows, const uint32_t *perm, int64_t b0, int nb) {
    MlpState *m = c->mlp;
    const int bp = (nb + MLP_BATCH_ALIGN - 1) / MLP_BATCH_ALIGN * MLP_BATCH_ALIGN;
    int rc = ensure_act(c, bp);
    if (rc) return rc;
    const MlpBuf Bf = mlp_buf(m->d_act, bp);
    k_mlp_fwdbwd<<<(unsigned)(bp / 64), 256, 0, c->stream>>>(rows, perm, b0, nb, m->d_params, m->d_norm, Bf,
                                                             m->d_loss);
    PSR_HIP(c, hipGetLastError());
    const int nchunk = bp / MLP_BATCH_ALIGN;
    k_mlp_wgrad<<<(unsigned)((MLP_NTILE * nchunk + 3) / 4), 256, 0, c->stream>>>(Bf, m->d_grad, nchunk);
    PSR_HIP(c, hipGetLastError());
    return PSR_OK;
}

extern "C" int psr_mlp_features(const psr_step *s, const psr_units *, float *x) {
    if (!s || !x) return PSR_ERR_ARG;
    mlp_features(s->lambdaH, s->lambdaHeI, s->lambdaHeII, s->T, s->nH, s->yGdepHI, s->yGdepHeI, s->yGdepHeII,
                 s->heat, s->yeMET, s->time, x);
    return PSR_OK;
}

extern "C" int psr_mlp_capture(psr_ctx *c, int64_t cap) {
    if (!c || cap < 0) return PSR_ERR_ARG;
    PSR_HIP(c, hipSetDevice(c->device));
    if (int rc_j = psr_join(c)) return rc_j;
    if (int rc = psr_resolve(c)) return rc;
    if (int rc = mlp_state(c)) return rc;
    MlpState *m = c->mlp;
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    if (m->d_cap) PSR_HIP(c, hipFree(m->d_cap));
    m->d_cap = nullptr;
    m->cap = 0;
    if (!m->d_cap_n) PSR_HIP(c, hipMalloc(&m->d_cap_n, sizeof(unsigned long long)));
    PSR_HIP(c, hipMemset(m->d_cap_n, 0, sizeof(unsigned long long)));
    if (cap > 0) {
        PSR_HIP(c, hipMalloc(&m->d_cap, sizeof(float) * MLP_ROW * cap));
        m->cap = cap;
    }
    return PSR_OK;
}

extern "C" int psr_mlp_captured(psr_ctx *c, int64_t *n, float *x, float *y, int64_t max_rows) {
    if (!c || !n) return PSR_ERR_ARG;
    PSR_HIP(c, hipSetDevice(c->device));
    if (int rc_j = psr_join(c)) return rc_j;
    if (int rc = psr_resolve(c)) return rc;
    *n = 0;
    MlpState *m = c->mlp;
    if (!m || !m->d_cap_n) return PSR_OK;
    unsigned long long k = 0;
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    PSR_HIP(c, hipMemcpy(&k, m->d_cap_n, sizeof(k), hipMemcpyDeviceToHost));
    *n = (int64_t)std::min<unsigned long long>(k, (unsigned long long)m->cap);
    if ((x || y) && max_rows > 0) {
        const int64_t nr = std::min(*n, max_rows);
        std::vector<float> h((size_t)nr * MLP_ROW);
        PSR_HIP(c, hipMemcpy(h.data(), m->d_cap, sizeof(float) * h.size(), hipMemcpyDeviceToHost));
        for (int64_t r = 0; r < nr; r++) {
            if (x) std::copy(h.begin() + r * MLP_ROW, h.begin() + r * MLP_ROW + MLP_IN, x + r * MLP_IN);
            if (y)
                std::copy(h.begin() + r * MLP_ROW + MLP_IN, h.begin() + (r + 1) * MLP_ROW, y + r * MLP_OUT);
        }
    }
    return PSR_OK;
}

extern "C" int psr_mlp_train(psr_ctx *c, const float *x, const float *y, int64_t n, int32_t epochs, int32_t batch,
                             float lr, uint64_t seed, float *loss_out) {
    if (!c || epochs < 1 || batch < 16 || !(lr > 0) || (!x) != (!y)) return PSR_ERR_ARG;
    PSR_HIP(c, hipSetDevice(c->device));
    if (int rc_j = psr_join(c)) return rc_j;
    if (int rc = psr_resolve(c)) return rc;
    if (int rc = mlp_state(c)) return rc;
    MlpState *m = c->mlp;
    const float *rows;
    if (x) {
        if (n < 1) return PSR_ERR_ARG;
        if (int rc = upload_rows(c, x, y, n)) return rc;
        rows = m->d_rows;
    } else {
        int64_t have = 0;
        if (int rc = psr_mlp_captured(c, &have, nullptr, nullptr, 0)) return rc;
        n = n > 0 ? std::min(n, have) : have;
        if (n < 1) PSR_FAIL(c, PSR_ERR_STATE, "psr_mlp_train: no captured tuples (psr_mlp_capture, exact ticks)");
        rows = m->d_cap;
    }
    if (!m->have_norm)
        if (int rc = fit_norm(c, rows, n)) return rc;
    if (!m->have_params)
        if (int rc = init_params(c, seed)) return rc;
    if (m->perm_cap < n) {
        if (m->d_perm) PSR_HIP(c, hipFree(m->d_perm));
        m->d_perm = nullptr;
        m->perm_cap = 0;
        PSR_HIP(c, hipMalloc(&m->d_perm, sizeof(uint32_t) * n));
        m->perm_cap = n;
    }
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0u);
    std::mt19937_64 rng(seed ^ 0x9e3779b97f4a7c15ull);
    const int64_t nbatch = (n + batch - 1) / batch;
    const int64_t total = nbatch * epochs;
    int64_t step = 0;
    double last = 0;
    for (int ep = 0; ep < epochs; ep++) {
        std::shuffle(perm.begin(), perm.end(), rng);
        PSR_HIP(c, hipMemcpyAsync(m->d_perm, perm.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, c->stream));
        PSR_HIP(c, hipMemsetAsync(m->d_loss, 0, sizeof(double), c->stream));
        for (int64_t b = 0; b < nbatch; b++, step++) {
            const int nb = (int)std::min<int64_t>(batch, n - b * batch);
            if (int rc = batch_grad(c, rows, m->d_perm, b * batch, nb)) return rc;
            // cosine decay to 5% of lr
            const double frac = (double)step / (double)std::max<int64_t>(total - 1, 1);
            const float lr_t = (float)(lr * (0.05 + 0.95 * 0.5 * (1.0 + std::cos(3.14159265358979323846 * frac))));
            m->adam_t++;
            const float c1 = (float)(1.0 - std::pow(0.9, (double)m->adam_t));
            const float c2 = (float)(1.0 - std::pow(0.999, (double)m->adam_t));
            k_mlp_adam<<<(MLP_NPARAM + 255) / 256, 256, 0, c->stream>>>(m->d_params, m->d_m, m->d_v, m->d_grad,
                                                                        MLP_NPARAM, lr_t, c1, c2);
            PSR_HIP(c, hipGetLastError());
        }
        if (ep == epochs - 1) {
            PSR_HIP(c, hipMemcpyAsync(&last, m->d_loss, sizeof(double), hipMemcpyDeviceToHost, c->stream));
            PSR_HIP(c, hipStreamSynchronize(c->stream));
        }
    }
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    if (loss_out) *loss_out = (float)(last / n);
    return PSR_OK;
}

extern "C" int psr_mlp_grad(psr_ctx *c, const float *x, const float *y, int64_t n, float *grad, float *loss) {
    if (!c || !x || !y || n < 1 || n > (1 << 24) || !grad) return PSR_ERR_ARG;
    PSR_HIP(c, hipSetDevice(c->device));
    if (int rc_j = psr_join(c)) return rc_j;
    if (int rc = psr_resolve(c)) return rc;
    if (!psr_mlp_ready(c)) PSR_FAIL(c, PSR_ERR_STATE, "psr_mlp_grad: set the parameters and normalisation first");
    MlpState *m = c->mlp;
    if (int rc = upload_rows(c, x, y, n)) return rc;
    PSR_HIP(c, hipMemsetAsync(m->d_grad, 0, sizeof(float) * MLP_NPARAM, c->stream));
    PSR_HIP(c, hipMemsetAsync(m->d_loss, 0, sizeof(double), c->stream));
    if (int rc = batch_grad(c, m->d_rows, nullptr, 0, (int)n)) return rc;
    double l = 0;
    PSR_HIP(c, hipMemcpyAsync(grad, m->d_grad, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, c->stream));
    PSR_HIP(c, hipMemcpyAsync(&l, m->d_loss, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PSR_HIP(c, hipMemsetAsync(m->d_grad, 0, sizeof(float) * MLP_NPARAM, c->stream));
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    if (loss) *loss = (float)(l / n);
    return PSR_OK;
}

extern "C" int psr_mlp_forward(psr_ctx *c, const float *x, int64_t n, float *y) {
    if (!c || !x || !y || n < 0) return PSR_ERR_ARG;
    if (n == 0) return PSR_OK;
    PSR_HIP(c, hipSetDevice(c->device));
    if (int rc_j = psr_join(c)) return rc_j;
    if (int rc = psr_resolve(c)) return rc;
    if (!psr_mlp_ready(c)) PSR_FAIL(c, PSR_ERR_STATE, "psr_mlp_forward: no trained surrogate");
    MlpState *m = c->mlp;
    float *dx = nullptr, *dy = nullptr;
    PSR_HIP(c, hipMalloc(&dx, sizeof(float) * MLP_IN * n));
    hipError_t e = hipMalloc(&dy, sizeof(float) * MLP_OUT * n);
    if (e == hipSuccess) e = hipMemcpyAsync(dx, x, sizeof(float) * MLP_IN * n, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        const unsigned grid = (unsigned)std::min<int64_t>((n + 255) / 256, (int64_t)c->cu_count * 8);
        k_mlp_forward<<<grid, 256, 0, c->stream>>>(dx, n, m->d_params, m->d_norm, dy);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(y, dy, sizeof(float) * MLP_OUT * n, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    if (e != hipSuccess) PSR_FAIL(c, PSR_ERR_HIP, std::string("psr_mlp_forward: ") + hipGetErrorString(e));
    return PSR_OK;
}

extern "C" int64_t psr_mlp_param_count(void) { return MLP_NPARAM; }

extern "C" int psr_mlp_get_params(psr_ctx *c, float *p) {
    if (!c || !p) return PSR_ERR_ARG;
    PSR_HIP(c, hipSetDevice(c->device));
    if (int rc_j = psr_join(c)) return rc_j;
    if (int rc = psr_resolve(c)) return rc;
    if (!c->mlp || !c->mlp->have_params) PSR_FAIL(c, PSR_ERR_STATE, "psr_mlp_get_params: no parameters yet");
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    PSR_HIP(c, hipMemcpy(p, c->mlp->d_params, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost));
    return PSR_OK;
}

extern "C" int psr_mlp_set_params(psr_ctx *c, const float *p) {
    if (!c || !p) return PSR_ERR_ARG;
    PSR_HIP(c, hipSetDevice(c->device));
    if (int rc_j = psr_join(c)) return rc_j;
    if (int rc = psr_resolve(c)) return rc;
    if (int rc = mlp_state(c)) return rc;
    MlpState *m = c->mlp;
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    PSR_HIP(c, hipMemcpy(m->d_params, p, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice));
    PSR_HIP(c, hipMemset(m->d_m, 0, sizeof(float) * 3 * MLP_NPARAM));
    m->adam_t = 0;
    m->have_params = true;
    return PSR_OK;
}

extern "C" int psr_mlp_set_norm(psr_ctx *c, const float *mean, const float *sigma, const float *lo,
                                const float *hi) {
    if (!c || !mean || !sigma) return PSR_ERR_ARG;
    for (int f = 0; f < MLP_IN; f++)
        if (!(sigma[f] > 0)) return PSR_ERR_ARG;
    PSR_HIP(c, hipSetDevice(c->device));
    if (int rc_j = psr_join(c)) return rc_j;
    if (int rc = psr_resolve(c)) return rc;
    if (int rc = mlp_state(c)) return rc;
    float h[MLP_NORM];
    for (int f = 0; f < MLP_IN; f++) {
        h[f] = mean[f];
        h[MLP_IN + f] = 1.0f / sigma[f];
        h[2 * MLP_IN + f] = lo ? lo[f] : -INFINITY;
        h[3 * MLP_IN + f] = hi ? hi[f] : INFINITY;
    }
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    PSR_HIP(c, hipMemcpy(c->mlp->d_norm, h, sizeof(h), hipMemcpyHostToDevice));
    c->mlp->have_norm = true;
    return PSR_OK;
}

extern "C" int psr_mlp_get_norm(psr_ctx *c, float *mean, float *sigma, float *lo, float *hi) {
    if (!c) return PSR_ERR_ARG;
    PSR_HIP(c, hipSetDevice(c->device));
    if (int rc_j = psr_join(c)) return rc_j;
    if (int rc = psr_resolve(c)) return rc;
    if (!c->mlp || !c->mlp->have_norm) PSR_FAIL(c, PSR_ERR_STATE, "psr_mlp_get_norm: no normalisation yet");
    float h[MLP_NORM];
    PSR_HIP(c, hipStreamSynchronize(c->stream));
    PSR_HIP(c, hipMemcpy(h, c->mlp->d_norm, sizeof(h), hipMemcpyDeviceToHost));
    for (int f = 0; f < MLP_IN; f++) {
        if (mean) mean[f] = h[f];
        if (sigma) sigma[f] = 1.0f / h[MLP_IN + f];
        if (lo) lo[f] = h[2 * MLP_IN + f];
        if (hi) hi[f] = h[3 * MLP_IN + f];
    }
    return PSR_OK;
}

extern "C" int psr_mlp_set_residual(psr_ctx *c, double tol) {
    if (!c || !(tol > 0)) return PSR_ERR_ARG;
    c->mlp_residual = tol;
    return PSR_OK;
}

extern "C" int psr_mlp_set_balance_check(psr_ctx *c, int32_t on) {
    if (!c) return PSR_ERR_ARG;
    c->mlp_balance = on != 0;
    return PSR_OK;
}
