// psr_mlp.h -- the chemistry surrogate's network on gfx950.
//
// A multilayer perceptron 22 -> 64 -> 64 -> 64 -> 4 with SiLU activations
// replaces step_evolve (src/solve.c:170-224) for one particle-step.  BF16
// weights and activations feed v_mfma_f32_16x16x16_bf16; sums accumulate in
// FP32.  No counterpart exists in the reference.
//
// Layers are evaluated transposed, Z^T (out x samples) = W^T (out x in) X^T
// (in x samples), so that a 16x16 output tile as the MFMA leaves it (lane l:
// rows 4(l/16)..+3, column l%16) is, after the activation and the BF16
// conversion, exactly the B operand the next layer needs for that k-block
// (lane l: k = 4(l/16)..+3, column l%16).  Activations never leave registers.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

constexpr int MLP_IN = 22;   // features
constexpr int MLP_INP = 32;  // padded to two 16-wide k-blocks
constexpr int MLP_HID = 64;
constexpr int MLP_OUT = 4;
// fp32 parameters, each weight [in][out] row-major followed by its bias
constexpr int MLP_W1 = 0;
constexpr int MLP_B1 = MLP_W1 + MLP_IN * MLP_HID;
constexpr int MLP_W2 = MLP_B1 + MLP_HID;
constexpr int MLP_B2 = MLP_W2 + MLP_HID * MLP_HID;
constexpr int MLP_W3 = MLP_B2 + MLP_HID;
constexpr int MLP_B3 = MLP_W3 + MLP_HID * MLP_HID;
constexpr int MLP_W4 = MLP_B3 + MLP_HID;
constexpr int MLP_B4 = MLP_W4 + MLP_HID * MLP_OUT;
constexpr int MLP_NPARAM = MLP_B4 + MLP_OUT;
// a training / capture row: 22 raw features, then the 4 targets
constexpr int MLP_ROW = MLP_IN + MLP_OUT;
// normalisation block: mean, 1/sigma, lowest and highest raw value seen in
// training (the surrogate's domain), MLP_IN each
constexpr int MLP_NORM = 4 * MLP_IN;

// log10 of a positive double in single precision: the exponent split off
// first, since internal-unit values (nH) overflow float
__host__ __device__ inline float mlp_lg(double v, double floor) {
    int e;
    const float m = (float)frexp(v > floor ? v : floor, &e);
    return log10f(m) + (float)e * 0.30102999566398120f;
}
// logit of a fraction in single precision, clamped to [1e-6, 1 - 1e-6]
__host__ __device__ inline float mlp_logitf(double x) {
    const float v = (float)(x < 1e-6 ? 1e-6 : (x > 1.0 - 1e-6 ? 1.0 - 1e-6 : x));
    return logf(v / (1.0f - v));
}

// log(x / (1 - x)) of a fraction clamped to [1e-9, 1 - 1e-9]
__host__ __device__ inline double mlp_logit(double x) {
    x = x < 1e-9 ? 1e-9 : (x > 1.0 - 1e-9 ? 1.0 - 1e-9 : x);
    return log(x / (1.0 - x));
}
__host__ __device__ inline double mlp_sigmoid(double z) { return 1.0 / (1.0 + exp(-z)); }

// The 22 inputs of one step, from the fields step_evolve reads
// (src/loop.c:662-711, src/solve.c:42-75):
//   0 log10 nH, 1 x_HI, 2 x_HeI, 3 x_HeII, 4 log10 T,
//   5..20 absorbed photons and the state they act on: 5..7 photo-
//         ionisations per atom this step (yGdepHI, yGdepHeI, yGdepHeII),
//         8 heat, 9 yeMET, 10..12 logit x_HI, x_HeI, x_HeII (clamped to
//         [1e-6, 1 - 1e-6]); 13..20 are zero
//         (the state keeps no per-frequency counts and step_evolve reads
//         the deposits only through 5..8),
//   21 log10 of the step time.
// The targets are x_HI, x_HeI, x_HeII and log10 T after the step; the
// network predicts the change of logit x (so predicted fractions stay in
// (0, 1)) and of log10 T.
__host__ __device__ inline void mlp_features(double lH, double lHeI, double lHeII, double T, double nH,
                                             double gHI, double gHeI, double gHeII, double heat, double yeMET,
                                             double time, float f[MLP_IN]) {
    f[0] = mlp_lg(nH, 1e-30);
    f[1] = (float)lH;
    f[2] = (float)lHeI;
    f[3] = (float)lHeII;
    f[4] = mlp_lg(T, 1e-30);
    f[5] = mlp_lg(gHI + 1e-12, 1e-30);
    f[6] = mlp_lg(gHeI + 1e-12, 1e-30);
    f[7] = mlp_lg(gHeII + 1e-12, 1e-30);
    f[8] = mlp_lg(heat + 1e-30, 1e-30);
    f[9] = (float)yeMET;
    f[10] = mlp_logitf(lH);
    f[11] = mlp_logitf(lHeI);
    f[12] = mlp_logitf(lHeII);
    for (int k = 13; k < 21; k++) f[k] = 0.0f;
    f[21] = mlp_lg(time, 1e-30);
}

// network targets of a (features f, step result y) row and their loss
// weights: the fractions' logit changes weighted by max(x (1 - x), 0.01)
// of the result (so the loss tracks absolute fraction errors away from 0
// and 1), and the change of log10 T with weight 1
__host__ __device__ inline void mlp_target(const float *f, const float *y, float t[MLP_OUT], float w[MLP_OUT]) {
    for (int j = 0; j < 3; j++) {
        const double x = y[j];
        t[j] = (float)(mlp_logit(x) - mlp_logit((double)f[1 + j]));
        const double v = x * (1.0 - x);
        w[j] = (float)(v > 0.01 ? v : 0.01);
    }
    t[3] = y[3] - f[4];
    w[3] = 1.0f;
}

// the step result the network predicts from features f and outputs o
__host__ __device__ inline void mlp_predict(const float *f, const float *o, double y[MLP_OUT]) {
    for (int j = 0; j < 3; j++) y[j] = mlp_sigmoid(mlp_logit((double)f[1 + j]) + (double)o[j]);
    y[3] = (double)f[4] + (double)o[3];
}

typedef short mlp_bf16x4 __attribute__((ext_vector_type(4)));
typedef float mlp_f32x4 __attribute__((ext_vector_type(4)));

// float -> bfloat16, round to nearest even (v_cvt_pk_bf16_f32 on gfx950)
__device__ __forceinline__ short mlp_bf16(float f) { return __builtin_bit_cast(short, (__bf16)f); }

// SiLU z / (1 + e^-z) with the hardware exp and reciprocal (their ulp-level
// error is far below the BF16 rounding that follows)
__device__ __forceinline__ float mlp_silu(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

// weights as MFMA A operands (tiles of W^T), held in registers:
// lane l of tile (mt, kb) holds W[i][o], o = 16 mt + l%16, i = 16 kb + 4(l/16) + j
struct MlpW {
    mlp_bf16x4 w1[4][2], w2[4][4], w3[4][4], w4[4];
};

__device__ __forceinline__ mlp_bf16x4 mlp_wtile(const float *__restrict__ W, int nin, int nout, int mt, int kb,
                                                int lane) {
    const int o = 16 * mt + (lane & 15);
    const int i0 = 16 * kb + 4 * (lane >> 4);
    mlp_bf16x4 t;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int i = i0 + j;
        t[j] = (i < nin && o < nout) ? mlp_bf16(W[i * nout + o]) : (short)0;
    }
    return t;
}

__device__ inline void mlp_load_w(const float *__restrict__ P, MlpW &w, int lane) {
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
#pragma unroll
        for (int kb = 0; kb < 2; kb++) w.w1[mt][kb] = mlp_wtile(P + MLP_W1, MLP_IN, MLP_HID, mt, kb, lane);
#pragma unroll
        for (int kb = 0; kb < 4; kb++) {
            w.w2[mt][kb] = mlp_wtile(P + MLP_W2, MLP_HID, MLP_HID, mt, kb, lane);
            w.w3[mt][kb] = mlp_wtile(P + MLP_W3, MLP_HID, MLP_HID, mt, kb, lane);
        }
        w.w4[mt] = mlp_wtile(P + MLP_W4, MLP_HID, MLP_OUT, 0, mt, lane);
    }
}

__device__ __forceinline__ mlp_f32x4 mlp_mfma(mlp_bf16x4 a, mlp_bf16x4 b, mlp_f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// bias rows of an output tile: lane l gets b[16 mt + 4(l/16) + j]
__device__ __forceinline__ mlp_f32x4 mlp_bias(const float *b, int n, int mt, int lane) {
    mlp_f32x4 c;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int o = 16 * mt + 4 * (lane >> 4) + j;
        c[j] = o < n ? b[o] : 0.0f;
    }
    return c;
}

__device__ __forceinline__ mlp_bf16x4 mlp_act(mlp_f32x4 z) {
    mlp_bf16x4 h;
#pragma unroll
    for (int j = 0; j < 4; j++) h[j] = mlp_bf16(mlp_silu(z[j]));
    return h;
}

// forward pass of 16 samples; x: the two k-blocks of the normalised inputs
// (B operand layout); bias: b1, b2, b3, b4 packed (64, 64, 64, 4).  Returns
// the output tile: lanes 0..15 hold outputs 0..3 of sample l.
__device__ __forceinline__ mlp_f32x4 mlp_forward16(const MlpW &w, const float *bias, const mlp_bf16x4 x[2],
                                                   int lane) {
    mlp_bf16x4 h[4], g[4];
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        mlp_f32x4 acc = mlp_bias(bias, MLP_HID, mt, lane);
        acc = mlp_mfma(w.w1[mt][0], x[0], acc);
        acc = mlp_mfma(w.w1[mt][1], x[1], acc);
        h[mt] = mlp_act(acc);
    }
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        mlp_f32x4 acc = mlp_bias(bias + MLP_HID, MLP_HID, mt, lane);
#pragma unroll
        for (int kb = 0; kb < 4; kb++) acc = mlp_mfma(w.w2[mt][kb], h[kb], acc);
        g[mt] = mlp_act(acc);
    }
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        mlp_f32x4 acc = mlp_bias(bias + 2 * MLP_HID, MLP_HID, mt, lane);
#pragma unroll
        for (int kb = 0; kb < 4; kb++) acc = mlp_mfma(w.w3[mt][kb], g[kb], acc);
        h[mt] = mlp_act(acc);
    }
    mlp_f32x4 acc = mlp_bias(bias + 3 * MLP_HID, MLP_OUT, 0, lane);
#pragma unroll
    for (int kb = 0; kb < 4; kb++) acc = mlp_mfma(w.w4[kb], h[kb], acc);
    return acc;
}

// B operand of 16 samples from normalised features in LDS (row stride ld):
// lane l takes sample l%16, features 16 kb + 4(l/16) + j (zero past MLP_IN)
__device__ __forceinline__ void mlp_inputs(const float *rows, int ld, int lane, mlp_bf16x4 x[2]) {
    const float *r = rows + (lane & 15) * ld;
#pragma unroll
    for (int kb = 0; kb < 2; kb++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int i = 16 * kb + 4 * (lane >> 4) + j;
            x[kb][j] = i < MLP_IN ? mlp_bf16(r[i]) : (short)0;
        }
}
