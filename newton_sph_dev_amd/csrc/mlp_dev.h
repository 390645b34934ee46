// Per-particle MLP 6 -> 64 -> 64 -> 1 (ReLU) for the pressure warm start, gfx950 BF16 MFMA.
//
// Parameters (FP32 master copy, flat, in this order; 4673 values):
//   W1[6][64] (k, n row-major), b1[64], W2[64][64] (k, n), b2[64], w3[64], b3[1]
// The kernels read BF16 copies of W1, W2, w3 and FP32 biases; products accumulate in FP32.
//
//   h1 = relu(bf16(x) W1 + b1)            x padded to 32 inputs (features 6, 7 are 0)
//   h2 = relu(bf16(h1) W2 + b2)
//   y  = sum_n h2[n] * w3[n] + b3         (FP32, w3 rounded to BF16)
//
// One wave evaluates 16 samples per MFMA row tile with V_MFMA_F32_16X16X32_BF16:
//   A (16 x 32):  lane l holds A[l % 16][8 (l / 16) + j], j = 0..7
//   B (32 x 16):  lane l holds B[8 (l / 16) + j][l % 16]
//   C (16 x 16):  lane l holds C[4 (l / 16) + r][l % 16], r = 0..3
// Layout changes between layers go through LDS (row-major [sample][hidden] BF16 tiles).
#pragma once

#include <hip/hip_runtime.h>

#include "common.h"

namespace nsph {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int MLP_IN = 6;
constexpr int MLP_XPAD = 8;  // feature row stride (floats)
constexpr int MLP_H = 64;
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_IN * MLP_H;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_H;
constexpr int MLP_NPARAM = MLP_OFF_B3 + 1;  // 4673

struct MlpDev {
    const __bf16* W1b;  // [6][64]
    const __bf16* W2b;  // [64][64]
    const __bf16* w3b;  // [64]
    const float* P;     // FP32 master (biases are read from here)
};

__device__ inline f32x4 mfma16(const bf16x8& a, const bf16x8& b, const f32x4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// B operands of layer 1 (4 column tiles) and layer 2 (2 K-halves x 4 column tiles).
struct MlpRegs {
    bf16x8 bw1[4];
    bf16x8 bw2[2][4];
    float b1[4], b2[4], w3[4], b3;
};

__device__ inline void mlp_load_regs(const MlpDev& M, MlpRegs& R) {
    const int l = threadIdx.x & 63;
    const int g = l >> 4, c = l & 15;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int n = 16 * t + c;
        bf16x8 v;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = 8 * g + j;
            v[j] = k < MLP_IN ? M.W1b[k * MLP_H + n] : (__bf16)0.0f;
        }
        R.bw1[t] = v;
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
            bf16x8 u;
#pragma unroll
            for (int j = 0; j < 8; ++j) u[j] = M.W2b[(32 * kk + 8 * g + j) * MLP_H + n];
            R.bw2[kk][t] = u;
        }
        R.b1[t] = M.P[MLP_OFF_B1 + n];
        R.b2[t] = M.P[MLP_OFF_B2 + n];
        R.w3[t] = (float)M.w3b[n];
    }
    R.b3 = M.P[MLP_OFF_B3];
}

// Forward of one 16-sample tile. xs: the tile's feature rows in LDS ([16][MLP_XPAD] f32);
// h1s: LDS scratch [16][64] bf16 for this wave. Returns, in every lane of group g = l/16,
// y for samples 4g + r in y[r] and the layer-2 pre-activations of column 16t + l%16 in c2.
__device__ inline void mlp_tile_forward(const MlpRegs& R, const float* xs, __bf16* h1s, float y[4],
                                        f32x4 c2[4]) {
    const int l = threadIdx.x & 63;
    const int g = l >> 4, c = l & 15;
    bf16x8 a;
#pragma unroll
    for (int j = 0; j < 8; ++j) a[j] = (g == 0 && j < MLP_IN) ? (__bf16)xs[c * MLP_XPAD + j] : (__bf16)0.0f;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f32x4 z = {0.f, 0.f, 0.f, 0.f};
        const f32x4 h = mfma16(a, R.bw1[t], z);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float v = fmaxf(h[r] + R.b1[t], 0.0f);
            h1s[(4 * g + r) * MLP_H + 16 * t + c] = (__bf16)v;
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    bf16x8 a2[2];
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) a2[kk] = *reinterpret_cast<const bf16x8*>(h1s + c * MLP_H + 32 * kk + 8 * g);
    float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma16(a2[0], R.bw2[0][t], acc);
        acc = mfma16(a2[1], R.bw2[1][t], acc);
        c2[t] = acc;
#pragma unroll
        for (int r = 0; r < 4; ++r) part[r] += fmaxf(acc[r] + R.b2[t], 0.0f) * R.w3[t];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        float v = part[r];
        v += __shfl_xor(v, 1, 64);
        v += __shfl_xor(v, 2, 64);
        v += __shfl_xor(v, 4, 64);
        v += __shfl_xor(v, 8, 64);
        y[r] = v + R.b3;
    }
    __builtin_amdgcn_wave_barrier();
}

}  // namespace nsph

// Opaque handle of the C ABI (shared by mlp.hip and sph.hip).
struct nsph_mlp {
    nsph_ctx* ctx = nullptr;
    nsph::DevBuf P, W1b, W2b, w3b, grad, m, v, partial;
    int partial_waves = 0;
    int steps = 0;
    float* h_loss = nullptr;  // pinned
};
