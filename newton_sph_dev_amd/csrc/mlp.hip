// Training and evaluation kernels of the warm-start MLP (BF16 MFMA, FP32 accumulate), and
// the nsph_mlp_* C ABI.  Network and operand layouts: mlp_dev.h.
//
// Gradient pass (one wave = 32 samples per round, two 16-sample tiles):
//   forward     h1, c2 (layer-2 pre-activation), y
//   dy          2 (y - t) / n_total          loss += (y - t)^2
//   dW3[n]   += h2[s][n] dy[s]               db3 += dy[s]
//   dH2[s][n] = dy[s] w3[n] [c2 + b2 > 0]    db2 += dH2 (FP32)
//   dW2      += bf16(h1)^T bf16(dH2)         (MFMA, K = 32 samples)
//   dH1[s][k] = (bf16(dH2) W2^T)[s][k] [h1 > 0]   (MFMA)      db1 += dH1 (FP32)
//   dW1      += bf16(x)^T bf16(dH1)          (MFMA, K = 32 samples)
// Each wave keeps its partial sums in registers and writes them once to a parameter-major
// partial buffer; k_mlp_reduce sums them in a fixed order, so gradients are bitwise
// repeatable.  Adam (FP32) then refreshes the BF16 copies.

#include <math.h>
#include <string.h>

#include <vector>

#include "mlp_dev.h"

namespace nsph {

constexpr int MLP_TRAIN_WAVES = 2048;  // waves of the gradient pass (fixed: fixes the sum order)

// LDS per wave for the gradient pass
struct GradLds {
    float xs[32 * MLP_XPAD];
    __bf16 xsT[16 * 32];     // x^T, rows k (6 real, 10 zero), columns = 32 samples
    __bf16 h1s[32 * MLP_H];  // h1 row-major [s][k]
    __bf16 h1T[MLP_H * 32];  // h1 column-major [k][s]
    __bf16 d2s[32 * MLP_H];  // bf16(dH2) [s][n]
    __bf16 d2T[MLP_H * 32];  // bf16(dH2) [n][s]
    __bf16 d1T[MLP_H * 32];  // bf16(dH1) [k][s]
    float t[32];
};

__device__ inline void wave_sync() {
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(64) void k_mlp_forward(MlpDev M, const float* __restrict__ X, int n,
                                                    float* __restrict__ Y) {
    __shared__ __align__(16) float xs[64 * MLP_XPAD];
    __shared__ __align__(16) __bf16 h1s[16 * MLP_H];
    const int l = threadIdx.x;
    const int base = blockIdx.x * 64;
    MlpRegs R;
    mlp_load_regs(M, R);
    for (int j = 0; j < MLP_XPAD; ++j) {
        const int idx = base + l;
        xs[l * MLP_XPAD + j] = idx < n ? X[(size_t)idx * MLP_XPAD + j] : 0.0f;
    }
    wave_sync();
    const int g = l >> 4, c = l & 15;
    for (int h = 0; h < 4; ++h) {
        float y[4];
        f32x4 c2[4];
        mlp_tile_forward(R, xs + 16 * h * MLP_XPAD, h1s, y, c2);
        if (c < 4) {
            const int idx = base + 16 * h + 4 * g + c;
            const float v = c == 0 ? y[0] : (c == 1 ? y[1] : (c == 2 ? y[2] : y[3]));
            if (idx < n) Y[idx] = v;
        }
        wave_sync();
    }
}

__global__ __launch_bounds__(64) void k_mlp_grad(MlpDev M, const float* __restrict__ X,
                                                 const float* __restrict__ T, int n, float inv_n,
                                                 float* __restrict__ partial, int G) {
    __shared__ __align__(16) GradLds L;
    const int l = threadIdx.x;
    const int g = l >> 4, c = l & 15;
    const int wave = blockIdx.x;
    MlpRegs R;
    mlp_load_regs(M, R);
    // W2^T as B operand: B[nn = 32kk + 8g + j][k = 16t + c] = W2[k][nn]
    bf16x8 bw2t[2][4];
#pragma unroll
    for (int kk = 0; kk < 2; ++kk)
#pragma unroll
        for (int t = 0; t < 4; ++t) bw2t[kk][t] = *reinterpret_cast<const bf16x8*>(M.W2b + (16 * t + c) * MLP_H + 32 * kk + 8 * g);

    f32x4 gw2[4][4];
    f32x4 gw1[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        gw1[a] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int b = 0; b < 4; ++b) gw2[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    float gw3[4] = {0.f, 0.f, 0.f, 0.f}, gb2[4] = {0.f, 0.f, 0.f, 0.f}, gb1[4] = {0.f, 0.f, 0.f, 0.f};
    float gb3 = 0.f, loss = 0.f;

    for (int blk = wave; blk * 32 < n; blk += G) {
        const int base = blk * 32;
        // stage inputs: 32 rows x 8 floats, and x^T (bf16)
        for (int e = l; e < 32 * MLP_XPAD; e += 64) {
            const int sidx = base + e / MLP_XPAD;
            L.xs[e] = sidx < n ? X[(size_t)base * MLP_XPAD + e] : 0.0f;
        }
        if (l < 32) L.t[l] = base + l < n ? T[base + l] : 0.0f;
        wave_sync();
        for (int e = l; e < 16 * 32; e += 64) {
            const int k = e / 32, s2 = e % 32;
            L.xsT[e] = k < MLP_IN ? (__bf16)L.xs[s2 * MLP_XPAD + k] : (__bf16)0.0f;
        }
        // forward both tiles; keep c2 and y
        f32x4 c2[2][4];
        float y[2][4];
        for (int h = 0; h < 2; ++h) {
            mlp_tile_forward(R, L.xs + 16 * h * MLP_XPAD, L.h1s + 16 * h * MLP_H, y[h], c2[h]);
        }
        wave_sync();
        for (int e = l; e < 32 * MLP_H; e += 64) {
            const int s2 = e / MLP_H, k = e % MLP_H;
            L.h1T[k * 32 + s2] = L.h1s[e];
        }
        // output layer gradients
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            float dy[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int s2 = 16 * h + 4 * g + r;
                const bool ok = base + s2 < n;
                const float e = y[h][r] - L.t[s2];
                dy[r] = ok ? 2.0f * e * inv_n : 0.0f;
                if (c == 0 && ok) {
                    loss += e * e;
                    gb3 += dy[r];
                }
            }
#pragma unroll
            for (int t = 0; t < 4; ++t) {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float pre = c2[h][t][r] + R.b2[t];
                    const float h2 = fmaxf(pre, 0.0f);
                    gw3[t] += h2 * dy[r];
                    const float d2 = pre > 0.0f ? dy[r] * R.w3[t] : 0.0f;
                    gb2[t] += d2;
                    const int s2 = 16 * h + 4 * g + r;
                    const __bf16 db = (__bf16)d2;
                    L.d2s[s2 * MLP_H + 16 * t + c] = db;
                    L.d2T[(16 * t + c) * 32 + s2] = db;
                }
            }
        }
        wave_sync();
        // dW2 += h1^T dH2 (K = 32 samples)
#pragma unroll
        for (int mi = 0; mi < 4; ++mi) {
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(L.h1T + (16 * mi + c) * 32 + 8 * g);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const bf16x8 b = *reinterpret_cast<const bf16x8*>(L.d2T + (16 * t + c) * 32 + 8 * g);
                gw2[mi][t] = mfma16(a, b, gw2[mi][t]);
            }
        }
        // dH1 = (dH2 W2^T) masked, per 16-sample tile
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            bf16x8 a2[2];
#pragma unroll
            for (int kk = 0; kk < 2; ++kk)
                a2[kk] = *reinterpret_cast<const bf16x8*>(L.d2s + (16 * h + c) * MLP_H + 32 * kk + 8 * g);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                f32x4 acc = {0.f, 0.f, 0.f, 0.f};
                acc = mfma16(a2[0], bw2t[0][t], acc);
                acc = mfma16(a2[1], bw2t[1][t], acc);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int s2 = 16 * h + 4 * g + r;
                    const float h1 = (float)L.h1s[s2 * MLP_H + 16 * t + c];
                    const float d1 = h1 > 0.0f ? acc[r] : 0.0f;
                    gb1[t] += d1;
                    L.d1T[(16 * t + c) * 32 + s2] = (__bf16)d1;
                }
            }
        }
        wave_sync();
        // dW1 += x^T dH1 (rows k = 0..15, only k < 6 nonzero)
        {
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(L.xsT + c * 32 + 8 * g);
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const bf16x8 b = *reinterpret_cast<const bf16x8*>(L.d1T + (16 * t + c) * 32 + 8 * g);
                gw1[t] = mfma16(a, b, gw1[t]);
            }
        }
        wave_sync();
    }
    // write this wave's partials (parameter-major: partial[p * G + wave])
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int k = 16 * mi + 4 * g + r, nn = 16 * t + c;
                partial[(size_t)(MLP_OFF_W2 + k * MLP_H + nn) * G + wave] = gw2[mi][t][r];
            }
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int k = 4 * g + r, nn = 16 * t + c;
            if (k < MLP_IN) partial[(size_t)(MLP_OFF_W1 + k * MLP_H + nn) * G + wave] = gw1[t][r];
        }
    // column sums over the 4 lane groups (same c), then lanes 0..15 write
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        float a = gw3[t], b = gb2[t], d = gb1[t];
        a += __shfl_xor(a, 16, 64);
        a += __shfl_xor(a, 32, 64);
        b += __shfl_xor(b, 16, 64);
        b += __shfl_xor(b, 32, 64);
        d += __shfl_xor(d, 16, 64);
        d += __shfl_xor(d, 32, 64);
        if (g == 0) {
            partial[(size_t)(MLP_OFF_W3 + 16 * t + c) * G + wave] = a;
            partial[(size_t)(MLP_OFF_B2 + 16 * t + c) * G + wave] = b;
            partial[(size_t)(MLP_OFF_B1 + 16 * t + c) * G + wave] = d;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        gb3 += __shfl_xor(gb3, off, 64);
        loss += __shfl_xor(loss, off, 64);
    }
    if (l == 0) {
        partial[(size_t)MLP_OFF_B3 * G + wave] = gb3;
        partial[(size_t)MLP_NPARAM * G + wave] = loss;
    }
}

// Sum of the G partials of each gradient entry: one wave per entry; lane l adds partials
// l, l + 64, ... in order (coalesced across the wave), then the wave adds its lanes in a
// fixed butterfly, so the result is bitwise repeatable.
__global__ __launch_bounds__(256) void k_mlp_reduce(const float* __restrict__ partial, int G, float* __restrict__ grad) {
    const int p = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    const int l = (int)(threadIdx.x & 63u);
    if (p > MLP_NPARAM) return;  // whole waves
    const float* row = partial + (size_t)p * G;
    float acc = 0.f;
    for (int w = l; w < G; w += 64) acc += row[w];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o);
    if (l == 0) grad[p] = acc;
}

__global__ void k_mlp_adam(float* __restrict__ P, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ grad, float lr, float b1, float b2, float eps,
                           float c1, float c2, __bf16* __restrict__ W1b, __bf16* __restrict__ W2b,
                           __bf16* __restrict__ w3b) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    const float gr = grad[p];
    const float mm = b1 * m[p] + (1.0f - b1) * gr;
    const float vv = b2 * v[p] + (1.0f - b2) * gr * gr;
    m[p] = mm;
    v[p] = vv;
    const float w = P[p] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    P[p] = w;
    if (p < MLP_OFF_B1) W1b[p - MLP_OFF_W1] = (__bf16)w;
    else if (p >= MLP_OFF_W2 && p < MLP_OFF_B2) W2b[p - MLP_OFF_W2] = (__bf16)w;
    else if (p >= MLP_OFF_W3 && p < MLP_OFF_B3) w3b[p - MLP_OFF_W3] = (__bf16)w;
}

__global__ void k_mlp_refresh(const float* __restrict__ P, __bf16* __restrict__ W1b,
                              __bf16* __restrict__ W2b, __bf16* __restrict__ w3b) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    const float w = P[p];
    if (p < MLP_OFF_B1) W1b[p - MLP_OFF_W1] = (__bf16)w;
    else if (p >= MLP_OFF_W2 && p < MLP_OFF_B2) W2b[p - MLP_OFF_W2] = (__bf16)w;
    else if (p >= MLP_OFF_W3 && p < MLP_OFF_B3) w3b[p - MLP_OFF_W3] = (__bf16)w;
}

// One 16x16x32 product with the documented lane layout (self-check of the MFMA operand
// mapping the MLP kernels rely on).
__global__ __launch_bounds__(64) void k_mfma_selftest(const float* __restrict__ A,
                                                      const float* __restrict__ B,
                                                      float* __restrict__ Cm) {
    const int l = threadIdx.x, g = l >> 4, c = l & 15;
    bf16x8 a, b;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        a[j] = (__bf16)A[c * 32 + 8 * g + j];
        b[j] = (__bf16)B[(8 * g + j) * 16 + c];
    }
    f32x4 z = {0.f, 0.f, 0.f, 0.f};
    const f32x4 d = mfma16(a, b, z);
#pragma unroll
    for (int r = 0; r < 4; ++r) Cm[(4 * g + r) * 16 + c] = d[r];
}

}  // namespace nsph

using namespace nsph;

static MlpDev mlp_dev(const nsph_mlp* m) {
    MlpDev M;
    M.W1b = bp<__bf16>(m->W1b);
    M.W2b = bp<__bf16>(m->W2b);
    M.w3b = bp<__bf16>(m->w3b);
    M.P = bp<float>(m->P);
    return M;
}

MlpDev nsph_mlp_device_view(const nsph_mlp* m) { return mlp_dev(m); }

extern "C" {

int nsph_mlp_num_params(void) { return MLP_NPARAM; }

int nsph_mlp_create(nsph_ctx* ctx, nsph_mlp** out) {
    NSPH_CHECK(ctx && out, NSPH_ERR_INVALID, "nsph_mlp_create: NULL argument");
    nsph_mlp* m = new nsph_mlp();
    m->ctx = ctx;
    const size_t np = MLP_NPARAM + 1;
    int rc = buf_ensure(m->P, np * 4);
    if (rc == NSPH_OK) rc = buf_ensure(m->grad, (np + 1) * 4);
    if (rc == NSPH_OK) rc = buf_ensure(m->m, np * 4);
    if (rc == NSPH_OK) rc = buf_ensure(m->v, np * 4);
    if (rc == NSPH_OK) rc = buf_ensure(m->W1b, MLP_IN * MLP_H * 2);
    if (rc == NSPH_OK) rc = buf_ensure(m->W2b, MLP_H * MLP_H * 2);
    if (rc == NSPH_OK) rc = buf_ensure(m->w3b, MLP_H * 2);
    if (rc == NSPH_OK) rc = buf_ensure(m->partial, (size_t)(MLP_NPARAM + 1) * MLP_TRAIN_WAVES * 4);
    if (rc == NSPH_OK && hipHostMalloc((void**)&m->h_loss, 64, hipHostMallocDefault) != hipSuccess) {
        set_error("nsph_mlp_create: pinned allocation failed");
        rc = NSPH_ERR_HIP;
    }
    if (rc != NSPH_OK) {
        buf_free(m->P); buf_free(m->grad); buf_free(m->m); buf_free(m->v);
        buf_free(m->W1b); buf_free(m->W2b); buf_free(m->w3b); buf_free(m->partial);
        delete m;
        return rc;
    }
    m->partial_waves = MLP_TRAIN_WAVES;
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(k_fill_i32, dim3(32), dim3(256), 0, st, bp<int>(m->P), np, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(32), dim3(256), 0, st, bp<int>(m->m), np, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(32), dim3(256), 0, st, bp<int>(m->v), np, 0);
    hipLaunchKernelGGL(k_mlp_refresh, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, st,
                       bp<float>(m->P), bp<__bf16>(m->W1b), bp<__bf16>(m->W2b), bp<__bf16>(m->w3b));
    NSPH_HIP(hipGetLastError());
    *out = m;
    return NSPH_OK;
}

int nsph_mlp_destroy(nsph_mlp* m) {
    if (!m) return NSPH_OK;
    const hipError_t e = nsph_ctx_drain(m->ctx);
    buf_free(m->P); buf_free(m->grad); buf_free(m->m); buf_free(m->v);
    buf_free(m->W1b); buf_free(m->W2b); buf_free(m->w3b); buf_free(m->partial);
    if (m->h_loss) (void)hipHostFree(m->h_loss);
    delete m;
    NSPH_HIP(e);
    return NSPH_OK;
}

int nsph_mlp_set_params(nsph_mlp* m, const float* h_params) {
    NSPH_CHECK(m && h_params, NSPH_ERR_INVALID, "nsph_mlp_set_params: NULL argument");
    hipStream_t st = m->ctx->stream;
    NSPH_HIP(hipMemcpyAsync(m->P.ptr, h_params, MLP_NPARAM * 4, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_fill_i32, dim3(32), dim3(256), 0, st, bp<int>(m->m), (size_t)MLP_NPARAM, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(32), dim3(256), 0, st, bp<int>(m->v), (size_t)MLP_NPARAM, 0);
    hipLaunchKernelGGL(k_mlp_refresh, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, st,
                       bp<float>(m->P), bp<__bf16>(m->W1b), bp<__bf16>(m->W2b), bp<__bf16>(m->w3b));
    NSPH_HIP(hipGetLastError());
    NSPH_HIP(hipStreamSynchronize(st));
    m->steps = 0;
    return NSPH_OK;
}

int nsph_mlp_get_params(nsph_mlp* m, float* h_params) {
    NSPH_CHECK(m && h_params, NSPH_ERR_INVALID, "nsph_mlp_get_params: NULL argument");
    NSPH_HIP(hipMemcpyAsync(h_params, m->P.ptr, MLP_NPARAM * 4, hipMemcpyDeviceToHost, m->ctx->stream));
    NSPH_HIP(hipStreamSynchronize(m->ctx->stream));
    return NSPH_OK;
}

int nsph_mlp_forward(nsph_mlp* m, const float* d_x, int n, float* d_y) {
    NSPH_CHECK(m && n >= 0 && (n == 0 || (d_x && d_y)), NSPH_ERR_INVALID, "nsph_mlp_forward: bad argument");
    if (n == 0) return NSPH_OK;
    hipLaunchKernelGGL(k_mlp_forward, dim3((n + 63) / 64), dim3(64), 0, m->ctx->stream, mlp_dev(m), d_x, n, d_y);
    NSPH_HIP(hipGetLastError());
    return NSPH_OK;
}

int nsph_mlp_grad(nsph_mlp* m, const float* d_x, const float* d_t, int n, float* h_loss,
                  float* d_grad) {
    NSPH_CHECK(m && n > 0 && d_x && d_t, NSPH_ERR_INVALID, "nsph_mlp_grad: bad argument");
    hipStream_t st = m->ctx->stream;
    const int G = m->partial_waves;
    hipLaunchKernelGGL(k_mlp_grad, dim3(G), dim3(64), 0, st, mlp_dev(m), d_x, d_t, n, 1.0f / (float)n,
                       bp<float>(m->partial), G);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NPARAM + 1 + 3) / 4), dim3(256), 0, st,
                       bp<float>(m->partial), G, bp<float>(m->grad));
    NSPH_HIP(hipGetLastError());
    if (d_grad)
        NSPH_HIP(hipMemcpyAsync(d_grad, m->grad.ptr, (MLP_NPARAM + 1) * 4, hipMemcpyDeviceToDevice, st));
    if (h_loss) {
        NSPH_HIP(hipMemcpyAsync(m->h_loss, bp<float>(m->grad) + MLP_NPARAM, 4, hipMemcpyDeviceToHost, st));
        NSPH_HIP(hipStreamSynchronize(st));
        *h_loss = m->h_loss[0] / (float)n;
    }
    return NSPH_OK;
}

int nsph_mlp_adam(nsph_mlp* m, float lr, float beta1, float beta2, float eps) {
    NSPH_CHECK(m != nullptr, NSPH_ERR_INVALID, "nsph_mlp_adam: NULL handle");
    m->steps += 1;
    const float c1 = 1.0f - powf(beta1, (float)m->steps);
    const float c2 = 1.0f - powf(beta2, (float)m->steps);
    hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, m->ctx->stream,
                       bp<float>(m->P), bp<float>(m->m), bp<float>(m->v), bp<float>(m->grad), lr,
                       beta1, beta2, eps, c1, c2, bp<__bf16>(m->W1b), bp<__bf16>(m->W2b),
                       bp<__bf16>(m->w3b));
    NSPH_HIP(hipGetLastError());
    return NSPH_OK;
}

int nsph_mfma_selftest(nsph_ctx* ctx, const float* h_a, const float* h_b, float* h_c) {
    NSPH_CHECK(ctx && h_a && h_b && h_c, NSPH_ERR_INVALID, "nsph_mfma_selftest: NULL argument");
    DevBuf a, b, c;
    int rc = buf_ensure(a, 16 * 32 * 4);
    if (rc == NSPH_OK) rc = buf_ensure(b, 32 * 16 * 4);
    if (rc == NSPH_OK) rc = buf_ensure(c, 16 * 16 * 4);
    hipStream_t st = ctx->stream;
    if (rc == NSPH_OK) {
        if (hipMemcpyAsync(a.ptr, h_a, 16 * 32 * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(b.ptr, h_b, 32 * 16 * 4, hipMemcpyHostToDevice, st) != hipSuccess) {
            rc = NSPH_ERR_HIP;
        } else {
            hipLaunchKernelGGL(k_mfma_selftest, dim3(1), dim3(64), 0, st, bp<float>(a), bp<float>(b), bp<float>(c));
            if (hipMemcpyAsync(h_c, c.ptr, 16 * 16 * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                rc = NSPH_ERR_HIP;
        }
        if (rc != NSPH_OK) set_error("nsph_mfma_selftest: HIP error");
    }
    buf_free(a);
    buf_free(b);
    buf_free(c);
    return rc;
}

}  // extern "C"
