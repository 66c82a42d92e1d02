// k_mlp.hip -- the sign-risk proposer: MLP 12 -> 64 -> 64 -> 64 -> 1 (ReLU)
// on BF16 matrix cores (v_mfma_f32_16x16x16_bf16), FP32 accumulation.
//
// Parameters (FP32 master copy, canonical layout, rhz_mlp_* in
// rhizome_sdf.h): W1[12][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64]
// w4[64] b4 | feature mean[12] and 1/std[12].  Every kernel rounds the
// weights and activations it feeds to the matrix cores to BF16 (round to
// nearest even).
//
// One wave handles 16 samples per step.  MFMA 16x16x16 operand layout
// (lane l, element q of the 4 per lane):
//   A[m][k]: m = l % 16, k = 4 (l / 16) + q
//   B[k][n]: n = l % 16, k = 4 (l / 16) + q
//   D[m][n]: n = l % 16, m = 4 (l / 16) + q
// A layer output in D layout (rows = samples) is re-read in A layout through
// a 16 x 64 LDS tile.  For the weight gradients dW = H^T dZ the sample index
// is the reduction axis: H and dZ in D layout are directly the A and B
// operands (m = unit, k = sample), so no transposition is needed there.
#include "rhz_device.hpp"
#include "rhz_internal.hpp"

namespace rhz {

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kIn = 12;
constexpr int kH = 64;
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + kIn * kH;
constexpr int OFF_W2 = OFF_B1 + kH;
constexpr int OFF_B2 = OFF_W2 + kH * kH;
constexpr int OFF_W3 = OFF_B2 + kH;
constexpr int OFF_B3 = OFF_W3 + kH * kH;
constexpr int OFF_W4 = OFF_B3 + kH;
constexpr int OFF_B4 = OFF_W4 + kH;
constexpr int NPARAM = OFF_B4 + 1;  // 9217
constexpr int OFF_MEAN = NPARAM;
constexpr int OFF_ISTD = OFF_MEAN + kIn;
constexpr int NBLOB = OFF_ISTD + kIn;  // 9241

constexpr int kWaves = 4;

__device__ __forceinline__ short f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}
__device__ __forceinline__ float bfr(float f) {  // round to BF16 and back
    return __uint_as_float(((uint32_t)(uint16_t)f2bf(f)) << 16);
}
__device__ __forceinline__ v4f mfma(v4s a, v4s b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4s pack4(float a, float b, float c, float d) {
    v4s r;
    r.x = f2bf(a);
    r.y = f2bf(b);
    r.z = f2bf(c);
    r.w = f2bf(d);
    return r;
}

struct Weights {
    v4s w1[4];      // B operands, layer 1 (k padded to 16)
    v4s w2[4][4];   // [kt][nt]
    v4s w3[4][4];
    float b1[4], b2[4], b3[4], w4[4];  // per unit 16 nt + l % 16
    float b4;
};

__device__ __forceinline__ void load_weights(const float* __restrict__ P, int lane, Weights& W) {
    const int c = lane & 15, g = lane >> 4;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        float v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = 4 * g + q;
            v[q] = (k < kIn) ? P[OFF_W1 + k * kH + 16 * nt + c] : 0.0f;
        }
        W.w1[nt] = pack4(v[0], v[1], v[2], v[3]);
        W.b1[nt] = P[OFF_B1 + 16 * nt + c];
        W.b2[nt] = P[OFF_B2 + 16 * nt + c];
        W.b3[nt] = P[OFF_B3 + 16 * nt + c];
        W.w4[nt] = P[OFF_W4 + 16 * nt + c];
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) {
            float a[4], b[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int k = 16 * kt + 4 * g + q;
                a[q] = P[OFF_W2 + k * kH + 16 * nt + c];
                b[q] = P[OFF_W3 + k * kH + 16 * nt + c];
            }
            W.w2[kt][nt] = pack4(a[0], a[1], a[2], a[3]);
            W.w3[kt][nt] = pack4(b[0], b[1], b[2], b[3]);
        }
    }
    W.b4 = P[OFF_B4];
}

// Transposed weights as B operands of dH = dZ W^T: B[k = out][n = in] = W[in][out].
__device__ __forceinline__ void load_wt(const float* __restrict__ P, int off, int lane, v4s wt[4][4]) {
    const int c = lane & 15, g = lane >> 4;
#pragma unroll
    for (int kt = 0; kt < 4; ++kt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            float a[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int out = 16 * kt + 4 * g + q, in = 16 * nt + c;
                a[q] = P[off + in * kH + out];
            }
            wt[kt][nt] = pack4(a[0], a[1], a[2], a[3]);
        }
}

// Sample tile -> A operand of layer 1 (sample l % 16, features 4 (l / 16) + q).
__device__ __forceinline__ v4s load_x_a(const float* __restrict__ X, const float* __restrict__ P, int64_t n,
                                        int64_t s0, int lane) {
    const int64_t i = s0 + (lane & 15);
    float v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int k = 4 * (lane >> 4) + q;
        v[q] = (i < n && k < kIn) ? (X[12 * i + k] - P[OFF_MEAN + k]) * P[OFF_ISTD + k] : 0.0f;
    }
    return pack4(v[0], v[1], v[2], v[3]);
}

// D-layout activations (rows = samples) -> A operands through LDS.
__device__ __forceinline__ void d_to_a(const v4f h[4], float (*L)[kH + 1], int lane, v4s a[4]) {
    const int c = lane & 15, g = lane >> 4;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int q = 0; q < 4; ++q) L[4 * g + q][16 * nt + c] = h[nt][q];
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): the wave's LDS writes are done
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int kt = 0; kt < 4; ++kt)
        a[kt] = pack4(L[c][16 * kt + 4 * g + 0], L[c][16 * kt + 4 * g + 1], L[c][16 * kt + 4 * g + 2],
                      L[c][16 * kt + 4 * g + 3]);
    __builtin_amdgcn_s_waitcnt(0xc07f);
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ void layer(const v4s a[4], const v4s w[4][4], const float b[4], v4f h[4]) {
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < 4; ++kt) acc = mfma(a[kt], w[kt][nt], acc);
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] = fmaxf(acc[q] + b[nt], 0.0f);
        h[nt] = acc;
    }
}

struct Fwd {
    v4f h1[4], h2[4], h3[4];
    float logit[4];  // samples 4 (l / 16) + q (identical across the 16 lanes of a group)
};

__device__ __forceinline__ void forward(const Weights& W, v4s x, float (*L)[kH + 1], int lane, Fwd& F) {
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        v4f acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma(x, W.w1[nt], acc);
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] = fmaxf(acc[q] + W.b1[nt], 0.0f);
        F.h1[nt] = acc;
    }
    v4s a[4];
    d_to_a(F.h1, L, lane, a);
    layer(a, W.w2, W.b2, F.h2);
    d_to_a(F.h2, L, lane, a);
    layer(a, W.w3, W.b3, F.h3);
    float part[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int q = 0; q < 4; ++q) part[q] += bfr(W.w4[nt]) * bfr(F.h3[nt][q]);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        float s = part[q];
        for (int o = 1; o < 16; o <<= 1) s += __shfl_xor(s, o);
        F.logit[q] = s + W.b4;
    }
}

__global__ __launch_bounds__(64 * kWaves) void mlp_forward_kernel(const float* __restrict__ P,
                                                                  const float* __restrict__ X, int64_t n,
                                                                  float* __restrict__ out) {
    __shared__ float L[kWaves][16][kH + 1];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    Weights W;
    load_weights(P, lane, W);
    const int64_t tiles = (n + 15) / 16;
    for (int64_t t = (int64_t)blockIdx.x * kWaves + wv; t < tiles; t += (int64_t)gridDim.x * kWaves) {
        Fwd F;
        forward(W, load_x_a(X, P, n, t * 16, lane), L[wv], lane, F);
        if ((lane & 15) == 0)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int64_t i = t * 16 + 4 * (lane >> 4) + q;
                if (i < n) out[i] = F.logit[q];
            }
    }
}

// Sum over the 4 lane groups (same l % 16): xor 16 and 32.
__device__ __forceinline__ float group_sum(float v) {
    v += __shfl_xor(v, 16);
    v += __shfl_xor(v, 32);
    return v;
}

// Weighted binary cross-entropy on logits: loss_i = c_i * [softplus(z) - y z]
// with c_i = pos_weight for y = 1, 1 otherwise; dL/dz = c_i (sigmoid(z) - y).
__global__ __launch_bounds__(64 * kWaves) void mlp_grad_kernel(const float* __restrict__ P,
                                                               const float* __restrict__ X,
                                                               const float* __restrict__ Y, int64_t n,
                                                               float pos_weight, float* __restrict__ G,
                                                               float* __restrict__ loss) {
    __shared__ float L[kWaves][16][kH + 1];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int c = lane & 15, g = lane >> 4;
    Weights W;
    load_weights(P, lane, W);
    v4s w3t[4][4], w2t[4][4];
    load_wt(P, OFF_W3, lane, w3t);
    load_wt(P, OFF_W2, lane, w2t);
    v4f gw1[4], gw2[4][4], gw3[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        gw1[a] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            gw2[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
            gw3[a][b] = v4f{0.f, 0.f, 0.f, 0.f};
        }
    }
    float gb1[4] = {0, 0, 0, 0}, gb2[4] = {0, 0, 0, 0}, gb3[4] = {0, 0, 0, 0}, gw4[4] = {0, 0, 0, 0};
    float gb4 = 0.f, lsum = 0.f;
    const int64_t tiles = (n + 15) / 16;
    for (int64_t t = (int64_t)blockIdx.x * kWaves + wv; t < tiles; t += (int64_t)gridDim.x * kWaves) {
        const int64_t s0 = t * 16;
        Fwd F;
        forward(W, load_x_a(X, P, n, s0, lane), L[wv], lane, F);
        float dz[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int64_t i = s0 + 4 * g + q;
            if (i < n) {
                const float y = Y[i];
                const float z = F.logit[q];
                const float cw = (y > 0.5f) ? pos_weight : 1.0f;
                const float sp = fmaxf(z, 0.0f) + log1pf(expf(-fabsf(z)));
                lsum += (c == 0) ? cw * (sp - y * z) : 0.0f;
                dz[q] = cw * (1.0f / (1.0f + expf(-z)) - y);
            } else {
                dz[q] = 0.0f;
            }
        }
        gb4 += (c == 0) ? (dz[0] + dz[1] + dz[2] + dz[3]) : 0.0f;
        // layer 4 and dZ3 (D layout)
        v4f d3[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                gw4[nt] += dz[q] * bfr(F.h3[nt][q]);
                d3[nt][q] = (F.h3[nt][q] > 0.0f) ? dz[q] * bfr(W.w4[nt]) : 0.0f;
                gb3[nt] += d3[nt][q];
            }
        }
        // dW3 += H2^T dZ3
        v4s ah[4], bd[4];
#pragma unroll
        for (int t4 = 0; t4 < 4; ++t4) {
            ah[t4] = pack4(F.h2[t4][0], F.h2[t4][1], F.h2[t4][2], F.h2[t4][3]);
            bd[t4] = pack4(d3[t4][0], d3[t4][1], d3[t4][2], d3[t4][3]);
        }
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) gw3[mt][nt] = mfma(ah[mt], bd[nt], gw3[mt][nt]);
        // dH2 = dZ3 W3^T, dZ2 = dH2 * [H2 > 0]
        v4s ad[4];
        d_to_a(d3, L[wv], lane, ad);
        v4f d2[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) acc = mfma(ad[kt], w3t[kt][nt], acc);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                acc[q] = (F.h2[nt][q] > 0.0f) ? acc[q] : 0.0f;
                gb2[nt] += acc[q];
            }
            d2[nt] = acc;
        }
        // dW2 += H1^T dZ2
#pragma unroll
        for (int t4 = 0; t4 < 4; ++t4) {
            ah[t4] = pack4(F.h1[t4][0], F.h1[t4][1], F.h1[t4][2], F.h1[t4][3]);
            bd[t4] = pack4(d2[t4][0], d2[t4][1], d2[t4][2], d2[t4][3]);
        }
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) gw2[mt][nt] = mfma(ah[mt], bd[nt], gw2[mt][nt]);
        // dH1 = dZ2 W2^T, dZ1 = dH1 * [H1 > 0]
        d_to_a(d2, L[wv], lane, ad);
        v4f d1[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) acc = mfma(ad[kt], w2t[kt][nt], acc);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                acc[q] = (F.h1[nt][q] > 0.0f) ? acc[q] : 0.0f;
                gb1[nt] += acc[q];
            }
            d1[nt] = acc;
        }
        // dW1 += X^T dZ1: X in D-like layout (m = feature l % 16, k = sample 4 (l / 16) + q)
        float xv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int64_t i = s0 + 4 * g + q;
            xv[q] = (i < n && c < kIn) ? (X[12 * i + c] - P[OFF_MEAN + c]) * P[OFF_ISTD + c] : 0.0f;
        }
        const v4s ax = pack4(xv[0], xv[1], xv[2], xv[3]);
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            bd[nt] = pack4(d1[nt][0], d1[nt][1], d1[nt][2], d1[nt][3]);
            gw1[nt] = mfma(ax, bd[nt], gw1[nt]);
        }
    }
    // flush: D[m][n] of the weight-gradient tiles has m = 4 (l / 16) + q + 16 mt, n = l % 16 + 16 nt
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int m = 4 * g + q;
            if (m < kIn) atomicAdd(&G[OFF_W1 + m * kH + 16 * nt + c], gw1[nt][q]);
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) {
                atomicAdd(&G[OFF_W2 + (16 * mt + m) * kH + 16 * nt + c], gw2[mt][nt][q]);
                atomicAdd(&G[OFF_W3 + (16 * mt + m) * kH + 16 * nt + c], gw3[mt][nt][q]);
            }
        }
        const float b1s = group_sum(gb1[nt]), b2s = group_sum(gb2[nt]), b3s = group_sum(gb3[nt]),
                    w4s = group_sum(gw4[nt]);
        if (g == 0) {
            atomicAdd(&G[OFF_B1 + 16 * nt + c], b1s);
            atomicAdd(&G[OFF_B2 + 16 * nt + c], b2s);
            atomicAdd(&G[OFF_B3 + 16 * nt + c], b3s);
            atomicAdd(&G[OFF_W4 + 16 * nt + c], w4s);
        }
    }
    float b4s = gb4, ls = lsum;
    for (int o = 1; o < 64; o <<= 1) {
        b4s += __shfl_xor(b4s, o);
        ls += __shfl_xor(ls, o);
    }
    if (lane == 0) {
        atomicAdd(&G[OFF_B4], b4s);
        atomicAdd(loss, ls);
    }
}

__global__ void adam_kernel(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ M,
                            float* __restrict__ V, int n, float lr, float b1, float b2, float eps, float bc1,
                            float bc2, float scale, float wd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = G[i] * scale + wd * P[i];
    const float m = b1 * M[i] + (1.0f - b1) * g;
    const float v = b2 * V[i] + (1.0f - b2) * g * g;
    M[i] = m;
    V[i] = v;
    P[i] -= lr * (m / bc1) / (sqrtf(v / bc2) + eps);
}

}  // namespace

int mlp_param_count() { return NPARAM; }
int mlp_blob_count() { return NBLOB; }

hipError_t launch_mlp_forward(const float* P, const float* X, int64_t n, float* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int64_t tiles = (n + 15) / 16;
    int64_t blocks = (tiles + kWaves - 1) / kWaves;
    if (blocks > 4096) blocks = 4096;
    mlp_forward_kernel<<<(unsigned)blocks, 64 * kWaves, 0, s>>>(P, X, n, out);
    return hipGetLastError();
}

hipError_t launch_mlp_grad(const float* P, const float* X, const float* Y, int64_t n, float pos_weight, float* G,
                           float* loss, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const int64_t tiles = (n + 15) / 16;
    int64_t blocks = (tiles + kWaves - 1) / kWaves;
    if (blocks > 1024) blocks = 1024;
    mlp_grad_kernel<<<(unsigned)blocks, 64 * kWaves, 0, s>>>(P, X, Y, n, pos_weight, G, loss);
    return hipGetLastError();
}

hipError_t launch_adam(float* P, const float* G, float* M, float* V, int n, float lr, float b1, float b2, float eps,
                       int step, float scale, float wd, hipStream_t s) {
    const float bc1 = 1.0f - powf(b1, (float)step), bc2 = 1.0f - powf(b2, (float)step);
    adam_kernel<<<(n + 255) / 256, 256, 0, s>>>(P, G, M, V, n, lr, b1, b2, eps, bc1, bc2, scale, wd);
    return hipGetLastError();
}

}  // namespace rhz
