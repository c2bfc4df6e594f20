// Per-node MLP of the learned warm start (no counterpart in the reference): features
// of a node and its 1-ring -> 128 -> 128 -> 128 -> 1, ReLU, bf16 MFMA with FP32
// accumulation (the forward pass in v_mfma_f32_16x16x32_bf16, the backward and weight-
// gradient products in v_mfma_f32_16x16x16_bf16), FP32 master weights and Adam.
//
// Products are formed transposed, H_out^T = W^T H_in^T (features x nodes), so that the
// 16x16 output fragment of a layer (lane l: column = node l % 16, rows = features
// 4 (l / 16) + e) is exactly the B operand of the next layer for the matching k-step:
// activations never leave registers between layers.  A wave handles 32 nodes (two
// 16-node column tiles) so that each weight fragment read from LDS feeds two MFMAs.
//
// Fragment layouts of v_mfma_f32_16x16x16_bf16 (64 lanes, e = 0..3):
//   A[i][k]: lane = i + 16 (k / 4), element k % 4
//   B[k][j]: lane = j + 16 (k / 4), element k % 4
//   D[i][j]: lane = j + 16 (i / 4), element i % 4
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.h"
#include "mlp.h"

namespace mm {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kWaves = 8;                 // waves per block of the forward / backward kernels
constexpr int kNT = kWaves * kWave;       // threads per block

__device__ __forceinline__ uint16_t f2bf(float x) {   // round to nearest, ties to even
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Two consecutive 16-wide k-steps in one v_mfma_f32_16x16x32_bf16 (gfx950: twice the
// products per instruction).  Its operands hold 8 k-values per lane; a lane's elements
// 0-3 take the first k-step's fragment (k = 4 g + e) and 4-7 the second's (k = 16 + 4 g + e),
// for A and B alike: the same permutation of the 32 summed k-values on both sides, so the
// 16x16x16 fragments above combine without moving data (checked lane by lane on the
// device against a CPU product).
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 nbf16x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ f32x4 mfma2(bf16x4 a0, bf16x4 a1, bf16x4 b0, bf16x4 b1, f32x4 c) {
    const bf16x8 a = __builtin_bit_cast(bf16x8, (s16x8)__builtin_shufflevector(a0, a1, 0, 1, 2, 3, 4, 5, 6, 7));
    const bf16x8 b = __builtin_bit_cast(bf16x8, (s16x8)__builtin_shufflevector(b0, b1, 0, 1, 2, 3, 4, 5, 6, 7));
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Cooperative copy of `bytes` (multiple of 16) from global to LDS.
__device__ __forceinline__ void copy_lds(void *dst, const void *src, int bytes) {
    const uint4 *s = (const uint4 *)src;
    uint4 *d = (uint4 *)dst;
    for (int i = threadIdx.x; i < bytes / 16; i += blockDim.x) d[i] = s[i];
}

// Weight fragment (A operand): rows 16 t + (lane % 16), k = 16 f + 4 (lane / 16) + 0..3
// of a row-major bf16 matrix with padded row length `ld`.
__device__ __forceinline__ bf16x4 afrag(const uint16_t *w, int ld, int t, int f, int col, int grp) {
    return *(const bf16x4 *)(w + (16 * t + col) * ld + 16 * f + 4 * grp);
}

// ------------------------------------------------------------------ forward
// X: [n][kMlpIn] node-major bf16.  y: [n] FP32.  TRAIN: also the hidden activations,
// feature-major bf16 [128][ldh] (row = feature, column = node).
template <bool TRAIN>
__global__ __launch_bounds__(kNT) void k_mlp_fwd(MlpDev w, int64_t n, const uint16_t *__restrict__ X,
                                                 float *__restrict__ y, uint16_t *__restrict__ h1,
                                                 uint16_t *__restrict__ h2, uint16_t *__restrict__ h3, int64_t ldh) {
    __shared__ uint16_t s1[kMlpW * kMlpP1];
    __shared__ uint16_t s2[kMlpW * kMlpPW];
    __shared__ uint16_t s3[kMlpW * kMlpPW];
    __shared__ __attribute__((aligned(16))) float sb[3][kMlpW];
    __shared__ __attribute__((aligned(16))) float sw4[kMlpW];
    copy_lds(s1, w.w1t, kMlpW * kMlpP1 * 2);
    copy_lds(s2, w.w2t, kMlpW * kMlpPW * 2);
    copy_lds(s3, w.w3t, kMlpW * kMlpPW * 2);
    for (int i = threadIdx.x; i < kMlpW; i += blockDim.x) {
        sb[0][i] = w.b1[i];
        sb[1][i] = w.b2[i];
        sb[2][i] = w.b3[i];
        sw4[i] = w.w4[i];
    }
    __syncthreads();
    const float b4 = w.b4[0];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int col = lane & 15, grp = lane >> 4;
    uint16_t *hs[3] = {h1, h2, h3};
    // training: every node of the padded batch [0, ldh) gets activations (zero inputs
    // past n), so that the weight-gradient GEMMs never read unset memory
    const int64_t nrun = TRAIN ? ldh : n;
    for (int64_t base = ((int64_t)blockIdx.x * kWaves + wave) * 32; base < nrun;
         base += (int64_t)gridDim.x * kWaves * 32) {
        int64_t node[2];
        bool ok[2], st[2];
        bf16x4 hb[8][2];   // B operands of the next layer, per k-step
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            node[c] = base + 16 * c + col;
            ok[c] = node[c] < n;
            st[c] = node[c] < nrun;
        }
        // layer 1 inputs: X[node][16 f + 4 grp + 0..3]
#pragma unroll
        for (int f = 0; f < kMlpIn / 16; ++f)
#pragma unroll
            for (int c = 0; c < 2; ++c)
                hb[f][c] = ok[c] ? *(const bf16x4 *)(X + node[c] * kMlpIn + 16 * f + 4 * grp) : bf16x4{0, 0, 0, 0};
        float yacc[2] = {0.f, 0.f};
#pragma unroll
        for (int layer = 0; layer < 3; ++layer) {
            const uint16_t *sw = layer == 0 ? s1 : layer == 1 ? s2 : s3;
            const int ld = layer == 0 ? kMlpP1 : kMlpPW;
            const int nk = layer == 0 ? kMlpIn / 16 : kMlpW / 16;
            // the accumulators start at the bias (features 16 t + 4 grp + 0..3 of this lane)
            f32x4 acc[8][2];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const f32x4 bias = *(const f32x4 *)&sb[layer][16 * t + 4 * grp];
#pragma unroll
                for (int c = 0; c < 2; ++c) acc[t][c] = bias;
            }
#pragma unroll
            for (int f = 0; f < nk; f += 2) {
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    const bf16x4 a0 = afrag(sw, ld, t, f, col, grp), a1 = afrag(sw, ld, t, f + 1, col, grp);
#pragma unroll
                    for (int c = 0; c < 2; ++c) acc[t][c] = mfma2(a0, a1, hb[f][c], hb[f + 1][c], acc[t][c]);
                }
                // keep the weight reads of one k-step pair in flight at a time (registers)
                __builtin_amdgcn_sched_barrier(0);
            }
            // ReLU + bf16 (round to nearest even, two values per conversion instruction): the
            // next layer's B operands (k-step t)
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const f32x4 w4 = layer == 2 ? *(const f32x4 *)&sw4[16 * t + 4 * grp] : f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    f32x4 z = acc[t][c];
#pragma unroll
                    for (int e = 0; e < 4; ++e) z[e] = fmaxf(z[e], 0.f);
                    const nbf16x4 hv = __builtin_convertvector(z, nbf16x4);
                    const bf16x4 v = __builtin_bit_cast(bf16x4, hv);
                    if (layer == 2) {
                        const f32x4 r = __builtin_convertvector(hv, f32x4);
                        yacc[c] += w4[0] * r[0] + w4[1] * r[1] + w4[2] * r[2] + w4[3] * r[3];
                    }
                    if (TRAIN && st[c]) {
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            hs[layer][(int64_t)(16 * t + 4 * grp + e) * ldh + node[c]] = (uint16_t)v[e];
                    }
                    hb[t][c] = v;
                }
            }
        }
        // layer 4: sum over the 4 lane groups holding the node's 128 features
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            float s = yacc[c];
            s += __shfl_xor(s, 16, kWave);
            s += __shfl_xor(s, 32, kWave);
            if (grp == 0 && ok[c]) y[node[c]] = s + b4;
        }
    }
}

// ------------------------------------------------------------------ loss
// dy = 2 (y - t) / n; partial sums of (y - t)^2 per block (atomically added).
__global__ void k_mlp_loss(int64_t n, const float *__restrict__ y, const float *__restrict__ t,
                           float *__restrict__ dy, double *__restrict__ loss) {
    double s = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float d = y[i] - t[i];
        dy[i] = 2.f * d / (float)n;
        s += (double)d * d;
    }
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) atomicAdd(loss, s / (double)n);
}

// ------------------------------------------------------------------ backward (data path)
// dz3 = w4 (x) dy * [h3 > 0];  dz2 = (W3 dz3) * [h2 > 0];  dz1 = (W2 dz2) * [h1 > 0];
// all stored feature-major bf16 [128][ldh].  Nodes >= n get dz = 0.
__global__ __launch_bounds__(kNT) void k_mlp_bwd(MlpDev w, int64_t n, const float *__restrict__ dy,
                                                 const uint16_t *__restrict__ h1, const uint16_t *__restrict__ h2,
                                                 const uint16_t *__restrict__ h3, uint16_t *__restrict__ dz1,
                                                 uint16_t *__restrict__ dz2, uint16_t *__restrict__ dz3,
                                                 int64_t ldh) {
    __shared__ uint16_t s2[kMlpW * kMlpPW];   // W2 row-major [in][out]
    __shared__ uint16_t s3[kMlpW * kMlpPW];   // W3
    __shared__ __attribute__((aligned(16))) float sw4[kMlpW];
    copy_lds(s2, w.w2n, kMlpW * kMlpPW * 2);
    copy_lds(s3, w.w3n, kMlpW * kMlpPW * 2);
    for (int i = threadIdx.x; i < kMlpW; i += blockDim.x) sw4[i] = w.w4[i];
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int col = lane & 15, grp = lane >> 4;
    for (int64_t base = ((int64_t)blockIdx.x * kWaves + wave) * 32; base < ldh;
         base += (int64_t)gridDim.x * kWaves * 32) {
        int64_t node[2];
        bool ok[2];
        float dyv[2];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            node[c] = base + 16 * c + col;
            ok[c] = node[c] < n;
            dyv[c] = ok[c] ? dy[node[c]] : 0.f;
        }
        bf16x4 dz[8][2];
#pragma unroll
        for (int t = 0; t < 8; ++t)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                bf16x4 v;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int feat = 16 * t + 4 * grp + e;
                    float g = 0.f;
                    if (ok[c] && bf2f(h3[(int64_t)feat * ldh + node[c]]) > 0.f) g = sw4[feat] * dyv[c];
                    v[e] = (short)f2bf(g);
                    if (node[c] < ldh) dz3[(int64_t)feat * ldh + node[c]] = (uint16_t)v[e];
                }
                dz[t][c] = v;
            }
#pragma unroll
        for (int layer = 0; layer < 2; ++layer) {   // layer 0: through W3 -> dz2; 1: through W2 -> dz1
            const uint16_t *sw = layer == 0 ? s3 : s2;
            const uint16_t *hin = layer == 0 ? h2 : h1;
            uint16_t *dzo = layer == 0 ? dz2 : dz1;
            f32x4 acc[8][2];
#pragma unroll
            for (int t = 0; t < 8; ++t)
#pragma unroll
                for (int c = 0; c < 2; ++c) acc[t][c] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int f = 0; f < kMlpW / 16; ++f) {
#pragma unroll
                for (int t = 0; t < 8; ++t) {
                    const bf16x4 a = afrag(sw, kMlpPW, t, f, col, grp);
#pragma unroll
                    for (int c = 0; c < 2; ++c) acc[t][c] = mfma(a, dz[f][c], acc[t][c]);
                }
                __builtin_amdgcn_sched_barrier(0);
            }
#pragma unroll
            for (int t = 0; t < 8; ++t)
#pragma unroll
                for (int c = 0; c < 2; ++c) {
                    bf16x4 v;
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int feat = 16 * t + 4 * grp + e;
                        float g = 0.f;
                        if (ok[c] && bf2f(hin[(int64_t)feat * ldh + node[c]]) > 0.f) g = acc[t][c][e];
                        v[e] = (short)f2bf(g);
                        if (node[c] < ldh) dzo[(int64_t)feat * ldh + node[c]] = (uint16_t)v[e];
                    }
                    dz[t][c] = v;
                }
        }
    }
}

// ------------------------------------------------------------------ weight gradients
// dW[i][j] += sum over nodes of A[i][node] * B[j][node] (i < KIN, j < 128), with A
// feature-major [KIN][lda] (XNODE: node-major [n][kMlpIn], the layer-1 inputs) and B
// feature-major [128][ldb].  One wave per (16 x 16 tile, node chunk); the chunk's
// partial tile is added atomically.  n is a multiple of 16 (zero-padded batches).
template <int KIN, bool XNODE>
__global__ __launch_bounds__(256) void k_mlp_dw(int64_t n, int64_t nx, int64_t chunk, const uint16_t *__restrict__ A,
                                                int64_t lda, const uint16_t *__restrict__ B, int64_t ldb,
                                                float *__restrict__ dW) {
    const int lane = threadIdx.x & (kWave - 1);
    const int col = lane & 15, grp = lane >> 4;
    const int gw = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    constexpr int ntiles = (KIN / 16) * (kMlpW / 16);
    const int tile = gw % ntiles;
    const int64_t c0 = (int64_t)(gw / ntiles) * chunk;
    if (c0 >= n) return;
    const int64_t c1 = c0 + chunk < n ? c0 + chunk : n;
    const int ti = tile / (kMlpW / 16), tj = tile % (kMlpW / 16);
    const int i = 16 * ti + col, j = 16 * tj + col;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int64_t k0 = c0; k0 < c1; k0 += 16) {
        const int64_t kk = k0 + 4 * grp;
        bf16x4 a;
        if (XNODE) {   // X has nx rows (the batch); zero past them
#pragma unroll
            for (int e = 0; e < 4; ++e) a[e] = kk + e < nx ? (short)A[(kk + e) * kMlpIn + i] : (short)0;
        } else {
            a = *(const bf16x4 *)(A + (int64_t)i * lda + kk);
        }
        const bf16x4 b = *(const bf16x4 *)(B + (int64_t)j * ldb + kk);
        acc = mfma(a, b, acc);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) atomicAdd(&dW[(16 * ti + 4 * grp + e) * kMlpW + 16 * tj + col], acc[e]);
}

// out[r] += sum over nodes of A[r][node] * (v ? v[node] : 1) for the 128 rows of a
// feature-major bf16 [128][lda] array (bias gradients: v = null; the layer-4 weight
// gradient: v = dy).  Block = one row x one node chunk.
__global__ void k_mlp_rowdot(int64_t n, int64_t chunk, const uint16_t *__restrict__ A, int64_t lda,
                             const float *__restrict__ v, float *__restrict__ out) {
    const int r = blockIdx.x;
    const int64_t c0 = (int64_t)blockIdx.y * chunk;
    const int64_t c1 = c0 + chunk < n ? c0 + chunk : n;
    float s = 0.f;
    for (int64_t k = c0 + threadIdx.x; k < c1; k += blockDim.x)
        s += bf2f(A[(int64_t)r * lda + k]) * (v ? v[k] : 1.f);
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) atomicAdd(&out[r], s);
}

// sum of dy (the layer-4 bias gradient)
__global__ void k_mlp_sum(int64_t n, const float *__restrict__ v, float *__restrict__ out) {
    float s = 0.f;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
        s += v[k];
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) atomicAdd(out, s);
}

// ------------------------------------------------------------------ Adam + publish
// Flat FP32 parameters (MlpLayout offsets); then the bf16 copies the kernels read.
__global__ void k_mlp_adam(int np, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m1,
                           float *__restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < np; i += gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float a = b1 * m1[i] + (1.f - b1) * gi;
        const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
        m1[i] = a;
        m2[i] = b;
        p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
    }
}

// bf16 copies: w1t [128][kMlpP1] = W1^T, w2t / w3t [128][kMlpPW] = W^T (forward),
// w2n / w3n [128][kMlpPW] = W (backward); pads zero.
__global__ void k_mlp_publish(const float *__restrict__ p, uint16_t *__restrict__ w1t, uint16_t *__restrict__ w2t,
                              uint16_t *__restrict__ w3t, uint16_t *__restrict__ w2n, uint16_t *__restrict__ w3n) {
    const int r = blockIdx.x;   // row 0..127
    for (int k = threadIdx.x; k < kMlpPW; k += blockDim.x) {
        if (k < kMlpP1) w1t[r * kMlpP1 + k] = k < kMlpIn ? f2bf(p[kMlpOffW1 + k * kMlpW + r]) : 0;
        const bool in = k < kMlpW;
        w2t[r * kMlpPW + k] = in ? f2bf(p[kMlpOffW2 + k * kMlpW + r]) : 0;
        w3t[r * kMlpPW + k] = in ? f2bf(p[kMlpOffW3 + k * kMlpW + r]) : 0;
        w2n[r * kMlpPW + k] = in ? f2bf(p[kMlpOffW2 + r * kMlpW + k]) : 0;
        w3n[r * kMlpPW + k] = in ? f2bf(p[kMlpOffW3 + r * kMlpW + k]) : 0;
    }
}

// ------------------------------------------------------------------ features
// Node features on the structured grid (node (i, j, k) = i + nx (j + ny k)):
//   0..20  m of the node and of its 6 axis neighbours (order: self, -x, +x, -y, +y, -z, +z;
//          0 outside the grid)
//   21..23 node coordinates / half box edge
//   24..26 mean magnetisation of the magnet (sum_i m_i meas_i / V; mbar[3] holds the sums)
//   27     fraction of the 8 adjacent voxels that are magnetic
//   28     mean Js of the adjacent magnetic voxels / Js_max
//   29..31 zero
__global__ void k_mlp_features(FeatGrid F, const MatCoef *__restrict__ mats, const double *__restrict__ m,
                               const double *__restrict__ mbar, uint16_t *__restrict__ X) {
    const int64_t N = (int64_t)F.nx * F.ny * F.nz;
    const double mb0 = mbar[0], mb1 = mbar[1], mb2 = mbar[2];
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < N; idx += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(idx % F.nx);
        const int j = (int)((idx / F.nx) % F.ny);
        const int k = (int)(idx / ((int64_t)F.nx * F.ny));
        float f[kMlpIn];
        const int64_t nb[7] = {idx, i > 0 ? idx - 1 : -1, i + 1 < F.nx ? idx + 1 : -1,
                               j > 0 ? idx - F.nx : -1, j + 1 < F.ny ? idx + F.nx : -1,
                               k > 0 ? idx - (int64_t)F.nx * F.ny : -1,
                               k + 1 < F.nz ? idx + (int64_t)F.nx * F.ny : -1};
#pragma unroll
        for (int q = 0; q < 7; ++q)
#pragma unroll
            for (int a = 0; a < 3; ++a) f[3 * q + a] = nb[q] >= 0 ? (float)m[3 * nb[q] + a] : 0.f;
        f[21] = (float)(F.x[i] * F.inv_half);
        f[22] = (float)(F.y[j] * F.inv_half);
        f[23] = (float)(F.z[k] * F.inv_half);
        f[24] = (float)(mb0 * F.inv_V);
        f[25] = (float)(mb1 * F.inv_V);
        f[26] = (float)(mb2 * F.inv_V);
        // adjacent voxels (vi, vj, vk) = (i - 1 + di, ...), magnetic when inside the magnet block
        int nmag = 0;
        double js = 0.0;
        for (int d = 0; d < 8; ++d) {
            const int vi = i - 1 + (d & 1) - F.m0x, vj = j - 1 + ((d >> 1) & 1) - F.m0y,
                      vk = k - 1 + ((d >> 2) & 1) - F.m0z;
            if (vi < 0 || vj < 0 || vk < 0 || vi >= F.mn || vj >= F.mn || vk >= F.mn) continue;
            const int mat = F.vmat[((int64_t)vk * F.mn + vj) * F.mn + vi];
            const double J = mats[mat].Js;
            if (J > 0.0) {
                ++nmag;
                js += J;
            }
        }
        f[27] = nmag * 0.125f;
        f[28] = nmag ? (float)(js / nmag * F.inv_js) : 0.f;
        f[29] = f[30] = f[31] = 0.f;
        uint16_t *o = X + idx * kMlpIn;
#pragma unroll
        for (int q = 0; q < kMlpIn; q += 4) {
            bf16x4 v = {(short)f2bf(f[q]), (short)f2bf(f[q + 1]), (short)f2bf(f[q + 2]), (short)f2bf(f[q + 3])};
            *(bf16x4 *)(o + q) = v;
        }
    }
}

// Direction of the learned warm-start step, FP64: d = s (pred - prev) when the previous
// prediction exists, else d = s pred - u; zero on the Dirichlet nodes (the six faces of
// the grid box), where u stays 0.
__global__ void k_mlp_direction(int nx, int ny, int nz, const float *__restrict__ pred,
                                const float *__restrict__ prev, const double *__restrict__ u, double s,
                                double *__restrict__ d) {
    const int64_t N = (int64_t)nx * ny * nz;
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < N; idx += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(idx % nx), j = (int)((idx / nx) % ny), k = (int)(idx / ((int64_t)nx * ny));
        const bool face = i == 0 || j == 0 || k == 0 || i == nx - 1 || j == ny - 1 || k == nz - 1;
        d[idx] = face ? 0.0 : prev ? s * ((double)pred[idx] - (double)prev[idx]) : s * (double)pred[idx] - u[idx];
    }
}

// Galerkin step u += gamma d with gamma = num / den (den = d.K d; skipped unless
// den > 0 and gamma is finite).  gamma is recorded in *gamma_out.
__global__ void k_mlp_galerkin(int64_t n, const double *__restrict__ num, const double *__restrict__ den,
                               const double *__restrict__ d, double *__restrict__ u, double *gamma_out) {
    const double dd = *den;
    double g = dd > 0.0 ? *num / dd : 0.0;
    if (!(fabs(g) < 1e300)) g = 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) *gamma_out = g;
    if (g == 0.0) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        u[i] += g * d[i];
}

int grid_blocks(int64_t work, int per_block, int cap) {
    const int64_t b = (work + per_block - 1) / per_block;
    return (int)(b < 1 ? 1 : b > cap ? cap : b);
}

}  // namespace

void launch_mlp_forward(const Launch &L, const MlpDev &w, int64_t n, const uint16_t *X, float *y) {
    const int nb = grid_blocks(n, kWaves * 32, 2048);
    hipLaunchKernelGGL(k_mlp_fwd<false>, dim3(nb), dim3(kNT), 0, L.s, w, n, X, y, nullptr, nullptr, nullptr,
                       (int64_t)0);
}

void launch_mlp_forward_train(const Launch &L, const MlpDev &w, int64_t n, const uint16_t *X, float *y,
                              uint16_t *h1, uint16_t *h2, uint16_t *h3, int64_t ldh) {
    const int nb = grid_blocks(n, kWaves * 32, 2048);
    hipLaunchKernelGGL(k_mlp_fwd<true>, dim3(nb), dim3(kNT), 0, L.s, w, n, X, y, h1, h2, h3, ldh);
}

void launch_mlp_loss(const Launch &L, int64_t n, const float *y, const float *t, float *dy, double *loss) {
    hipLaunchKernelGGL(k_mlp_loss, dim3(grid_blocks(n, 256, 1024)), dim3(256), 0, L.s, n, y, t, dy, loss);
}

void launch_mlp_backward(const Launch &L, const MlpDev &w, int64_t n, const float *dy, const uint16_t *h1,
                         const uint16_t *h2, const uint16_t *h3, uint16_t *dz1, uint16_t *dz2, uint16_t *dz3,
                         int64_t ldh) {
    const int nb = grid_blocks(ldh, kWaves * 32, 2048);
    hipLaunchKernelGGL(k_mlp_bwd, dim3(nb), dim3(kNT), 0, L.s, w, n, dy, h1, h2, h3, dz1, dz2, dz3, ldh);
}

// All parameter gradients into grad (flat, zeroed by the caller): n = batch, ldh = padded
// batch (multiple of 32; activations set and dz = 0 on [n, ldh)); dy has n entries.
void launch_mlp_param_grads(const Launch &L, int64_t n, int64_t ldh, const uint16_t *X, const float *dy,
                            const uint16_t *h1, const uint16_t *h2, const uint16_t *h3, const uint16_t *dz1,
                            const uint16_t *dz2, const uint16_t *dz3, float *grad) {
    const int64_t chunk = 1024;
    const int nchunk = (int)((ldh + chunk - 1) / chunk);
    {
        const int waves = (kMlpIn / 16) * 8 * nchunk;
        hipLaunchKernelGGL((k_mlp_dw<kMlpIn, true>), dim3((waves + 3) / 4), dim3(256), 0, L.s, ldh, n, chunk, X,
                           (int64_t)0, dz1, ldh, grad + kMlpOffW1);
    }
    const int waves = 64 * nchunk;
    hipLaunchKernelGGL((k_mlp_dw<kMlpW, false>), dim3((waves + 3) / 4), dim3(256), 0, L.s, ldh, ldh, chunk, h1, ldh,
                       dz2, ldh, grad + kMlpOffW2);
    hipLaunchKernelGGL((k_mlp_dw<kMlpW, false>), dim3((waves + 3) / 4), dim3(256), 0, L.s, ldh, ldh, chunk, h2, ldh,
                       dz3, ldh, grad + kMlpOffW3);
    const dim3 rg(kMlpW, nchunk);
    hipLaunchKernelGGL(k_mlp_rowdot, rg, dim3(256), 0, L.s, ldh, chunk, dz1, ldh, (const float *)nullptr,
                       grad + kMlpOffB1);
    hipLaunchKernelGGL(k_mlp_rowdot, rg, dim3(256), 0, L.s, ldh, chunk, dz2, ldh, (const float *)nullptr,
                       grad + kMlpOffB2);
    hipLaunchKernelGGL(k_mlp_rowdot, rg, dim3(256), 0, L.s, ldh, chunk, dz3, ldh, (const float *)nullptr,
                       grad + kMlpOffB3);
    const dim3 rgn(kMlpW, (unsigned)((n + chunk - 1) / chunk));
    hipLaunchKernelGGL(k_mlp_rowdot, rgn, dim3(256), 0, L.s, n, chunk, h3, ldh, dy, grad + kMlpOffW4);
    hipLaunchKernelGGL(k_mlp_sum, dim3(grid_blocks(n, 256, 256)), dim3(256), 0, L.s, n, dy, grad + kMlpOffB4);
}

void launch_mlp_adam(const Launch &L, float *p, const float *g, float *m1, float *m2, float lr, int step) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_mlp_adam, dim3(grid_blocks(kMlpParams, 256, 256)), dim3(256), 0, L.s, kMlpParams, p, g, m1,
                       m2, lr, b1, b2, eps, c1, c2);
}

void launch_mlp_publish(const Launch &L, const float *p, uint16_t *w1t, uint16_t *w2t, uint16_t *w3t, uint16_t *w2n,
                        uint16_t *w3n) {
    hipLaunchKernelGGL(k_mlp_publish, dim3(kMlpW), dim3(64), 0, L.s, p, w1t, w2t, w3t, w2n, w3n);
}

void launch_mlp_features(const Launch &L, const FeatGrid &F, const double *m, const double *mbar, uint16_t *X) {
    const int64_t N = (int64_t)F.nx * F.ny * F.nz;
    hipLaunchKernelGGL(k_mlp_features, dim3(grid_blocks(N, 256, 8192)), dim3(256), 0, L.s, F, L.mats, m, mbar, X);
}

void launch_mlp_direction(const Launch &L, const FeatGrid &F, const float *pred, const float *prev, const double *u,
                          double s, double *d) {
    const int64_t N = (int64_t)F.nx * F.ny * F.nz;
    hipLaunchKernelGGL(k_mlp_direction, dim3(grid_blocks(N, 256, 4096)), dim3(256), 0, L.s, F.nx, F.ny, F.nz, pred,
                       prev, u, s, d);
}

void launch_mlp_galerkin(const Launch &L, int64_t n, const double *num, const double *den, const double *d,
                         double *u, double *gamma_out) {
    hipLaunchKernelGGL(k_mlp_galerkin, dim3(grid_blocks(n, 256, 4096)), dim3(256), 0, L.s, n, num, den, d, u,
                       gamma_out);
}

// ------------------------------------------------------------------ learned preconditioner
// Features of the residual network (a second instance of the same MLP): per node, the
// Jacobi-scaled residual w = r / diag(K) of the node and of its neighbours one and two
// nodes away along each axis (13 values, 0 outside the grid), divided by s = RMS(w); the
// six adjacent spacings over their mean; the face (Dirichlet) flag; the magnetic fraction
// of the 8 adjacent voxels; the coordinates over the half box edge.  Output y: the
// predicted error e = K^-1 r at the node in units of c s, c the network's output scale
// (z = c s y, 0 on the faces).
__global__ void k_lp_scaled(int64_t n, const double *__restrict__ r, const double *__restrict__ invD,
                            double *__restrict__ w) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        w[i] = r[i] * invD[i];
}

__global__ void k_lp_features(FeatGrid F, PoissonGrid<double> P, const MatCoef *__restrict__ mats,
                              const double *__restrict__ w, const double *__restrict__ ss, uint16_t *__restrict__ X) {
    const int64_t N = (int64_t)F.nx * F.ny * F.nz;
    const double s2 = *ss / (double)N;
    const double inv_s = s2 > 0.0 ? 1.0 / sqrt(s2) : 0.0;
    const int64_t sy = F.nx, sz = (int64_t)F.nx * F.ny;
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < N; idx += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(idx % F.nx);
        const int j = (int)((idx / F.nx) % F.ny);
        const int k = (int)(idx / sz);
        float f[kMlpIn];
        auto at = [&](int di, int dj, int dk) -> float {
            const int a = i + di, b = j + dj, c = k + dk;
            if (a < 0 || b < 0 || c < 0 || a >= F.nx || b >= F.ny || c >= F.nz) return 0.f;
            return (float)(w[idx + di + sy * dj + sz * dk] * inv_s);
        };
        f[0] = at(0, 0, 0);
        f[1] = at(-1, 0, 0); f[2] = at(1, 0, 0);
        f[3] = at(0, -1, 0); f[4] = at(0, 1, 0);
        f[5] = at(0, 0, -1); f[6] = at(0, 0, 1);
        f[7] = at(-2, 0, 0); f[8] = at(2, 0, 0);
        f[9] = at(0, -2, 0); f[10] = at(0, 2, 0);
        f[11] = at(0, 0, -2); f[12] = at(0, 0, 2);
        const double h[6] = {P.hxp[i], P.hxp[i + 1], P.hyp[j], P.hyp[j + 1], P.hzp[k], P.hzp[k + 1]};
        double hs = 0.0;
        int hc = 0;
#pragma unroll
        for (int q = 0; q < 6; ++q)
            if (h[q] > 0.0) {
                hs += h[q];
                ++hc;
            }
        const double ih = hc ? hc / hs : 0.0;
#pragma unroll
        for (int q = 0; q < 6; ++q) f[13 + q] = (float)(h[q] * ih);
        const bool face = i == 0 || j == 0 || k == 0 || i == F.nx - 1 || j == F.ny - 1 || k == F.nz - 1;
        f[19] = face ? 1.f : 0.f;
        int nmag = 0;
        for (int d = 0; d < 8; ++d) {
            const int vi = i - 1 + (d & 1) - F.m0x, vj = j - 1 + ((d >> 1) & 1) - F.m0y,
                      vk = k - 1 + ((d >> 2) & 1) - F.m0z;
            if (vi < 0 || vj < 0 || vk < 0 || vi >= F.mn || vj >= F.mn || vk >= F.mn) continue;
            if (mats[F.vmat[((int64_t)vk * F.mn + vj) * F.mn + vi]].Js > 0.0) ++nmag;
        }
        f[20] = nmag * 0.125f;
        f[21] = (float)(F.x[i] * F.inv_half);
        f[22] = (float)(F.y[j] * F.inv_half);
        f[23] = (float)(F.z[k] * F.inv_half);
#pragma unroll
        for (int q = 24; q < kMlpIn; ++q) f[q] = 0.f;
        uint16_t *o = X + idx * kMlpIn;
#pragma unroll
        for (int q = 0; q < kMlpIn; q += 4) {
            bf16x4 v = {(short)f2bf(f[q]), (short)f2bf(f[q + 1]), (short)f2bf(f[q + 2]), (short)f2bf(f[q + 3])};
            *(bf16x4 *)(o + q) = v;
        }
    }
}

// z = c s y on the interior, 0 on the faces (s = RMS of the scaled residual, from ss; c the
// network's output scale).
__global__ void k_lp_apply(int nx, int ny, int nz, const float *__restrict__ y, const double *__restrict__ ss,
                           double c, double *__restrict__ z) {
    const int64_t N = (int64_t)nx * ny * nz;
    const double s2 = *ss / (double)N;
    const double s = s2 > 0.0 ? c * sqrt(s2) : 0.0;
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < N; idx += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(idx % nx), j = (int)((idx / nx) % ny), k = (int)(idx / ((int64_t)nx * ny));
        const bool face = i == 0 || j == 0 || k == 0 || i == nx - 1 || j == ny - 1 || k == nz - 1;
        z[idx] = face ? 0.0 : s * (double)y[idx];
    }
}

void launch_lp_scaled(const Launch &L, int64_t n, const double *r, const double *invD, double *w) {
    hipLaunchKernelGGL(k_lp_scaled, dim3(grid_blocks(n, 256, 4096)), dim3(256), 0, L.s, n, r, invD, w);
}
void launch_lp_features(const Launch &L, const FeatGrid &F, const PoissonGrid<double> &P, const double *w,
                        const double *ss, uint16_t *X) {
    const int64_t N = (int64_t)F.nx * F.ny * F.nz;
    hipLaunchKernelGGL(k_lp_features, dim3(grid_blocks(N, 256, 8192)), dim3(256), 0, L.s, F, P, L.mats, w, ss, X);
}
void launch_lp_apply(const Launch &L, const FeatGrid &F, const float *y, const double *ss, double c, double *z) {
    const int64_t N = (int64_t)F.nx * F.ny * F.nz;
    hipLaunchKernelGGL(k_lp_apply, dim3(grid_blocks(N, 256, 4096)), dim3(256), 0, L.s, F.nx, F.ny, F.nz, y, ss, c,
                       z);
}

}  // namespace mm
