// V-cycle kernels of the smoothed-aggregation AMG (FP32).  Short rows (level 0's K, the
// prolongations) in SELL-64: one wave per 64-row slice, one lane per row (the slice's
// entries column-major, so every load of values and column indices is coalesced; padding
// entries hold 0 and a valid column).  Long rows (the restrictions: tens to thousands of
// entries; the coarse levels' matrices) in CSR with a group of G lanes per row, the
// group's lanes striding over the row's entries and summed by shuffles: one lane per such
// row would leave a few hundred waves, each walking a thousand-entry row serially.
// Each kernel returns at once when *skip is set (iterations queued past convergence).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>

#include "amg.h"

namespace mm {
namespace {

constexpr int kSlicesPerBlock = kBlock / kWave;

__device__ __forceinline__ float sell_row(const SellDev<float> &A, int slice, int lane, const float *__restrict__ x) {
    const int beg = A.slice_ptr[slice], end = A.slice_ptr[slice + 1];
    float acc = 0.f;
    for (int k = beg + lane; k < end; k += kWave) acc += A.val[k] * x[A.col[k]];
    return acc;
}

// MODE 0: out = w dinv b (x unused); 1: out = x + w dinv (b - A x); 2: out = b - A x
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_amg_smooth(SellDev<float> A, const float *__restrict__ dinv,
                                                       const float *__restrict__ b, const float *__restrict__ x,
                                                       float *__restrict__ out, float w, const int *skip) {
    if (skip && *skip) return;
    const int lane = threadIdx.x & (kWave - 1);
    for (int slice = blockIdx.x * kSlicesPerBlock + threadIdx.x / kWave; slice < A.n_slices;
         slice += gridDim.x * kSlicesPerBlock) {
        const int row = slice * kWave + lane;
        const float ax = MODE == 0 ? 0.f : sell_row(A, slice, lane, x);
        if (row >= A.n_rows) continue;
        if (MODE == 0) out[row] = w * dinv[row] * b[row];
        else if (MODE == 1) out[row] = x[row] + w * dinv[row] * (b[row] - ax);
        else out[row] = b[row] - ax;
    }
}

// CSR rows, G lanes each (64 / G rows per wave).  MODE 1: out = x + w dinv (b - A x);
// 2: out = b - A x; 3 (restriction, x = the fine residual): out = A x and, when out2 is
// given, out2 = w dinv out (dinv, w: the coarse level's first smoothing step)
template <int G, int MODE>
__global__ __launch_bounds__(kBlock) void k_amg_csr(CsrDevF A, const float *__restrict__ dinv,
                                                    const float *__restrict__ b, const float *__restrict__ x,
                                                    float *__restrict__ out, float *__restrict__ out2, float w,
                                                    const int *skip) {
    if (skip && *skip) return;
    constexpr int kRows = kWave / G;
    const int lane = threadIdx.x & (kWave - 1), sub = lane & (G - 1);
    // the loop bound depends on the wave only: every lane reaches the shuffles
    for (int t = blockIdx.x * kSlicesPerBlock + threadIdx.x / kWave; t * kRows < A.n_rows;
         t += gridDim.x * kSlicesPerBlock) {
        const int row = t * kRows + lane / G;
        float acc = 0.f;
        if (row < A.n_rows)
            for (int k = A.ptr[row] + sub; k < A.ptr[row + 1]; k += G) acc += A.val[k] * x[A.col[k]];
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kWave);
        if (row >= A.n_rows || sub != 0) continue;
        if (MODE == 1) out[row] = x[row] + w * dinv[row] * (b[row] - acc);
        else if (MODE == 2) out[row] = b[row] - acc;
        else {
            out[row] = acc;
            if (out2) out2[row] = w * dinv[row] * acc;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_amg_prolong(SellDev<float> P, const float *__restrict__ xc,
                                                        float *__restrict__ x, const int *skip) {
    if (skip && *skip) return;
    const int lane = threadIdx.x & (kWave - 1);
    for (int slice = blockIdx.x * kSlicesPerBlock + threadIdx.x / kWave; slice < P.n_slices;
         slice += gridDim.x * kSlicesPerBlock) {
        const int row = slice * kWave + lane;
        const float v = sell_row(P, slice, lane, xc);
        if (row < P.n_rows) x[row] += v;
    }
}

// x = inv b: one wave per row, the wave's lanes over the columns, then a wave sum
__global__ __launch_bounds__(kBlock) void k_amg_dense(const float *__restrict__ inv, int n,
                                                      const float *__restrict__ b, float *__restrict__ x,
                                                      const int *skip) {
    if (skip && *skip) return;
    const int lane = threadIdx.x & (kWave - 1);
    for (int row = blockIdx.x * kSlicesPerBlock + threadIdx.x / kWave; row < n; row += gridDim.x * kSlicesPerBlock) {
        float acc = 0.f;
        for (int c = lane; c < n; c += kWave) acc += inv[(size_t)row * n + c] * b[c];
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kWave);
        if (lane == 0) x[row] = acc;
    }
}

int slice_blocks(int n_slices) {
    return std::max(1, std::min((n_slices + kSlicesPerBlock - 1) / kSlicesPerBlock, 4096));
}

template <int MODE>
void launch_csr(const Launch &L, const CsrDevF &A, const float *dinv, const float *b, const float *x, float *out,
                float *out2, float w, const int *skip) {
    const int waves = (A.n_rows + kWave / A.group - 1) / (kWave / A.group);
    const dim3 g(slice_blocks(waves)), blk(kBlock);
    switch (A.group) {
        case 4: hipLaunchKernelGGL((k_amg_csr<4, MODE>), g, blk, 0, L.s, A, dinv, b, x, out, out2, w, skip); break;
        case 8: hipLaunchKernelGGL((k_amg_csr<8, MODE>), g, blk, 0, L.s, A, dinv, b, x, out, out2, w, skip); break;
        case 16: hipLaunchKernelGGL((k_amg_csr<16, MODE>), g, blk, 0, L.s, A, dinv, b, x, out, out2, w, skip); break;
        case 32: hipLaunchKernelGGL((k_amg_csr<32, MODE>), g, blk, 0, L.s, A, dinv, b, x, out, out2, w, skip); break;
        default: hipLaunchKernelGGL((k_amg_csr<64, MODE>), g, blk, 0, L.s, A, dinv, b, x, out, out2, w, skip); break;
    }
}

}  // namespace

void launch_amg_smooth(const Launch &L, const SellDev<float> &A, const float *dinv, const float *b,
                       const float *x, float *out, float w, const int *skip) {
    const dim3 g(slice_blocks(A.n_slices)), blk(kBlock);
    if (x && L.t0)   // timed: events recorded on the dispatch itself
        hipExtLaunchKernelGGL(k_amg_smooth<1>, g, blk, 0, L.s, L.t0, L.t1, 0, A, dinv, b, x, out, w, skip);
    else if (x)
        hipLaunchKernelGGL(k_amg_smooth<1>, g, blk, 0, L.s, A, dinv, b, x, out, w, skip);
    else
        hipLaunchKernelGGL(k_amg_smooth<0>, g, blk, 0, L.s, A, dinv, b, x, out, w, skip);
}
void launch_amg_resid(const Launch &L, const SellDev<float> &A, const float *b, const float *x, float *out,
                      const int *skip) {
    hipLaunchKernelGGL(k_amg_smooth<2>, dim3(slice_blocks(A.n_slices)), dim3(kBlock), 0, L.s, A, nullptr, b, x,
                       out, 0.f, skip);
}
void launch_amg_smooth(const Launch &L, const CsrDevF &A, const float *dinv, const float *b, const float *x,
                       float *out, float w, const int *skip) {
    launch_csr<1>(L, A, dinv, b, x, out, nullptr, w, skip);
}
void launch_amg_resid(const Launch &L, const CsrDevF &A, const float *b, const float *x, float *out,
                      const int *skip) {
    launch_csr<2>(L, A, nullptr, b, x, out, nullptr, 0.f, skip);
}
void launch_amg_restrict(const Launch &L, const CsrDevF &R, const float *r, float *bc, float *x1c,
                         const float *dinvc, float w, const int *skip) {
    launch_csr<3>(L, R, dinvc, nullptr, r, bc, x1c, w, skip);
}
void launch_amg_prolong(const Launch &L, const SellDev<float> &P, const float *xc, float *x, const int *skip) {
    hipLaunchKernelGGL(k_amg_prolong, dim3(slice_blocks(P.n_slices)), dim3(kBlock), 0, L.s, P, xc, x, skip);
}
void launch_amg_dense(const Launch &L, const float *inv, int n, const float *b, float *x, const int *skip) {
    hipLaunchKernelGGL(k_amg_dense, dim3(slice_blocks(n)), dim3(kBlock), 0, L.s, inv, n, b, x, skip);   // a wave per row
}

}  // namespace mm
