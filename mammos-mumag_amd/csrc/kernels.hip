// HIP kernels for gfx950 (CDNA4): the per-field-step hot path of mammos-mumag.
//
// Streaming kernels use 256-thread blocks (4 wavefronts of 64) with grid-stride loops
// over at most kMaxPartials blocks.  Reductions are deterministic and free of
// contended atomics: every block writes one partial per output slot, and a one-block
// finisher kernel (k_fin) sums the partials in a fixed order and derives the scalars
// an iteration needs (alpha, beta, convergence flags) on the device, so solver
// iterations are queued back to back without a host round trip.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.h"
#include "kuhn_tables.h"

namespace mm {

// ------------------------------------------------------------------ reductions

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}

// Block-level reduction of NV values; thread 0 writes partials[i * kMaxPartials + block]
// with the linear block index of a 1-, 2- or 3-D grid.  Bit i of maxmask selects max
// (identity 0, used on absolute values) instead of sum.
template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], unsigned maxmask, double *partials) {
    __shared__ double sh[NV][kBlock / kWave];
    const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = ((maxmask >> i) & 1u) ? wave_max(v[i]) : wave_sum(v[i]);
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) sh[i][w] = v[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            double a = sh[i][0];
#pragma unroll
            for (int k = 1; k < kBlock / kWave; ++k)
                a = ((maxmask >> i) & 1u) ? fmax(a, sh[i][k]) : a + sh[i][k];
            partials[i * kMaxPartials + bid] = a;
        }
    }
}

static inline int grid_for(int64_t n, int per_thread = 1) {
    int64_t b = (n + (int64_t)kBlock * per_thread - 1) / ((int64_t)kBlock * per_thread);
    if (b < 1) b = 1;
    if (b > kMaxPartials) b = kMaxPartials;
    return (int)b;
}

// The total of n partials in k_fin's order (per thread over j = t, t + 256, ..., then the
// wave, then the waves in order), in every thread of a kFinThreads-thread block; sh holds
// kFinThreads / kWave doubles.  Contains a barrier.
__device__ double fin_order_sum(const double *__restrict__ parts, int n, double *sh) {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    double acc = 0.0;
    for (int j = threadIdx.x; j < n; j += kBlock) acc = acc + parts[j];
    acc = wave_sum(acc);
    if (lane == 0) sh[w] = acc;
    __syncthreads();
    double t = sh[0];
    for (int k = 1; k < kBlock / kWave; ++k) t = t + sh[k];
    return t;
}

// One block: totals of the partials, then the scalar logic selected by op.  All values'
// partials are loaded in one pass (each value summed in the same order as one value at a
// time: per thread over j = t, t + 256, ..., then the wave, then the waves in order), one
// barrier.
constexpr int kFinThreads = kBlock;
__global__ __launch_bounds__(kFinThreads) void k_fin(const double *__restrict__ partials, FinArgs a,
                                                     DevScalars *dsc) {
    if (a.skip && *a.skip) return;
    __shared__ double sh[4][kFinThreads / kWave];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int j = threadIdx.x; j < a.nparts; j += kFinThreads) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (i >= a.nv) break;
            const double p = partials[i * kMaxPartials + j];
            acc[i] = ((a.maxmask >> i) & 1u) ? fmax(acc[i], p) : acc[i] + p;
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        acc[i] = ((a.maxmask >> i) & 1u) ? wave_max(acc[i]) : wave_sum(acc[i]);
        if (lane == 0) sh[i][w] = acc[i];
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double tot[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = 0; i < a.nv; ++i) {
        const bool mx = (a.maxmask >> i) & 1u;
        double t = sh[i][0];
        for (int k = 1; k < kFinThreads / kWave; ++k) t = mx ? fmax(t, sh[i][k]) : t + sh[i][k];
        tot[i] = t;
    }
    switch (a.op) {
        case FIN_STORE:
            for (int i = 0; i < a.nv; ++i) a.out[i] = tot[i];
            break;
        case FIN_TCG_INIT: {  // tot = (g.g, r.z)
            const double nrm = sqrt(tot[0]);
            dsc->tcg_eps = fmin(0.5, sqrt(nrm) * nrm);
            dsc->tcg_rz = tot[1];
            dsc->tcg_rzp[0] = tot[1];
            dsc->tcg_done = 0;
            dsc->tcg_neg = 0;
            dsc->tcg_iter = 0;
            break;
        }
        case FIN_TCG_UPDATE: {  // tot = (r'.r', r'.z')
            const bool neg = dsc->tcg_den <= 0.0;
            const bool conv = sqrt(tot[0]) <= dsc->tcg_eps;
            const double rz = dsc->tcg_rz;
            dsc->tcg_beta = (!neg && !conv) ? tot[1] / (rz + 1e-16) : 0.0;
            if (!neg) dsc->tcg_rz = tot[1];
            dsc->tcg_neg = neg;
            dsc->tcg_done = neg || conv;
            dsc->tcg_iter += 1;
            dsc->tcg_total += 1;
            break;
        }
        case FIN_RU_INIT:  // tot = (r.r FP64, r32.z32, b.b): fixes the bound
        case FIN_RU_REPLACE: {
            if (a.op == FIN_RU_INIT) {
                const double bound = fmax(a.tol * a.tol * tot[2], a.atol * a.atol);
                dsc->ir_bound = bound;
                dsc->pcg_bound = bound;
                dsc->pcg_iter = 0;
                dsc->ru_maxiter = a.maxiter;
                dsc->ir_outer = 0;
            }
            dsc->ir_rr = tot[0];
            dsc->ir_rmax2 = tot[0];
            dsc->ir_done = tot[0] <= dsc->ir_bound;
            if (a.delta2 > 0) dsc->pcg_bound = fmax(dsc->ir_bound, a.delta2 * tot[0]);
            dsc->ir_outer += 1;
            dsc->ru_gamma = tot[1];
            dsc->pcg_rr = tot[0];
            dsc->pcg_done = dsc->ir_done || dsc->pcg_iter >= dsc->ru_maxiter;
            break;
        }
    }
}

void launch_fin(const Launch &L, const FinArgs &a) {
    hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinThreads), 0, L.s, L.red.partials, a, L.dsc);
}

// The device scalars copied word by word into the host mirror by one wave; the wave's
// system-scope fence makes the copy visible before lane 0 writes the sequence number.
constexpr int kScalarWords = (int)(sizeof(DevScalars) / 8);
static_assert(sizeof(DevScalars) % 8 == 0 && kScalarWords <= 2 * kWave, "two words per lane at most");
__global__ __launch_bounds__(kWave) void k_publish(const unsigned long long *__restrict__ d,
                                                   unsigned long long *h, unsigned long long *hseq,
                                                   unsigned long long seq) {
    for (int w = threadIdx.x; w < kScalarWords; w += kWave) h[w] = d[w];
    __threadfence_system();
    if (threadIdx.x == 0) __hip_atomic_store(hseq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}
void launch_publish(const Launch &L, const DevScalars *d, DevScalars *h, unsigned long long *hseq,
                    unsigned long long seq) {
    hipLaunchKernelGGL(k_publish, dim3(1), dim3(kWave), 0, L.s, (const unsigned long long *)d,
                       (unsigned long long *)h, hseq, seq);
}

// ------------------------------------------------------------------ SELL SpMV

template <typename T>
__global__ __launch_bounds__(kBlock) void k_sell_spmv(SellDev<T> A, const T *__restrict__ x,
                                                      T *__restrict__ y, int want_dot,
                                                      const int *skip, double *partials) {
    if (skip && *skip) return;
    const int lane = threadIdx.x & (kWave - 1);
    double dot = 0.0;
    for (int slice = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; slice < A.n_slices;
         slice += gridDim.x * (kBlock / kWave)) {
        const int beg = A.slice_ptr[slice], end = A.slice_ptr[slice + 1];
        const int row = slice * kWave + lane;
        T acc = 0;
        for (int k = beg + lane; k < end; k += kWave) acc += A.val[k] * x[A.col[k]];
        if (row < A.n_rows) {
            y[row] = acc;
            dot += (double)acc * (double)x[row];
        }
    }
    if (want_dot) {
        double v[1] = {dot};
        block_partials<1>(v, 0u, partials);
    }
}

template <typename T>
static int sell_blocks(const SellDev<T> &A) {
    return std::max(1, std::min((A.n_slices + kBlock / kWave - 1) / (kBlock / kWave), kMaxPartials));
}

void launch_sell_spmv(const Launch &L, const SellDev<double> &A, const double *x, double *y,
                      double *dot_out, const int *skip_flag) {
    const int nb = sell_blocks(A);
    hipLaunchKernelGGL(k_sell_spmv<double>, dim3(nb), dim3(kBlock), 0, L.s, A, x, y, dot_out != nullptr,
                       skip_flag, L.red.partials);
    if (dot_out) launch_fin(L, FinArgs::store(dot_out, 1, 0u, nb, skip_flag));
}
void launch_sell_spmv32(const Launch &L, const SellDev<float> &A, const float *x, float *y,
                        double *dot_out, const int *skip_flag) {
    const int nb = sell_blocks(A);
    hipLaunchKernelGGL(k_sell_spmv<float>, dim3(nb), dim3(kBlock), 0, L.s, A, x, y, dot_out != nullptr,
                       skip_flag, L.red.partials);
    if (dot_out) launch_fin(L, FinArgs::store(dot_out, 1, 0u, nb, skip_flag));
}

// ------------------------------------------------------------------ 7-point Poisson stencil
// On the Kuhn-split rectilinear grid every tetrahedron couples its nodes only along
// axis-aligned edges (the barycentric gradients of non-consecutive path vertices are
// orthogonal).  The edge (P, P + e_d) lies in 4 voxels; in each it belongs to 2 Kuhn
// tetrahedra if its two transverse local coordinates are equal, else to 1, and each
// contributes -vol_t / h_d^2 = -(h_d h_d1 h_d2 / 6) / h_d^2.  The diagonal is minus the
// sum of the six couplings; Dirichlet rows (the box faces) are identity and couplings
// to Dirichlet nodes are dropped (escript nullifyRowsAndCols semantics).
//
// Voxel sizes are stored padded with a zero at both ends (hp[i] = size of voxel i-1),
// and ip[i] = 1 / (6 hp[i]) (0 for the pads), so no branch is needed for missing
// voxels.  Blocks are 64 x 4 tiles of an (x, y) plane that march through a contiguous
// range of z planes, so the z neighbours come from cache.

template <typename T>
__device__ __forceinline__ T edge_sum(const T *__restrict__ h1p, const T *__restrict__ h2p, int a, int b) {
    // transverse voxels a-1 / a (local coordinate 1 / 0) and b-1 / b; 2 tets when equal
    const T l1 = h1p[a], u1 = h1p[a + 1], l2 = h2p[b], u2 = h2p[b + 1];
    return (T)2 * l1 * l2 + l1 * u2 + u1 * l2 + (T)2 * u1 * u2;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_grid_poisson(PoissonGrid<T> P, const T *__restrict__ x,
                                                         T *__restrict__ y, int want_dot,
                                                         const int *skip, double *partials) {
    if (skip && *skip) return;
    const int nx = P.nx, ny = P.ny, nz = P.nz;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int chunk = (nz + gridDim.z - 1) / gridDim.z;
    const int k0 = blockIdx.z * chunk, k1 = min(nz, k0 + chunk);
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    double dot = 0.0;
    // the last chunks can be empty (k0 >= nz) when gridDim.z * chunk > nz
    if (i < nx && j < ny && k0 < k1) {
        const bool edge_ij = i == 0 || j == 0 || i == nx - 1 || j == ny - 1;
        // x-edge sums depend on (j, k), y-edge sums on (k, i), z-edge sums on (i, j)
        const T sz_ij = edge_sum(P.hxp, P.hyp, i, j);
        const T mxm = (i - 1 > 0) ? (T)1 : (T)0, mxp = (i + 1 < nx - 1) ? (T)1 : (T)0;
        const T mym = (j - 1 > 0) ? (T)1 : (T)0, myp = (j + 1 < ny - 1) ? (T)1 : (T)0;
        const T ixm = P.ixp[i], ixp = P.ixp[i + 1], iym = P.iyp[j], iyp = P.iyp[j + 1];
        const int64_t base = i + sy * j;
        // sliding window along z: (k-1, k, k+1) values of this column stay in registers
        T xm = k0 > 0 ? x[base + sz * (k0 - 1)] : (T)0;
        T xc = x[base + sz * k0];
        for (int k = k0; k < k1; ++k) {
            const int64_t idx = base + sz * k;
            const T xp = k + 1 < nz ? x[idx + sz] : (T)0;
            T out;
            if (edge_ij || k == 0 || k == nz - 1) {
                out = xc;  // Dirichlet row
            } else {
                const T xw = x[idx - 1], xe = x[idx + 1], xs = x[idx - sy], xn = x[idx + sy];
                const T sx = edge_sum(P.hyp, P.hzp, j, k);
                const T syy = edge_sum(P.hzp, P.hxp, k, i);
                const T cxm = -sx * ixm, cxp = -sx * ixp;
                const T cym = -syy * iym, cyp = -syy * iyp;
                const T czm = -sz_ij * P.izp[k], czp = -sz_ij * P.izp[k + 1];
                const T diag = -(cxm + cxp + cym + cyp + czm + czp);
                const T mzm = (k - 1 > 0) ? (T)1 : (T)0, mzp = (k + 1 < nz - 1) ? (T)1 : (T)0;
                out = diag * xc + mxm * cxm * xw + mxp * cxp * xe + mym * cym * xs + myp * cyp * xn +
                      mzm * czm * xm + mzp * czp * xp;
            }
            y[idx] = out;
            dot += (double)out * (double)xc;
            xm = xc;
            xc = xp;
        }
    }
    if (want_dot) {
        double v[1] = {dot};
        block_partials<1>(v, 0u, partials);
    }
}

template <typename T>
static dim3 poisson_grid(const PoissonGrid<T> &P) {
    const int gx = (P.nx + 63) / 64, gy = (P.ny + 3) / 4;
    int gz = std::max(1, std::min(P.nz, kMaxPartials / std::max(1, gx * gy)));
    while (gx * gy * gz > kMaxPartials && gz > 1) --gz;
    return dim3(gx, gy, gz);
}
template <typename T>
static int nblocks(const dim3 &g) {
    return (int)(g.x * g.y * g.z);
}

void launch_grid_poisson(const Launch &L, const PoissonGrid<double> &P, const double *x, double *y,
                         double *dot_out, const int *skip_flag) {
    const dim3 g = poisson_grid(P);
    hipLaunchKernelGGL(k_grid_poisson<double>, g, dim3(kBlock), 0, L.s, P, x, y, dot_out != nullptr,
                       skip_flag, L.red.partials);
    if (dot_out) launch_fin(L, FinArgs::store(dot_out, 1, 0u, nblocks<double>(g), skip_flag));
}
void launch_grid_poisson32(const Launch &L, const PoissonGrid<float> &P, const float *x, float *y,
                           double *dot_out, const int *skip_flag) {
    const dim3 g = poisson_grid(P);
    hipLaunchKernelGGL(k_grid_poisson<float>, g, dim3(kBlock), 0, L.s, P, x, y, dot_out != nullptr,
                       skip_flag, L.red.partials);
    if (dot_out) launch_fin(L, FinArgs::store(dot_out, 1, 0u, nblocks<float>(g), skip_flag));
}

// ------------------------------------------------------------------ reliable updates (mixed precision)
// Mixed precision, reliable updates: the FP32 iteration runs on the correction dx;
// the true residual is recomputed in FP64 from x whenever the FP32 recursive residual
// has dropped by a factor, keeping the search direction p.
//   r = b - Ax (FP64), r32 = (float) r, z32 = r32 / D (as r32 * invD32); with INIT also
//   p32 = z32.  Partials: (r.r in FP64, r32.z32, b.b).
template <bool INIT>
__global__ __launch_bounds__(kBlock) void k_ru_residual(int64_t n, const double *__restrict__ b,
                                                        const double *__restrict__ Ax,
                                                        float *__restrict__ r32, float *__restrict__ z32,
                                                        float *__restrict__ p32,
                                                        const float *__restrict__ invD32,
                                                        const int *skip, double *partials) {
    if (skip && *skip) return;
    double s[3] = {0.0, 0.0, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const double bi = b[i];
        const double ri = bi - Ax[i];
        const float rf = (float)ri;
        const float zf = rf * invD32[i];
        r32[i] = rf;
        z32[i] = zf;
        if (INIT) p32[i] = zf;
        s[0] += ri * ri;
        s[1] += (double)rf * (double)zf;
        s[2] += bi * bi;
    }
    block_partials<3>(s, 0u, partials);
}
// The finalize of a true residual from nb blocks' partials (r.r FP64, r32.z32, b.b).
static void launch_ru_fin(const Launch &L, int nb, double tol, double atol, int maxiter, bool init,
                          double delta2) {
    const int *skip = init ? nullptr : &L.dsc->ir_done;
    FinArgs a{};
    a.op = init ? FIN_RU_INIT : FIN_RU_REPLACE;
    a.nv = 3;
    a.nparts = nb;
    a.tol = tol;
    a.atol = atol;
    a.maxiter = maxiter;
    a.delta2 = delta2;
    a.skip = skip;
    launch_fin(L, a);
}
void launch_ru_residual(const Launch &L, int64_t n, const double *b, const double *Ax, float *r32,
                        float *z32, float *p32, const float *invD32, double tol, double atol,
                        int maxiter, bool init, double delta2) {
    const int nb = grid_for(n, 4);
    const int *skip = init ? nullptr : &L.dsc->ir_done;
    if (init)
        hipLaunchKernelGGL(k_ru_residual<true>, dim3(nb), dim3(kBlock), 0, L.s, n, b, Ax, r32, z32, p32,
                           invD32, skip, L.red.partials);
    else
        hipLaunchKernelGGL(k_ru_residual<false>, dim3(nb), dim3(kBlock), 0, L.s, n, b, Ax, r32, z32, p32,
                           invD32, skip, L.red.partials);
    launch_ru_fin(L, nb, tol, atol, maxiter, init, delta2);
}

// The true residual on the structured grid in one pass: K x formed as k_grid_poisson forms
// it (same tiles, same expressions, the edge sums written out), r = b - K x, r32,
// z32 = r32 invD32 (INIT: p32 = z32) and the partials (r.r FP64, r32.z32, b.b) of
// k_ru_residual; skipped when *skip.  Each plane's loads are issued one plane ahead
// (in-plane neighbours clamped to the node itself on Dirichlet rows, where they are
// unused).
template <bool INIT>
__global__ __launch_bounds__(kBlock) void k_grid_ru_residual(PoissonGrid<double> P, const double *__restrict__ x,
                                                             const double *__restrict__ b,
                                                             float *__restrict__ r32, float *__restrict__ z32,
                                                             float *__restrict__ p32,
                                                             const float *__restrict__ invD32, const int *skip,
                                                             double *partials) {
    if (skip && *skip) return;
    const int nx = P.nx, ny = P.ny, nz = P.nz;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int chunk = (nz + gridDim.z - 1) / gridDim.z;
    const int k0 = blockIdx.z * chunk, k1 = min(nz, k0 + chunk);
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    double s[3] = {0.0, 0.0, 0.0};
    if (i < nx && j < ny && k0 < k1) {
        const bool edge_ij = i == 0 || j == 0 || i == nx - 1 || j == ny - 1;
        const double sz_ij = edge_sum(P.hxp, P.hyp, i, j);
        const double mxm = (i - 1 > 0) ? 1.0 : 0.0, mxp = (i + 1 < nx - 1) ? 1.0 : 0.0;
        const double mym = (j - 1 > 0) ? 1.0 : 0.0, myp = (j + 1 < ny - 1) ? 1.0 : 0.0;
        const double ixm = P.ixp[i], ixp = P.ixp[i + 1], iym = P.iyp[j], iyp = P.iyp[j + 1];
        const double hxl = P.hxp[i], hxu = P.hxp[i + 1], hyl = P.hyp[j], hyu = P.hyp[j + 1];
        const int64_t base = i + sy * j;
        // plane k: x at k + 1, its in-plane neighbours, b, invD, hz and iz at k and k + 1
        // (the spacing arrays hold n + 1 entries)
        struct Plane {
            double xp, xw, xe, xs, xn, bi, hzl, hzu, izl, izu;
            float iv;
        };
        auto load = [&](int k) {
            Plane p;
            const int64_t idx = base + sz * k;
            const bool in = !(edge_ij || k == 0 || k == nz - 1);
            const int64_t ox = in ? 1 : 0, oy = in ? sy : 0;
            p.xp = x[k + 1 < nz ? idx + sz : idx];
            if (k + 1 >= nz) p.xp = 0.0;
            p.xw = x[idx - ox];
            p.xe = x[idx + ox];
            p.xs = x[idx - oy];
            p.xn = x[idx + oy];
            p.bi = b[idx];
            p.iv = invD32[idx];
            p.hzl = P.hzp[k];
            p.hzu = P.hzp[k + 1];
            p.izl = P.izp[k];
            p.izu = P.izp[k + 1];
            return p;
        };
        double xm = k0 > 0 ? x[base + sz * (k0 - 1)] : 0.0;
        double xc = x[base + sz * k0];
        Plane cur = load(k0);
        for (int k = k0; k < k1; ++k) {
            const int64_t idx = base + sz * k;
            Plane nxt = cur;
            if (k + 1 < k1) nxt = load(k + 1);
            const double xp = cur.xp;
            double out;
            if (edge_ij || k == 0 || k == nz - 1) {
                out = xc;  // Dirichlet row
            } else {
                // edge_sum(hy, hz; j, k) and edge_sum(hz, hx; k, i)
                const double sx = 2.0 * hyl * cur.hzl + hyl * cur.hzu + hyu * cur.hzl + 2.0 * hyu * cur.hzu;
                const double syy = 2.0 * cur.hzl * hxl + cur.hzl * hxu + cur.hzu * hxl + 2.0 * cur.hzu * hxu;
                const double cxm = -sx * ixm, cxp = -sx * ixp;
                const double cym = -syy * iym, cyp = -syy * iyp;
                const double czm = -sz_ij * cur.izl, czp = -sz_ij * cur.izu;
                const double diag = -(cxm + cxp + cym + cyp + czm + czp);
                const double mzm = (k - 1 > 0) ? 1.0 : 0.0, mzp = (k + 1 < nz - 1) ? 1.0 : 0.0;
                out = diag * xc + mxm * cxm * cur.xw + mxp * cxp * cur.xe + mym * cym * cur.xs +
                      myp * cyp * cur.xn + mzm * czm * xm + mzp * czp * xp;
            }
            const double bi = cur.bi;
            const double ri = bi - out;
            const float rf = (float)ri;
            const float zf = rf * cur.iv;
            r32[idx] = rf;
            z32[idx] = zf;
            if (INIT) p32[idx] = zf;
            s[0] += ri * ri;
            s[1] += (double)rf * (double)zf;
            s[2] += bi * bi;
            xm = xc;
            xc = xp;
            cur = nxt;
        }
    }
    block_partials<3>(s, 0u, partials);
}

bool grid_ru_residual_fits(const PoissonGrid<double> &P) {
    return nblocks<double>(poisson_grid(P)) <= kMaxPartials;
}
void launch_grid_ru_residual(const Launch &L, const PoissonGrid<double> &P, const double *x, const double *b,
                             float *r32, float *z32, float *p32, const float *invD32, double tol, double atol,
                             int maxiter, bool init, double delta2) {
    const dim3 g = poisson_grid(P);
    const int *skip = init ? nullptr : &L.dsc->ir_done;
    if (init)
        hipLaunchKernelGGL(k_grid_ru_residual<true>, g, dim3(kBlock), 0, L.s, P, x, b, r32, z32, p32, invD32,
                           skip, L.red.partials);
    else
        hipLaunchKernelGGL(k_grid_ru_residual<false>, g, dim3(kBlock), 0, L.s, P, x, b, r32, z32, p32, invD32,
                           skip, L.red.partials);
    launch_ru_fin(L, nblocks<double>(g), tol, atol, maxiter, init, delta2);
}

// x += (double) dx; dx = 0.  Skipped once the true residual has converged.
__global__ __launch_bounds__(kBlock) void k_add_correction(int64_t n, double *__restrict__ x,
                                                           float *__restrict__ dx,
                                                           const DevScalars *dsc) {
    if (dsc->ir_done) return;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        x[i] += (double)dx[i];
        dx[i] = 0.0f;
    }
}
void launch_add_correction(const Launch &L, int64_t n, double *x, float *dx) {
    hipLaunchKernelGGL(k_add_correction, dim3(grid_for(n, 4)), dim3(kBlock), 0, L.s, n, x, dx, L.dsc);
}

// ------------------------------------------------------------------ element kernels (general meshes)
// Element e with nodes t0..t3, barycentric gradients G_k, volume vol, material m:
//   exchange:    y_k,a += (2A/V) vol  G_k . grad(v_a)
//   anisotropy:  y_k   += (-2K/V) (vol/20) u (u.v_k + sum_l u.v_l)
//   divergence:  b_k   += Js vol G_k . mean(v)
//   gradient:    g_k   += Js vol/(4V) grad(u)        (same for the 4 nodes)
// The first three are assembled per node on the host (k_pair_c below).  The gradient runs
// element by element: each element writes its vector once and a node-centric pass sums
// them in a fixed order, so the result is the same in every run (atomic accumulation
// would depend on the order of arrival).

struct ElemLocal {
    double G[4][3];
};

__device__ __forceinline__ void load_G(const ElemDev &E, int e, ElemLocal &L) {
    const int n = E.n;
#pragma unroll
    for (int k = 1; k < 4; ++k)
#pragma unroll
        for (int a = 0; a < 3; ++a) L.G[k][a] = E.G[(3 * (k - 1) + a) * (int64_t)n + e];
#pragma unroll
    for (int a = 0; a < 3; ++a) L.G[0][a] = -(L.G[1][a] + L.G[2][a] + L.G[3][a]);
}

__global__ __launch_bounds__(kBlock) void k_ebe_grad(ElemDev E, const MatCoef *__restrict__ mats,
                                                     const double *__restrict__ u,
                                                     double *__restrict__ cg3, double invV) {
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= E.n) return;
    const int4 t = E.tets[e];
    const int tn[4] = {t.x, t.y, t.z, t.w};
    ElemLocal L;
    load_G(E, e, L);
    const double cg = mats[E.mat[e]].Js * E.vol[e] * invV * 0.25;
    double ul[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) ul[k] = u[tn[k]];
    double gu[3];
#pragma unroll
    for (int c = 0; c < 3; ++c)
        gu[c] = cg * (L.G[0][c] * ul[0] + L.G[1][c] * ul[1] + L.G[2][c] * ul[2] + L.G[3][c] * ul[3]);
#pragma unroll
    for (int a = 0; a < 3; ++a) cg3[3 * (int64_t)e + a] = gu[a];
}

// Node-centric sums of the element contributions, in a fixed order: out[C i + a] = sum over
// the slots s of node i (4 e + k for its magnet elements e, ascending) of c[C (s / D) + a]
// (D = 1: one contribution per element and local node; D = 4: one per element, the same
// for its four nodes).  kGatherLanes lanes share a node: lane l takes the node's slots l,
// l + kGatherLanes, ..., then a butterfly over the lanes adds the partial sums.  Every
// node is written.
constexpr int kGatherLanes = 8;

template <int C, int D>
__global__ __launch_bounds__(kBlock) void k_ebe_gather(int64_t n, const int *__restrict__ ptr,
                                                       const int *__restrict__ slot, const double *__restrict__ c,
                                                       double *__restrict__ out, const int *skip) {
    if (skip && *skip) return;
    const int lane = threadIdx.x & (kGatherLanes - 1);
    const int64_t stride = (int64_t)gridDim.x * (kBlock / kGatherLanes);
    // every lane of a group holds the same i: the butterfly stays inside active groups
    for (int64_t i = blockIdx.x * (int64_t)(kBlock / kGatherLanes) + threadIdx.x / kGatherLanes; i < n;
         i += stride) {
        double acc[C];
#pragma unroll
        for (int a = 0; a < C; ++a) acc[a] = 0.0;
        const int q1 = ptr[i + 1];
        for (int q = ptr[i] + lane; q < q1; q += kGatherLanes) {
            const int64_t s = slot[q] / D;
#pragma unroll
            for (int a = 0; a < C; ++a) acc[a] += c[C * s + a];
        }
#pragma unroll
        for (int off = kGatherLanes / 2; off > 0; off >>= 1)
#pragma unroll
            for (int a = 0; a < C; ++a) acc[a] += __shfl_xor(acc[a], off, kGatherLanes);
        if (lane == 0)
#pragma unroll
            for (int a = 0; a < C; ++a) out[C * i + a] = acc[a];
    }
}

// Assembled C: kPairLanes lanes share a node; per row of the node the lanes stride over
// its entries (S v and M v in registers), a butterfly adds the lanes' sums, and every lane
// adds S v + u (u . M v) in row order (so lane 0 holds the node's sum in a fixed order).
// With DIV, b = D v from the same gathered v.
constexpr int kPairLanes = 8;

template <bool DIV>
__global__ __launch_bounds__(kBlock) void k_pair_c(PairCDev P, const MatCoef *__restrict__ mats,
                                                   const double *__restrict__ v, double *__restrict__ y,
                                                   double *__restrict__ b, const int *skip) {
    if (skip && *skip) return;
    const int lane = threadIdx.x & (kPairLanes - 1);
    const int64_t stride = (int64_t)gridDim.x * (kBlock / kPairLanes);
    for (int64_t i = blockIdx.x * (int64_t)(kBlock / kPairLanes) + threadIdx.x / kPairLanes; i < P.nn;
         i += stride) {
        double y3[3] = {0.0, 0.0, 0.0}, bi = 0.0;
        const int r1 = P.node_ptr[i + 1];
        for (int r = P.node_ptr[i]; r < r1; ++r) {   // the same rows for every lane of the group
            double sv[3] = {0.0, 0.0, 0.0}, mv[3] = {0.0, 0.0, 0.0}, dv = 0.0;
            const int q1 = P.row_ptr[r + 1];
            for (int q = P.row_ptr[r] + lane; q < q1; q += kPairLanes) {
                const int64_t j = P.col[q];
                const double2 w = P.sm[q];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const double vj = v[3 * j + a];
                    sv[a] += w.x * vj;
                    mv[a] += w.y * vj;
                    if (DIV) dv += P.d3[3 * (int64_t)q + a] * vj;
                }
            }
#pragma unroll
            for (int off = kPairLanes / 2; off > 0; off >>= 1) {
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    sv[a] += __shfl_xor(sv[a], off, kPairLanes);
                    mv[a] += __shfl_xor(mv[a], off, kPairLanes);
                }
                if (DIV) dv += __shfl_xor(dv, off, kPairLanes);
            }
            const double *u = mats[P.row_mat[r]].u;
            const double um = u[0] * mv[0] + u[1] * mv[1] + u[2] * mv[2];
#pragma unroll
            for (int a = 0; a < 3; ++a) y3[a] += sv[a] + u[a] * um;
            bi += dv;
        }
        if (lane == 0) {
#pragma unroll
            for (int a = 0; a < 3; ++a) y[3 * i + a] = y3[a];
            if (DIV) b[i] = bi;
        }
    }
}

static inline int gather_grid(int64_t nodes) {
    const int64_t b = (nodes * kGatherLanes + kBlock - 1) / kBlock;
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, 16384));
}

void launch_pair_c(const Launch &L, const PairCDev &P, const double *v, double *y, double *b, const int *skip) {
    const int64_t nb = (P.nn * kPairLanes + kBlock - 1) / kBlock;
    const int g = (int)std::max<int64_t>(1, std::min<int64_t>(nb, 16384));
    if (b)
        hipLaunchKernelGGL(k_pair_c<true>, dim3(g), dim3(kBlock), 0, L.s, P, L.mats, v, y, b, skip);
    else
        hipLaunchKernelGGL(k_pair_c<false>, dim3(g), dim3(kBlock), 0, L.s, P, L.mats, v, y, nullptr, skip);
}
void launch_ebe_grad(const Launch &L, const ElemDev &E, const double *u, double *g, double invV) {
    if (E.n > 0)
        hipLaunchKernelGGL(k_ebe_grad, dim3((E.n + kBlock - 1) / kBlock), dim3(kBlock), 0, L.s, E,
                           L.mats, u, E.cy, invV);
    hipLaunchKernelGGL((k_ebe_gather<3, 4>), dim3(gather_grid(E.nn)), dim3(kBlock), 0, L.s, E.nn, E.inc_ptr,
                       E.inc_slot, E.cy, g, nullptr);
}

// ------------------------------------------------------------------ structured magnet kernels
// Magnet block of mn^3 voxels (edge h) inside the rectilinear grid; magnet node
// (i,j,k), 0 <= i,j,k <= mn, is grid node (m0x+i, m0y+j, m0z+k).  Node-centric: each
// thread sums the contributions of the (up to) 8 voxels around its node with the exact
// Kuhn voxel tables (kuhn_tables.h); voxels outside the block use the all-zero
// material n_mats, so neighbour loads are clamped instead of branched.  No atomics.

__device__ __forceinline__ int64_t gnode(const GridDev &G, int i, int j, int k) {
    return (int64_t)(G.m0x + i) + (int64_t)G.nx * ((G.m0y + j) + (int64_t)G.ny * (G.m0z + k));
}
// Grid node of magnet-block node t = i + n1 (j + n1 k), n1 = mn + 1 (n1^3 < 2^32)
__device__ __forceinline__ int64_t block_gnode(const GridDev &G, uint32_t t) {
    const uint32_t n1 = (uint32_t)G.mn + 1u;
    const uint32_t tj = t / n1, k = tj / n1;
    return gnode(G, (int)(t - tj * n1), (int)(tj - k * n1), (int)k);
}

// Material coefficients of one voxel as the node kernels use them
template <typename TC>
struct VoxCoef {
    TC cex, cK, Js, u0, u1, u2;
};

// (C v)_node from the values nv[(dx+1) + 3 (dy+1) + 9 (dz+1)][a] of v at the node's 27
// neighbours (clamped to the block: clamped entries only meet zero-material voxels) and
// vox(w), the coefficients of voxel (i-1+vx, j-1+vy, k-1+vz), w = vx + 2 vy + 4 vz; with
// DIV also the node's divergence term sum Js (B v).  Unscaled: y = fS ex + fM an.
template <typename TC, bool DIV, typename Vox>
__device__ __forceinline__ void grid_c_core(const GridDev &G, const TC (&nv)[27][3], Vox vox, TC (&y)[3],
                                            TC &bdiv) {
    TC ex[3] = {0, 0, 0}, an[3] = {0, 0, 0}, bd = 0;
#pragma unroll
    for (int w = 0; w < 8; ++w) {
        const int vx = w & 1, vy = (w >> 1) & 1, vz = (w >> 2) & 1;
        const VoxCoef<TC> mc = vox(w);
        const int p = (1 - vx) + 2 * (1 - vy) + 4 * (1 - vz);  // corner of the node in this voxel
        TC sS[3] = {0, 0, 0}, sM = 0, sB = 0;
        const TC u0 = mc.u0, u1 = mc.u1, u2 = mc.u2;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int cx = c & 1, cy = (c >> 1) & 1, cz = (c >> 2) & 1;
            const int s = (cx - (1 - vx) + 1) + 3 * (cy - (1 - vy) + 1) + 9 * (cz - (1 - vz) + 1);
            const int KS = kKS[p][c], KM = kKM[p][c];
            if (KS != 0) {
#pragma unroll
                for (int a = 0; a < 3; ++a) sS[a] += (TC)KS * nv[s][a];
            }
            if (KM != 0) sM += (TC)KM * (u0 * nv[s][0] + u1 * nv[s][1] + u2 * nv[s][2]);
            if (DIV) {
                const int B0 = kKB[p][c][0], B1 = kKB[p][c][1], B2 = kKB[p][c][2];
                if (B0 != 0) sB += (TC)B0 * nv[s][0];
                if (B1 != 0) sB += (TC)B1 * nv[s][1];
                if (B2 != 0) sB += (TC)B2 * nv[s][2];
            }
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) ex[a] += mc.cex * sS[a];
        an[0] += mc.cK * sM * u0;
        an[1] += mc.cK * sM * u1;
        an[2] += mc.cK * sM * u2;
        if (DIV) bd += mc.Js * sB;
    }
    const TC fS = (TC)G.fS, fM = (TC)G.fM;
#pragma unroll
    for (int a = 0; a < 3; ++a) y[a] = fS * ex[a] + fM * an[a];
    if (DIV) bdiv = (TC)G.fB * bd;
}

template <typename TC>
__device__ __forceinline__ VoxCoef<TC> vox_coef(const MatCoef &mc) {
    return VoxCoef<TC>{(TC)mc.cex, (TC)mc.cK, (TC)mc.Js, (TC)mc.u[0], (TC)mc.u[1], (TC)mc.u[2]};
}

// grid_c_core with the voxel materials read from the block's material ids
template <typename TC, bool DIV>
__device__ __forceinline__ void grid_c_node(const GridDev &G, const MatCoef *__restrict__ mats,
                                            const TC (&nv)[27][3], int i, int j, int k, TC (&y)[3], TC &bdiv) {
    auto vox = [&](int w) {
        const int xi = i - 1 + (w & 1), yj = j - 1 + ((w >> 1) & 1), zk = k - 1 + ((w >> 2) & 1);
        const bool inside = xi >= 0 && yj >= 0 && zk >= 0 && xi < G.mn && yj < G.mn && zk < G.mn;
        const int mat = inside ? (int)G.vmat[xi + (int64_t)G.mn * (yj + (int64_t)G.mn * zk)] : G.n_mats;
        return vox_coef<TC>(mats[mat]);
    };
    grid_c_core<TC, DIV>(G, nv, vox, y, bdiv);
}

// y = C v at magnet nodes; with DIV also b = Dx vx + Dy vy + Dz vz at magnet nodes.
// A block owns a kCTX x kCTY tile of block nodes and marches a z chunk: v (in TC) is
// staged once per node into a three-plane LDS ring with a one-node halo (clamped to the
// block, as the 27-point gather clamps), the voxels' coefficients into a two-plane ring;
// each node then reads its 27 neighbours and 8 voxels from LDS.
constexpr int kCTX = 32, kCTY = 8;
struct TcgTiles {
    int gx, gy, gz;
};
TcgTiles block_tiles(const GridDev &G) {
    const int n1 = G.mn + 1;
    TcgTiles tl;
    tl.gx = (n1 + kCTX - 1) / kCTX;
    tl.gy = (n1 + kCTY - 1) / kCTY;
    tl.gz = std::max(1, std::min(n1, std::min(1024, kMaxPartials) / (tl.gx * tl.gy)));
    return tl;
}
bool tcg_tiles_fit(int mn) {   // the truncated CG's d'.H d' writes one partial per tile
    GridDev G;
    G.mn = mn;
    const TcgTiles tl = block_tiles(G);
    return (int64_t)tl.gx * tl.gy * tl.gz <= kMaxPartials;
}
template <typename TC, bool DIV>
__global__ __launch_bounds__(kBlock) void k_grid_exani(GridDev G, TcgTiles tl, const MatCoef *__restrict__ mats,
                                                       const double *__restrict__ v,
                                                       double *__restrict__ y, double *__restrict__ b,
                                                       const int *skip) {
    static_assert(kCTX * kCTY == kBlock, "one node per thread and plane");
    constexpr int NV = DIV ? 6 : 5;
    __shared__ TC nd[3][kCTY + 2][kCTX + 2][3];
    __shared__ TC vc[2][kCTY + 1][kCTX + 1][NV];
    if (skip && *skip) return;   // block-uniform
    const int mn = G.mn, n1 = mn + 1;
    const int bl = blockIdx.x;
    const int cz = bl % tl.gz, bx = (bl / tl.gz) % tl.gx, by = bl / (tl.gz * tl.gx);
    const int tx = threadIdx.x % kCTX, ty = threadIdx.x / kCTX;
    const int i0 = bx * kCTX, j0 = by * kCTY, i = i0 + tx, j = j0 + ty;
    const int chunk = (n1 + tl.gz - 1) / tl.gz;
    const int k0 = cz * chunk, k1 = min(n1, k0 + chunk);
    auto ring3 = [](int p) { return ((p % 3) + 3) % 3; };
    // A thread's ring entries: nodes e = t, t + kBlock < (kCTY + 2)(kCTX + 2), voxels
    // e = t, t + kBlock < (kCTY + 1)(kCTX + 1).  Plane k + 2's node values and plane
    // k + 1's voxel materials are loaded into registers before plane k is computed and
    // written to the rings at the next step.
    constexpr int kNodeE = (kCTY + 2) * (kCTX + 2), kVoxE = (kCTY + 1) * (kCTX + 1);
    static_assert(kNodeE <= 2 * kBlock && kVoxE <= 2 * kBlock, "two ring entries per thread");
    struct NodeRegs {
        TC v[2][3];
    };
    struct VoxRegs {
        int mat[2];
    };
    auto fetch_nodes = [&](int kp) {
        NodeRegs r{};
        const int kk = min(max(kp, 0), mn);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int e = threadIdx.x + q * kBlock;
            if (e >= kNodeE) break;
            const int ex = e % (kCTX + 2), ey = e / (kCTX + 2);
            const int ii = min(max(i0 - 1 + ex, 0), mn), jj = min(max(j0 - 1 + ey, 0), mn);
            const int64_t g = gnode(G, ii, jj, kk);
#pragma unroll
            for (int a = 0; a < 3; ++a) r.v[q][a] = (TC)v[3 * g + a];
        }
        return r;
    };
    auto store_nodes = [&](const NodeRegs &r, int kp) {
        const int slot = ring3(kp);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int e = threadIdx.x + q * kBlock;
            if (e >= kNodeE) break;
            const int ex = e % (kCTX + 2), ey = e / (kCTX + 2);
#pragma unroll
            for (int a = 0; a < 3; ++a) nd[slot][ey][ex][a] = r.v[q][a];
        }
    };
    auto fetch_voxels = [&](int kz) {
        VoxRegs r{};
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int e = threadIdx.x + q * kBlock;
            if (e >= kVoxE) break;
            const int ex = e % (kCTX + 1), ey = e / (kCTX + 1);
            const int xi = i0 - 1 + ex, yj = j0 - 1 + ey;
            const bool inside = xi >= 0 && yj >= 0 && kz >= 0 && xi < mn && yj < mn && kz < mn;
            r.mat[q] = inside ? (int)G.vmat[xi + (int64_t)mn * (yj + (int64_t)mn * kz)] : G.n_mats;
        }
        return r;
    };
    auto store_voxels = [&](const VoxRegs &r, int kz) {
        const int slot = kz & 1;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int e = threadIdx.x + q * kBlock;
            if (e >= kVoxE) break;
            const int ex = e % (kCTX + 1), ey = e / (kCTX + 1);
            const VoxCoef<TC> c = vox_coef<TC>(mats[r.mat[q]]);
            TC *s = vc[slot][ey][ex];
            s[0] = c.cex;
            s[1] = c.cK;
            s[2] = c.u0;
            s[3] = c.u1;
            s[4] = c.u2;
            if (DIV) s[NV - 1] = c.Js;
        }
    };
    if (k0 >= k1) return;   // block-uniform, before any barrier
    store_nodes(fetch_nodes(k0 - 1), k0 - 1);
    store_nodes(fetch_nodes(k0), k0);
    store_voxels(fetch_voxels(k0 - 1), k0 - 1);
    NodeRegs rn = fetch_nodes(k0 + 1);
    VoxRegs rv = fetch_voxels(k0);
    for (int k = k0; k < k1; ++k) {
        store_nodes(rn, k + 1);
        store_voxels(rv, k);
        if (k + 1 < k1) {
            rn = fetch_nodes(k + 2);
            rv = fetch_voxels(k + 1);
        }
        __syncthreads();
        if (i < n1 && j < n1) {
            TC nv[27][3];
#pragma unroll
            for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
                for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int sl = ring3(k + dz), s = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
#pragma unroll
                        for (int a = 0; a < 3; ++a) nv[s][a] = nd[sl][ty + 1 + dy][tx + 1 + dx][a];
                    }
            auto vox = [&](int w) {
                const int vx = w & 1, vy = (w >> 1) & 1, sl = (k - 1 + ((w >> 2) & 1)) & 1;
                const TC *q = vc[sl][ty + vy][tx + vx];
                return VoxCoef<TC>{q[0], q[1], DIV ? q[NV - 1] : (TC)0, q[2], q[3], q[4]};
            };
            TC yv[3], bd = 0;
            grid_c_core<TC, DIV>(G, nv, vox, yv, bd);
            const int64_t g = gnode(G, i, j, k);
#pragma unroll
            for (int a = 0; a < 3; ++a) y[3 * g + a] = (double)yv[a];
            if (DIV) b[g] = (double)bd;
        }
        __syncthreads();
    }
}

// g = interleave(Gx u, Gy u, Gz u) at magnet nodes.  Tiles of kCTX x kCTY nodes marching
// a z chunk, as k_grid_exani: u in a three-plane LDS ring with a one-node halo (clamped to
// the block), the voxels' Js in a two-plane ring, the next planes' loads in registers
// ahead of the compute; the 27 neighbours and 8 voxels then come from LDS.
template <typename TC>
__global__ __launch_bounds__(kBlock) void k_grid_grad(GridDev G, TcgTiles tl, const MatCoef *__restrict__ mats,
                                                      const double *__restrict__ u,
                                                      double *__restrict__ g) {
    static_assert(kCTX * kCTY == kBlock, "one node per thread and plane");
    __shared__ TC nd[3][kCTY + 2][kCTX + 2];
    __shared__ TC vj[2][kCTY + 1][kCTX + 1];
    const int mn = G.mn, n1 = mn + 1;
    const int bl = blockIdx.x;
    const int cz = bl % tl.gz, bx = (bl / tl.gz) % tl.gx, by = bl / (tl.gz * tl.gx);
    const int tx = threadIdx.x % kCTX, ty = threadIdx.x / kCTX;
    const int i0 = bx * kCTX, j0 = by * kCTY, i = i0 + tx, j = j0 + ty;
    const int chunk = (n1 + tl.gz - 1) / tl.gz;
    const int k0 = cz * chunk, k1 = min(n1, k0 + chunk);
    auto ring3 = [](int p) { return ((p % 3) + 3) % 3; };
    constexpr int kNodeE = (kCTY + 2) * (kCTX + 2), kVoxE = (kCTY + 1) * (kCTX + 1);
    static_assert(kNodeE <= 2 * kBlock && kVoxE <= 2 * kBlock, "two ring entries per thread");
    struct NodeRegs {
        TC v[2];
    };
    struct VoxRegs {
        int mat[2];
    };
    auto fetch_nodes = [&](int kp) {
        NodeRegs r{};
        const int kk = min(max(kp, 0), mn);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int e = threadIdx.x + q * kBlock;
            if (e >= kNodeE) break;
            const int ex = e % (kCTX + 2), ey = e / (kCTX + 2);
            const int ii = min(max(i0 - 1 + ex, 0), mn), jj = min(max(j0 - 1 + ey, 0), mn);
            r.v[q] = (TC)u[gnode(G, ii, jj, kk)];
        }
        return r;
    };
    auto store_nodes = [&](const NodeRegs &r, int kp) {
        const int slot = ring3(kp);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int e = threadIdx.x + q * kBlock;
            if (e >= kNodeE) break;
            nd[slot][e / (kCTX + 2)][e % (kCTX + 2)] = r.v[q];
        }
    };
    auto fetch_voxels = [&](int kz) {
        VoxRegs r{};
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int e = threadIdx.x + q * kBlock;
            if (e >= kVoxE) break;
            const int xi = i0 - 1 + e % (kCTX + 1), yj = j0 - 1 + e / (kCTX + 1);
            const bool inside = xi >= 0 && yj >= 0 && kz >= 0 && xi < mn && yj < mn && kz < mn;
            r.mat[q] = inside ? (int)G.vmat[xi + (int64_t)mn * (yj + (int64_t)mn * kz)] : G.n_mats;
        }
        return r;
    };
    auto store_voxels = [&](const VoxRegs &r, int kz) {
        const int slot = kz & 1;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int e = threadIdx.x + q * kBlock;
            if (e >= kVoxE) break;
            vj[slot][e / (kCTX + 1)][e % (kCTX + 1)] = (TC)mats[r.mat[q]].Js;
        }
    };
    if (k0 >= k1) return;   // block-uniform, before any barrier
    store_nodes(fetch_nodes(k0 - 1), k0 - 1);
    store_nodes(fetch_nodes(k0), k0);
    store_voxels(fetch_voxels(k0 - 1), k0 - 1);
    NodeRegs rn = fetch_nodes(k0 + 1);
    VoxRegs rv = fetch_voxels(k0);
    const TC f = (TC)G.fG;
    for (int k = k0; k < k1; ++k) {
        store_nodes(rn, k + 1);
        store_voxels(rv, k);
        if (k + 1 < k1) {
            rn = fetch_nodes(k + 2);
            rv = fetch_voxels(k + 1);
        }
        __syncthreads();
        if (i < n1 && j < n1) {
            TC nu[27];
#pragma unroll
            for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
                for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                    for (int dx = -1; dx <= 1; ++dx)
                        nu[(dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)] = nd[ring3(k + dz)][ty + 1 + dy][tx + 1 + dx];
            TC acc[3] = {0, 0, 0};
#pragma unroll
            for (int w = 0; w < 8; ++w) {
                const int vx = w & 1, vy = (w >> 1) & 1, vz = (w >> 2) & 1;
                const TC js = vj[(k - 1 + vz) & 1][ty + vy][tx + vx];
                const int p = (1 - vx) + 2 * (1 - vy) + 4 * (1 - vz);
                TC s[3] = {0, 0, 0};
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    const int cx = c & 1, cy = (c >> 1) & 1, cz2 = (c >> 2) & 1;
                    const int sidx = (cx - (1 - vx) + 1) + 3 * (cy - (1 - vy) + 1) + 9 * (cz2 - (1 - vz) + 1);
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        const int B = kKB[c][p][a];
                        if (B != 0) s[a] += (TC)B * nu[sidx];
                    }
                }
#pragma unroll
                for (int a = 0; a < 3; ++a) acc[a] += js * s[a];
            }
            const int64_t gn = gnode(G, i, j, k);
#pragma unroll
            for (int a = 0; a < 3; ++a) g[3 * gn + a] = (double)(f * acc[a]);
        }
        __syncthreads();
    }
}

void launch_grid_exani(const Launch &L, const GridDev &G, bool fp32, const double *v, double *y,
                       double *b, const int *skip_flag) {
    const TcgTiles tl = block_tiles(G);
    const dim3 g(tl.gx * tl.gy * tl.gz), blk(kBlock);
    if (fp32) {
        if (b)
            hipLaunchKernelGGL((k_grid_exani<float, true>), g, blk, 0, L.s, G, tl, L.mats, v, y, b, skip_flag);
        else
            hipLaunchKernelGGL((k_grid_exani<float, false>), g, blk, 0, L.s, G, tl, L.mats, v, y, b, skip_flag);
    } else {
        if (b)
            hipLaunchKernelGGL((k_grid_exani<double, true>), g, blk, 0, L.s, G, tl, L.mats, v, y, b, skip_flag);
        else
            hipLaunchKernelGGL((k_grid_exani<double, false>), g, blk, 0, L.s, G, tl, L.mats, v, y, b, skip_flag);
    }
}
void launch_grid_grad(const Launch &L, const GridDev &G, bool fp32, const double *u, double *g) {
    const TcgTiles tl = block_tiles(G);
    const dim3 gr(tl.gx * tl.gy * tl.gz), blk(kBlock);
    if (fp32)
        hipLaunchKernelGGL(k_grid_grad<float>, gr, blk, 0, L.s, G, tl, L.mats, u, g);
    else
        hipLaunchKernelGGL(k_grid_grad<double>, gr, blk, 0, L.s, G, tl, L.mats, u, g);
}

// ------------------------------------------------------------------ node kernels

__device__ __forceinline__ void cross3(const double *a, const double *b, double *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// g = gz + gC (+ gd); E = m.gz + m.gC/2 (+ m.gd/2); g_proj = m x (g x m)
// With a structured grid (G.nx > 0) the nodes outside the magnet block, where m, meas, gC
// and gd are zero, get zeros without being read.
__global__ __launch_bounds__(kBlock) void k_finish_eg(int64_t n, GridDev G, const double *__restrict__ m,
                                                      const double *__restrict__ meas, double zc0,
                                                      double zc1, double zc2,
                                                      const double *__restrict__ gC,
                                                      const double *__restrict__ gd,
                                                      double *__restrict__ g_full,
                                                      double *__restrict__ g_proj, int with_hmag,
                                                      int write_off, double *partials) {
    double e = 0.0;
    const double zc[3] = {zc0, zc1, zc2};
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        if (G.nx > 0) {   // 32-bit index arithmetic (a structured grid has fewer than 2^32 nodes)
            const uint32_t iu = (uint32_t)i, r = iu / (uint32_t)G.nx, k = r / (uint32_t)G.ny;
            const int gi = (int)(iu - r * (uint32_t)G.nx) - G.m0x, gj = (int)(r - k * (uint32_t)G.ny) - G.m0y,
                      gk = (int)k - G.m0z;
            if (gi < 0 || gj < 0 || gk < 0 || gi > G.mn || gj > G.mn || gk > G.mn) {
                if (write_off) {
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        g_full[3 * i + a] = 0.0;
                        g_proj[3 * i + a] = 0.0;
                    }
                }
                continue;
            }
        }
        const double mi = meas[i];
        double mm[3], g[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            mm[a] = m[3 * i + a];
            const double gz = zc[a] * mi;
            const double gc = gC[3 * i + a];
            const double gdm = with_hmag ? gd[3 * i + a] : 0.0;
            g[a] = gz + gc + gdm;
            e += mm[a] * gz + 0.5 * mm[a] * gc + 0.5 * mm[a] * gdm;
            g_full[3 * i + a] = g[a];
        }
        double c[3], pr[3];
        cross3(g, mm, c);
        cross3(mm, c, pr);
#pragma unroll
        for (int a = 0; a < 3; ++a) g_proj[3 * i + a] = pr[a];
    }
    double v[1] = {e};
    block_partials<1>(v, 0u, partials);
}

void launch_finish_eg(const Launch &L, int64_t n, const GridDev *G, const double *m, const double *meas,
                      double zc0, double zc1, double zc2, const double *gC, const double *gd, double *g_full,
                      double *g_proj, double *E_out, int with_hmag, bool write_off) {
    const int nb = grid_for(n, 2);
    const GridDev g = G ? *G : GridDev{};
    hipLaunchKernelGGL(k_finish_eg, dim3(nb), dim3(kBlock), 0, L.s, n, g, m, meas, zc0, zc1, zc2, gC, gd,
                       g_full, g_proj, with_hmag, (int)write_off, L.red.partials);
    launch_fin(L, FinArgs::store(E_out, 1, 0u, nb, nullptr));
}

__global__ __launch_bounds__(kBlock) void k_external(int64_t n, const double *__restrict__ m,
                                                     const double *__restrict__ meas, double zc0,
                                                     double zc1, double zc2,
                                                     double *__restrict__ g, double *partials) {
    double e = 0.0;
    const double zc[3] = {zc0, zc1, zc2};
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const double mi = meas[i];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double gz = zc[a] * mi;
            if (g) g[3 * i + a] = gz;
            e += m[3 * i + a] * gz;
        }
    }
    double v[1] = {e};
    block_partials<1>(v, 0u, partials);
}
void launch_external(const Launch &L, int64_t n, const double *m, const double *meas, double zc0,
                     double zc1, double zc2, double *g, double *E_out) {
    const int nb = grid_for(n, 2);
    hipLaunchKernelGGL(k_external, dim3(nb), dim3(kBlock), 0, L.s, n, m, meas, zc0, zc1, zc2, g,
                       L.red.partials);
    launch_fin(L, FinArgs::store(E_out, 1, 0u, nb, nullptr));
}

// H v = m0 x ((C v) x m0) - (m0 . gF)_i v_i ;  den = d . H d
__global__ __launch_bounds__(kBlock) void k_hvec(int64_t n, const double *__restrict__ m0,
                                                 const double *__restrict__ gF,
                                                 const double *__restrict__ Cd,
                                                 const double *__restrict__ d,
                                                 double *__restrict__ Hd, const int *skip,
                                                 double *partials) {
    if (skip && *skip) return;
    double den = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        double mm[3], cd[3], dd[3], c[3], pr[3];
        double mg = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            mm[a] = m0[3 * i + a];
            cd[a] = Cd[3 * i + a];
            dd[a] = d[3 * i + a];
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) mg += mm[a] * gF[3 * i + a];
        cross3(cd, mm, c);
        cross3(mm, c, pr);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double h = pr[a] - mg * dd[a];
            Hd[3 * i + a] = h;
            den += dd[a] * h;
        }
    }
    double v[1] = {den};
    block_partials<1>(v, 0u, partials);
}
void launch_hvec(const Launch &L, int64_t n, const double *m0, const double *gF, const double *Cd,
                 const double *d, double *Hd, double *den_out, const int *skip_flag) {
    const int nb = grid_for(n, 2);
    hipLaunchKernelGGL(k_hvec, dim3(nb), dim3(kBlock), 0, L.s, n, m0, gF, Cd, d, Hd, skip_flag,
                       L.red.partials);
    launch_fin(L, FinArgs::store(den_out, 1, 0u, nb, skip_flag));
}

// grad_fk = -g: p = 0, r = grad_fk, z = r / D, d = -z
__global__ __launch_bounds__(kBlock) void k_tcg_init(int64_t n3, const double *__restrict__ g,
                                                     const double *__restrict__ D,
                                                     double *__restrict__ p, double *__restrict__ r,
                                                     double *__restrict__ z, double *__restrict__ d,
                                                     double *partials) {
    double s[2] = {0.0, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n3; i += (int64_t)gridDim.x * kBlock) {
        const double ri = -g[i];
        const double zi = ri / D[i];
        p[i] = 0.0;
        r[i] = ri;
        z[i] = zi;
        d[i] = -zi;
        s[0] += ri * ri;
        s[1] += ri * zi;
    }
    block_partials<2>(s, 0u, partials);
}

__global__ __launch_bounds__(kBlock) void k_tcg_update(int64_t n3, double *__restrict__ p,
                                                       double *__restrict__ r,
                                                       double *__restrict__ z,
                                                       const double *__restrict__ d,
                                                       const double *__restrict__ Hd,
                                                       const double *__restrict__ D,
                                                       const DevScalars *dsc, double *partials) {
    if (dsc->tcg_done) return;
    const double den = dsc->tcg_den;
    const bool neg = den <= 0.0;
    const double alpha = neg ? 0.0 : dsc->tcg_rz / den;
    double s[2] = {0.0, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n3; i += (int64_t)gridDim.x * kBlock) {
        const double rn = r[i] + alpha * Hd[i];
        const double zn = rn / D[i];
        if (!neg) {
            p[i] = p[i] + alpha * d[i];
            r[i] = rn;
            z[i] = zn;
        }
        s[0] += rn * rn;
        s[1] += rn * zn;
    }
    block_partials<2>(s, 0u, partials);
}

__global__ __launch_bounds__(kBlock) void k_tcg_dupdate(int64_t n3, double *__restrict__ d,
                                                        const double *__restrict__ z,
                                                        const DevScalars *dsc) {
    if (dsc->tcg_done) return;
    const double beta = dsc->tcg_beta;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n3; i += (int64_t)gridDim.x * kBlock)
        d[i] = -z[i] + beta * d[i];
}

// z_out = (all p == 0) ? -grad_fk (= g) : p
__global__ __launch_bounds__(kBlock) void k_tcg_final(int64_t n3, const double *__restrict__ p,
                                                      const double *__restrict__ g,
                                                      double *__restrict__ zout,
                                                      const DevScalars *dsc) {
    const bool zero = dsc->tcg_pmax == 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n3; i += (int64_t)gridDim.x * kBlock)
        zout[i] = zero ? g[i] : p[i];
}

void launch_tcg_init(const Launch &L, int64_t n3, const double *g, const double *D, double *p,
                     double *r, double *z, double *d) {
    const int nb = grid_for(n3, 4);
    hipLaunchKernelGGL(k_tcg_init, dim3(nb), dim3(kBlock), 0, L.s, n3, g, D, p, r, z, d, L.red.partials);
    FinArgs a{};
    a.op = FIN_TCG_INIT;
    a.nv = 2;
    a.nparts = nb;
    launch_fin(L, a);
}
void launch_tcg_update(const Launch &L, int64_t n3, double *p, double *r, double *z, double *d,
                       const double *Hd, const double *D) {
    const int nb = grid_for(n3, 4);
    hipLaunchKernelGGL(k_tcg_update, dim3(nb), dim3(kBlock), 0, L.s, n3, p, r, z, d, Hd, D, L.dsc,
                       L.red.partials);
    FinArgs a{};
    a.op = FIN_TCG_UPDATE;
    a.nv = 2;
    a.nparts = nb;
    a.skip = &L.dsc->tcg_done;
    launch_fin(L, a);
}
void launch_tcg_dupdate(const Launch &L, int64_t n3, double *d, const double *z) {
    hipLaunchKernelGGL(k_tcg_dupdate, dim3(grid_for(n3, 4)), dim3(kBlock), 0, L.s, n3, d, z, L.dsc);
}
void launch_tcg_final(const Launch &L, int64_t n3, const double *p, const double *g, double *zout) {
    hipLaunchKernelGGL(k_tcg_final, dim3(grid_for(n3, 4)), dim3(kBlock), 0, L.s, n3, p, g, zout, L.dsc);
}

// ------------------------------------------------------------------ truncated CG on the magnet block
// Structured grids: every node outside the magnet block has m = 0 and a zero gradient
// (all its voxels carry the zero material), so the truncated CG above leaves it at zero
// and adds nothing to its dot products.  These kernels run the same recursion on the
// (mn+1)^3 block nodes only, magnet index t = i + n1 (j + n1 k), vectors [3 t + a] in the
// storage type T (FP32 for the mixed form, FP64 otherwise; dot products in FP64), with
// d = -z + beta d, C d and H d fused into one kernel:
//   k_tcgc_init   r = -g, z = r / D, d = -z (block copies of m0, m0.gF, D); (r.r, r.z)
//   k_tcgc_dh     d' = -z + beta d (27-neighbour gather), C d', H d'; d'.H d'
//   k_tcgc_upd    r += alpha H d', z = r / D; (r.r, r.z); alpha kept per iteration
//   k_tcgc_accum  p = sum_k alpha_k d'_k over the iterations' directions (kept in a ring
//                 of slots, so no iteration reads and writes p); max |p|
//   k_tcgc_final  zout = (max|p| == 0) ? g : p scattered to the grid (0 off the block)

__device__ __forceinline__ void block_ijk(int n1, int64_t t, int &i, int &j, int &k) {
    const uint32_t tu = (uint32_t)t, n1u = (uint32_t)n1, tj = tu / n1u, kk = tj / n1u;   // n1^3 < 2^32
    i = (int)(tu - tj * n1u);
    j = (int)(tj - kk * n1u);
    k = (int)kk;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_tcgc_init(GridDev G, const double *__restrict__ g,
                                                      const double *__restrict__ gF,
                                                      const double *__restrict__ m0,
                                                      const double *__restrict__ D, T *__restrict__ m0c,
                                                      T *__restrict__ mgc, T *__restrict__ Dc, T *__restrict__ r,
                                                      T *__restrict__ z, T *__restrict__ d,
                                                      double *__restrict__ alpha, int nalpha, double *partials) {
    const int n1 = G.mn + 1;
    const int64_t nb = (int64_t)n1 * n1 * n1;
    if (blockIdx.x == 0)
        for (int q = threadIdx.x; q < nalpha; q += kBlock) alpha[q] = 0.0;
    double s[2] = {0.0, 0.0};
    for (int64_t t = blockIdx.x * (int64_t)kBlock + threadIdx.x; t < nb; t += (int64_t)gridDim.x * kBlock) {
        int i, j, k;
        block_ijk(n1, t, i, j, k);
        const int64_t gn = gnode(G, i, j, k);
        double mg = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double ma = m0[3 * gn + a];
            mg += ma * gF[3 * gn + a];
            m0c[3 * t + a] = (T)ma;
            const T Da = (T)D[3 * gn + a];
            const T ri = (T)(-g[3 * gn + a]);
            const T zi = (T)((double)ri / (double)Da);
            Dc[3 * t + a] = Da;
            r[3 * t + a] = ri;
            z[3 * t + a] = zi;
            d[3 * t + a] = -zi;
            s[0] += (double)ri * (double)ri;
            s[1] += (double)ri * (double)zi;
        }
        mgc[t] = (T)mg;
    }
    block_partials<2>(s, 0u, partials);
}

// d' = FIRST ? d : -z + beta d, dout = d', Hd = H d'; partial d'.H d'.
// H v = m0 x ((C v) x m0) - (m0 . gF) v.  A block owns a kCTX x kCTY tile of block nodes
// and marches a z chunk: d' is formed once per node into a three-plane LDS ring with a
// one-node halo (clamped to the block, as the 27-point gather clamps), the voxels'
// coefficients into a two-plane ring; the node then reads its 27 neighbours and 8
// voxels from LDS.
template <typename T, bool FIRST>
__global__ __launch_bounds__(kBlock) void k_tcgc_dh(GridDev G, TcgTiles tl, const MatCoef *__restrict__ mats,
                                                    const T *__restrict__ z, const T *__restrict__ din,
                                                    T *__restrict__ dout, const T *__restrict__ m0c,
                                                    const T *__restrict__ mgc, T *__restrict__ Hd,
                                                    DevScalars *dsc, int it, const double *__restrict__ uparts,
                                                    int nuparts, double *partials) {
    static_assert(kCTX * kCTY == kBlock, "one node per thread and plane");
    static_assert(kFinThreads == kBlock, "the finalize order needs the finalize's block width");
    __shared__ T nd[3][kCTY + 2][kCTX + 2][3];
    __shared__ T vc[2][kCTY + 1][kCTX + 1][5];
    __shared__ double shf[kBlock / kWave];
    if (dsc->tcg_done) return;   // block-uniform
    // The previous iteration's finalize (FIN_TCG_UPDATE), formed in every block from the
    // update's partials (r'.r', r'.z') in k_fin's order: stop test and beta; block 0
    // publishes them, writing r.z into the other parity slot than the one read here.
    double beta_d = 0.0;
    if (!FIRST) {
        const double rr = fin_order_sum(uparts, nuparts, shf);
        __syncthreads();   // shf is reused by the second sum
        const double rz = fin_order_sum(uparts + kMaxPartials, nuparts, shf);
        const bool neg = dsc->tcg_den <= 0.0;
        const bool conv = sqrt(rr) <= dsc->tcg_eps;
        beta_d = (!neg && !conv) ? rz / (dsc->tcg_rzp[(it - 1) & 1] + 1e-16) : 0.0;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            if (!neg) dsc->tcg_rzp[it & 1] = rz;
            dsc->tcg_beta = beta_d;
            dsc->tcg_neg = neg;
            dsc->tcg_done = neg || conv;
            dsc->tcg_iter += 1;
            dsc->tcg_total += 1;
        }
        if (neg || conv) return;   // block-uniform
    }
    const int mn = G.mn, n1 = mn + 1;
    const int b = blockIdx.x;
    const int cz = b % tl.gz, bx = (b / tl.gz) % tl.gx, by = b / (tl.gz * tl.gx);
    const int tx = threadIdx.x % kCTX, ty = threadIdx.x / kCTX;
    const int i0 = bx * kCTX, j0 = by * kCTY, i = i0 + tx, j = j0 + ty;
    const int chunk = (n1 + tl.gz - 1) / tl.gz;
    const int k0 = cz * chunk, k1 = min(n1, k0 + chunk);
    const T beta = (T)beta_d;
    auto ring3 = [](int p) { return ((p % 3) + 3) % 3; };
    // A thread's ring entries: nodes e = t, t + kBlock < (kCTY + 2)(kCTX + 2), voxels
    // e = t, t + kBlock < (kCTY + 1)(kCTX + 1).  Plane k + 2's raw node values (d, and z
    // unless FIRST) and plane k + 1's voxel materials are loaded into registers before
    // plane k is computed; d' = -z + beta d is formed when they are written to the ring.
    constexpr int kNodeE = (kCTY + 2) * (kCTX + 2), kVoxE = (kCTY + 1) * (kCTX + 1);
    static_assert(kNodeE <= 2 * kBlock && kVoxE <= 2 * kBlock, "two ring entries per thread");
    struct NodeRegs {
        T d[2][3], z[2][3];
    };
    struct VoxRegs {
        int mat[2];
    };
    auto fetch_nodes = [&](int kp) {   // plane kp (clamped) with its halo
        NodeRegs r{};
        const int kk = min(max(kp, 0), mn);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int e = threadIdx.x + q * kBlock;
            if (e >= kNodeE) break;
            const int ex = e % (kCTX + 2), ey = e / (kCTX + 2);
            const int ii = min(max(i0 - 1 + ex, 0), mn), jj = min(max(j0 - 1 + ey, 0), mn);
            const int64_t tn = ii + (int64_t)n1 * (jj + (int64_t)n1 * kk);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                r.d[q][a] = din[3 * tn + a];
                if (!FIRST) r.z[q][a] = z[3 * tn + a];
            }
        }
        return r;
    };
    auto store_nodes = [&](const NodeRegs &r, int kp) {
        const int slot = ring3(kp);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int e = threadIdx.x + q * kBlock;
            if (e >= kNodeE) break;
            const int ex = e % (kCTX + 2), ey = e / (kCTX + 2);
#pragma unroll
            for (int a = 0; a < 3; ++a) nd[slot][ey][ex][a] = FIRST ? r.d[q][a] : -r.z[q][a] + beta * r.d[q][a];
        }
    };
    auto fetch_voxels = [&](int kz) {   // voxel plane kz (zero material outside the block)
        VoxRegs r{};
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int e = threadIdx.x + q * kBlock;
            if (e >= kVoxE) break;
            const int ex = e % (kCTX + 1), ey = e / (kCTX + 1);
            const int xi = i0 - 1 + ex, yj = j0 - 1 + ey;
            const bool inside = xi >= 0 && yj >= 0 && kz >= 0 && xi < mn && yj < mn && kz < mn;
            r.mat[q] = inside ? (int)G.vmat[xi + (int64_t)mn * (yj + (int64_t)mn * kz)] : G.n_mats;
        }
        return r;
    };
    auto store_voxels = [&](const VoxRegs &r, int kz) {
        const int slot = kz & 1;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int e = threadIdx.x + q * kBlock;
            if (e >= kVoxE) break;
            const int ex = e % (kCTX + 1), ey = e / (kCTX + 1);
            const VoxCoef<T> c = vox_coef<T>(mats[r.mat[q]]);
            vc[slot][ey][ex][0] = c.cex;
            vc[slot][ey][ex][1] = c.cK;
            vc[slot][ey][ex][2] = c.u0;
            vc[slot][ey][ex][3] = c.u1;
            vc[slot][ey][ex][4] = c.u2;
        }
    };
    double den = 0.0;
    if (k0 < k1) {   // block-uniform
        store_nodes(fetch_nodes(k0 - 1), k0 - 1);
        store_nodes(fetch_nodes(k0), k0);
        store_voxels(fetch_voxels(k0 - 1), k0 - 1);
        NodeRegs rn = fetch_nodes(k0 + 1);
        VoxRegs rv = fetch_voxels(k0);
        for (int k = k0; k < k1; ++k) {
            store_nodes(rn, k + 1);
            store_voxels(rv, k);
            if (k + 1 < k1) {
                rn = fetch_nodes(k + 2);
                rv = fetch_voxels(k + 1);
            }
            __syncthreads();
            if (i < n1 && j < n1) {
                T nv[27][3];
#pragma unroll
                for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
                    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                        for (int dx = -1; dx <= 1; ++dx) {
                            const int sl = ring3(k + dz), s = (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1);
#pragma unroll
                            for (int a = 0; a < 3; ++a) nv[s][a] = nd[sl][ty + 1 + dy][tx + 1 + dx][a];
                        }
                auto vox = [&](int w) {
                    const int vx = w & 1, vy = (w >> 1) & 1, sl = (k - 1 + ((w >> 2) & 1)) & 1;
                    const T *q = vc[sl][ty + vy][tx + vx];
                    return VoxCoef<T>{q[0], q[1], (T)0, q[2], q[3], q[4]};
                };
                T cv[3], unused = 0;
                grid_c_core<T, false>(G, nv, vox, cv, unused);
                const int64_t t = i + (int64_t)n1 * (j + (int64_t)n1 * k);
                double mm[3], cd[3], dd[3], c[3], pr[3];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    mm[a] = (double)m0c[3 * t + a];
                    cd[a] = (double)cv[a];
                    dd[a] = (double)nv[13][a];
                    dout[3 * t + a] = nv[13][a];
                }
                const double mg = (double)mgc[t];
                cross3(cd, mm, c);
                cross3(mm, c, pr);
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const T h = (T)(pr[a] - mg * dd[a]);
                    Hd[3 * t + a] = h;
                    den += dd[a] * (double)h;
                }
            }
            __syncthreads();
        }
    }
    double v[1] = {den};
    block_partials<1>(v, 0u, partials);
}

// den = d'.H d' from k_tcgc_dh's nden partials, summed in every block in k_fin's order
// (so alpha is what a finalize launch would give); block 0 publishes it and the
// iteration's step length (0 on negative curvature; an iteration skipped after the stop
// keeps the 0 the init wrote).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_tcgc_upd(int64_t n3, T *__restrict__ r, T *__restrict__ z,
                                                     const T *__restrict__ Hd, const T *__restrict__ Dc,
                                                     DevScalars *dsc, const double *__restrict__ dparts,
                                                     int nden, double *__restrict__ alpha_out, int it,
                                                     double *partials) {
    static_assert(kFinThreads == kBlock, "the finalize order needs the finalize's block width");
    __shared__ double sh[kBlock / kWave];
    const int done = dsc->tcg_done;   // block-uniform; tested after the barrier
    const double den = fin_order_sum(dparts, nden, sh);
    if (done) return;
    const bool neg = den <= 0.0;
    const double alpha = neg ? 0.0 : dsc->tcg_rzp[it & 1] / den;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        dsc->tcg_den = den;
        *alpha_out = alpha;
    }
    double s[2] = {0.0, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n3; i += (int64_t)gridDim.x * kBlock) {
        const T rn = (T)((double)r[i] + alpha * (double)Hd[i]);
        const T zn = (T)((double)rn / (double)Dc[i]);
        if (!neg) {
            r[i] = rn;
            z[i] = zn;
        }
        s[0] += (double)rn * (double)rn;
        s[1] += (double)rn * (double)zn;
    }
    block_partials<2>(s, 0u, partials);
}

// p = sum_k alpha_k d'_k in iteration order (the sum an update p += alpha_k d'_k per
// iteration forms, rounded to T once), skipping the iterations without an update
// (alpha_k = 0); max |p|.  Only the iterations that ran (tcg_iter, counted by the
// finalize) wrote a direction; their loads are issued four slots at a time.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_tcgc_accum(int64_t n3, const T *__restrict__ dring,
                                                       const double *__restrict__ alpha, int nit,
                                                       const DevScalars *dsc, T *__restrict__ p, double *partials) {
    const int nrun = min(dsc->tcg_iter, nit);   // uniform
    double s[1] = {0.0};
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n3; i += (int64_t)gridDim.x * kBlock) {
        double acc = 0.0;
        int k = 0;
        for (; k + 4 <= nrun; k += 4) {
            double dv[4], av[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                av[q] = alpha[k + q];
                dv[q] = (double)dring[(int64_t)(k + 1 + q) * n3 + i];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) acc = av[q] != 0.0 ? acc + av[q] * dv[q] : acc;
        }
        for (; k < nrun; ++k) {
            const double a = alpha[k], dv = (double)dring[(int64_t)(k + 1) * n3 + i];
            acc = a != 0.0 ? acc + a * dv : acc;
        }
        const T pv = (T)acc;
        p[i] = pv;
        s[0] = fmax(s[0], fabs((double)pv));
    }
    block_partials<1>(s, 1u, partials);
}

// The same on the block's nodes only (zout keeps its entries off the block)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_tcgc_final_block(GridDev G, const T *__restrict__ p,
                                                             const double *__restrict__ g,
                                                             double *__restrict__ zout, const DevScalars *dsc) {
    const bool zero = dsc->tcg_pmax == 0.0;
    const uint32_t n1 = (uint32_t)G.mn + 1u, nb = n1 * n1 * n1;
    for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < nb; t += gridDim.x * kBlock) {
        const int64_t gn = block_gnode(G, t);
#pragma unroll
        for (int a = 0; a < 3; ++a) zout[3 * gn + a] = zero ? g[3 * gn + a] : (double)p[3 * (int64_t)t + a];
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_tcgc_final(GridDev G, int64_t n, const T *__restrict__ p,
                                                       const double *__restrict__ g, double *__restrict__ zout,
                                                       const DevScalars *dsc) {
    const bool zero = dsc->tcg_pmax == 0.0;
    const int n1 = G.mn + 1;
    for (int64_t gn = blockIdx.x * (int64_t)kBlock + threadIdx.x; gn < n; gn += (int64_t)gridDim.x * kBlock) {
        const uint32_t gu = (uint32_t)gn, r = gu / (uint32_t)G.nx, kz = r / (uint32_t)G.ny;   // N < 2^32
        const int i = (int)(gu - r * (uint32_t)G.nx) - G.m0x, j = (int)(r - kz * (uint32_t)G.ny) - G.m0y,
                  k = (int)kz - G.m0z;
        const bool in = i >= 0 && j >= 0 && k >= 0 && i < n1 && j < n1 && k < n1;
        const int64_t t = in ? i + (int64_t)n1 * (j + (int64_t)n1 * k) : 0;
#pragma unroll
        for (int a = 0; a < 3; ++a)
            zout[3 * gn + a] = zero ? g[3 * gn + a] : (in ? (double)p[3 * t + a] : 0.0);
    }
}

template <typename T>
void tcgc_run(const Launch &L, const GridDev &G, int64_t n, const TcgBlock<T> &B, const double *m0,
              const double *gF, const double *g, const double *D, int max_iter, double *zout, bool write_off) {
    const int64_t n1 = G.mn + 1, nb = n1 * n1 * n1, n3 = 3 * nb;
    // the update's partials are summed by every block of the next stencil launch: at most
    // 1024 of them
    const int gi = grid_for(nb, 2), gu = std::min(grid_for(n3, 4), 1024);
    const TcgTiles tl = block_tiles(G);
    const int gd = tl.gx * tl.gy * tl.gz;
    hipLaunchKernelGGL(k_tcgc_init<T>, dim3(gi), dim3(kBlock), 0, L.s, G, g, gF, m0, D, B.m0, B.mg, B.D, B.r, B.z,
                       B.dring, B.alpha, max_iter, L.red.partials);
    FinArgs ai{};
    ai.op = FIN_TCG_INIT;
    ai.nv = 2;
    ai.nparts = gi;
    launch_fin(L, ai);
    for (int it = 0; it < max_iter; ++it) {
        const T *din = B.dring + (int64_t)it * n3;
        T *dout = B.dring + (int64_t)(it + 1) * n3;
        // d'.H d' partials in slot 3, read by the update kernel while it writes slots 0, 1
        double *const dparts = L.red.partials + 3 * kMaxPartials;
        // from the second iteration on, the stencil launch also finalizes the previous
        // update (the update's partials in slots 0, 1)
        if (it == 0)
            hipLaunchKernelGGL((k_tcgc_dh<T, true>), dim3(gd), dim3(kBlock), 0, L.s, G, tl, L.mats, B.z, din, dout,
                               B.m0, B.mg, B.Hd, L.dsc, it, L.red.partials, gu, dparts);
        else
            hipLaunchKernelGGL((k_tcgc_dh<T, false>), dim3(gd), dim3(kBlock), 0, L.s, G, tl, L.mats, B.z, din, dout,
                               B.m0, B.mg, B.Hd, L.dsc, it, L.red.partials, gu, dparts);
        hipLaunchKernelGGL(k_tcgc_upd<T>, dim3(gu), dim3(kBlock), 0, L.s, n3, B.r, B.z, B.Hd, B.D, L.dsc, dparts, gd,
                           B.alpha + it, it, L.red.partials);
    }
    if (max_iter > 0) {   // the last update's finalize (its r.z is not needed)
        FinArgs a{};
        a.op = FIN_TCG_UPDATE;
        a.nv = 2;
        a.nparts = gu;
        a.skip = &L.dsc->tcg_done;
        launch_fin(L, a);
    }
    const int ga = grid_for(n3, 4);
    hipLaunchKernelGGL(k_tcgc_accum<T>, dim3(ga), dim3(kBlock), 0, L.s, n3, B.dring, B.alpha, max_iter, L.dsc,
                       B.p, L.red.partials);
    launch_fin(L, FinArgs::store(&L.dsc->tcg_pmax, 1, 1u, ga, nullptr));
    if (write_off)
        hipLaunchKernelGGL(k_tcgc_final<T>, dim3(grid_for(n, 2)), dim3(kBlock), 0, L.s, G, n, B.p, g, zout, L.dsc);
    else
        hipLaunchKernelGGL(k_tcgc_final_block<T>, dim3(grid_for(nb, 2)), dim3(kBlock), 0, L.s, G, B.p, g, zout,
                           L.dsc);
}
template void tcgc_run<float>(const Launch &, const GridDev &, int64_t, const TcgBlock<float> &, const double *,
                              const double *, const double *, const double *, int, double *, bool);
template void tcgc_run<double>(const Launch &, const GridDev &, int64_t, const TcgBlock<double> &, const double *,
                               const double *, const double *, const double *, int, double *, bool);

// ------------------------------------------------------------------ vector kernels

// out_i = normalize(x_i + alpha d_i), zero-length vectors divided by 1 (jax_tools.py:4-18)
__global__ __launch_bounds__(kBlock) void k_update_m(int64_t n, const double *__restrict__ x,
                                                     double alpha, const double *__restrict__ d,
                                                     double *__restrict__ out) {
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        double v[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) v[a] = x[3 * i + a] + alpha * d[3 * i + a];
        double nrm = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (nrm == 0.0) nrm = 1.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) out[3 * i + a] = v[a] / nrm;
    }
}
// The same on the magnet block's nodes only (out keeps its entries off the block)
__global__ __launch_bounds__(kBlock) void k_update_m_block(GridDev G, const double *__restrict__ x,
                                                           double alpha, const double *__restrict__ d,
                                                           double *__restrict__ out) {
    const uint32_t n1 = (uint32_t)G.mn + 1u, nb = n1 * n1 * n1;
    for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < nb; t += gridDim.x * kBlock) {
        const int64_t i = block_gnode(G, t);
        double v[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) v[a] = x[3 * i + a] + alpha * d[3 * i + a];
        double nrm = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (nrm == 0.0) nrm = 1.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) out[3 * i + a] = v[a] / nrm;
    }
}
void launch_update_m(const Launch &L, int64_t n, const double *x, double alpha, const double *d,
                     double *out, const GridDev *G) {
    if (G) {
        const int64_t n1 = G->mn + 1;
        hipLaunchKernelGGL(k_update_m_block, dim3(grid_for(n1 * n1 * n1, 2)), dim3(kBlock), 0, L.s, *G, x, alpha,
                           d, out);
        return;
    }
    hipLaunchKernelGGL(k_update_m, dim3(grid_for(n, 2)), dim3(kBlock), 0, L.s, n, x, alpha, d, out);
}

// out may alias y (d = -z + beta d): each element is read before it is written.
__global__ __launch_bounds__(kBlock) void k_axpby(int64_t n, double a, const double *__restrict__ x,
                                                  double b, const double *y, double *out) {
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        out[i] = (y ? a * x[i] + b * y[i] : a * x[i]);
}
// The same over interleaved 3-vectors on the magnet block's nodes only
__global__ __launch_bounds__(kBlock) void k_axpby3_block(GridDev G, double a, const double *__restrict__ x,
                                                         double b, const double *y, double *out) {
    const uint32_t n1 = (uint32_t)G.mn + 1u, nb = n1 * n1 * n1;
    for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < nb; t += gridDim.x * kBlock) {
        const int64_t i = 3 * block_gnode(G, t);
#pragma unroll
        for (int c = 0; c < 3; ++c) out[i + c] = (y ? a * x[i + c] + b * y[i + c] : a * x[i + c]);
    }
}
void launch_axpby(const Launch &L, int64_t n, double a, const double *x, double b, const double *y,
                  double *out, const GridDev *G3) {
    if (G3) {
        const int64_t n1 = G3->mn + 1;
        hipLaunchKernelGGL(k_axpby3_block, dim3(grid_for(n1 * n1 * n1, 2)), dim3(kBlock), 0, L.s, *G3, a, x, b, y,
                           out);
        return;
    }
    hipLaunchKernelGGL(k_axpby, dim3(grid_for(n, 4)), dim3(kBlock), 0, L.s, n, a, x, b, y, out);
}

// ------------------------------------------------------------------ generic reductions

// With G.nx > 0 (structured grid; the vectors are zero off the magnet block) only the
// block's entries are visited: element e of the block range maps to the grid element
// comp * node(e / comp) + e % comp (comp = 3 for interleaved 3-vectors, 1 for nodes).
template <int OP>
__global__ __launch_bounds__(kBlock) void k_reduce(int64_t n, GridDev G, int comp, const double *__restrict__ a,
                                                   const double *__restrict__ b,
                                                   const double *__restrict__ c,
                                                   const double *__restrict__ d, double h0, double h1,
                                                   double h2, double *partials) {
    constexpr int NV = (OP == RED_STOP3 || OP == RED_DG3 || OP == RED_H3) ? 3
                       : (OP == RED_GD2 || OP == RED_YZD)                 ? 2
                                                                          : 1;
    constexpr unsigned MASK = OP == RED_NORMINF ? 1u : OP == RED_STOP3 ? 7u : OP == RED_H3 ? 4u : 0u;
    double s[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) s[k] = 0.0;
    const int n1 = G.mn + 1;
    const int64_t ne = G.nx > 0 ? (int64_t)n1 * n1 * n1 * comp : n;
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < ne; e += (int64_t)gridDim.x * kBlock) {
        int64_t i = e;
        if (G.nx > 0) {   // 32-bit index arithmetic (the block holds fewer than 2^32 entries)
            const uint32_t eu = (uint32_t)e, n1u = (uint32_t)n1;
            const uint32_t t = comp == 3 ? eu / 3u : eu;
            const uint32_t r = eu - t * (uint32_t)comp;
            const uint32_t tj = t / n1u, k = tj / n1u;
            i = comp * gnode(G, (int)(t - tj * n1u), (int)(tj - k * n1u), (int)k) + r;
        }
        if constexpr (OP == RED_DOT) {
            s[0] += a[i] * b[i];
        } else if constexpr (OP == RED_NORMINF) {
            s[0] = fmax(s[0], fabs(a[i]));
        } else if constexpr (OP == RED_STOP3) {
            s[0] = fmax(s[0], fabs(a[i] - b[i]));
            s[1] = fmax(s[1], fabs(b[i]));
            s[2] = fmax(s[2], fabs(c[i]));
        } else if constexpr (OP == RED_DG3) {
            s[0] += a[i] * b[i];
            s[1] += a[i] * a[i];
            s[2] += b[i] * b[i];
        } else if constexpr (OP == RED_H3) {
            s[0] += a[i] * a[i];
            s[1] += b[i] * b[i];
            s[2] = fmax(s[2], fabs(b[i]));
        } else if constexpr (OP == RED_GD2) {
            s[0] += a[i] * c[i];
            s[1] += b[i] * c[i];
        } else if constexpr (OP == RED_YZD) {
            const double y = b[i] - a[i];
            s[0] += y * c[i];
            s[1] += y * d[i];
        } else if constexpr (OP == RED_MH) {
            s[0] += (h0 * a[3 * i] + h1 * a[3 * i + 1] + h2 * a[3 * i + 2]) * b[i];
        } else {
            s[0] += a[i] * a[i];
        }
    }
    block_partials<NV>(s, MASK, partials);
}

// max |m| over the nodes off the magnet block of a structured grid (m interleaved [3 N])
__global__ __launch_bounds__(kBlock) void k_outside_max(int64_t n, GridDev G, const double *__restrict__ m,
                                                        double *partials) {
    double s[1] = {0.0};
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const uint32_t iu = (uint32_t)i, r = iu / (uint32_t)G.nx, k = r / (uint32_t)G.ny;
        const int gi = (int)(iu - r * (uint32_t)G.nx) - G.m0x, gj = (int)(r - k * (uint32_t)G.ny) - G.m0y,
                  gk = (int)k - G.m0z;
        if (gi < 0 || gj < 0 || gk < 0 || gi > G.mn || gj > G.mn || gk > G.mn)
#pragma unroll
            for (int a = 0; a < 3; ++a) s[0] = fmax(s[0], fabs(m[3 * i + a]));
    }
    block_partials<1>(s, 1u, partials);
}
void launch_outside_max(const Launch &L, int64_t n, const GridDev &G, const double *m, double *out) {
    const int nb = grid_for(n, 4);
    hipLaunchKernelGGL(k_outside_max, dim3(nb), dim3(kBlock), 0, L.s, n, G, m, L.red.partials);
    launch_fin(L, FinArgs::store(out, 1, 1u, nb, nullptr));
}

void launch_reduce(const Launch &L, RedOp op, int64_t n, const double *a, const double *b,
                   const double *c, const double *d, double h0, double h1, double h2, double *out,
                   const GridDev *G, int comp) {
    const GridDev gd = G ? *G : GridDev{};
    const int64_t n1 = gd.mn + 1;
    const int nb = grid_for(G ? n1 * n1 * n1 * comp : n, 8);
    const dim3 g(nb), blk(kBlock);
    int nv = 1;
    unsigned mask = 0;
    switch (op) {
#define MM_RED_CASE(OPC, NV, MASK)                                                                        \
    case OPC:                                                                                             \
        hipLaunchKernelGGL(k_reduce<OPC>, g, blk, 0, L.s, n, gd, comp, a, b, c, d, h0, h1, h2, L.red.partials); \
        nv = NV;                                                                                          \
        mask = MASK;                                                                                      \
        break;
        MM_RED_CASE(RED_DOT, 1, 0u)
        MM_RED_CASE(RED_NORMINF, 1, 1u)
        MM_RED_CASE(RED_STOP3, 3, 7u)
        MM_RED_CASE(RED_DG3, 3, 0u)
        MM_RED_CASE(RED_H3, 3, 4u)
        MM_RED_CASE(RED_GD2, 2, 0u)
        MM_RED_CASE(RED_YZD, 2, 0u)
        MM_RED_CASE(RED_MH, 1, 0u)
        MM_RED_CASE(RED_SS, 1, 0u)
#undef MM_RED_CASE
    }
    launch_fin(L, FinArgs::store(out, nv, mask, nb, nullptr));
}

// Totals of NV values' partials into out[0 .. NV), each summed in k_fin's order (per
// thread over j = t, t + 256, ..., then the wave, then the waves in order).
template <int NV>
__global__ __launch_bounds__(kFinThreads) void k_fin_store(const double *__restrict__ partials, int nparts,
                                                           unsigned maxmask, double *__restrict__ out) {
    __shared__ double sh[NV][kFinThreads / kWave];
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    double acc[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) acc[i] = 0.0;
    for (int j = threadIdx.x; j < nparts; j += kFinThreads) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const double p = partials[i * kMaxPartials + j];
            acc[i] = ((maxmask >> i) & 1u) ? fmax(acc[i], p) : acc[i] + p;
        }
    }
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        acc[i] = ((maxmask >> i) & 1u) ? wave_max(acc[i]) : wave_sum(acc[i]);
        if (lane == 0) sh[i][w] = acc[i];
    }
    __syncthreads();
    if (threadIdx.x >= NV) return;
    const int i = threadIdx.x;
    const bool mx = (maxmask >> i) & 1u;
    double t = sh[i][0];
    for (int k = 1; k < kFinThreads / kWave; ++k) t = mx ? fmax(t, sh[i][k]) : t + sh[i][k];
    out[i] = t;
}

// One pass over the NCG state between two iterations.  With UPD, the HS+ direction
// d = -z + beta d (minimizers.py:175-181), beta = max(yzd[0] / yzd[1], 0) from the totals
// of (g1 - g).z and (g1 - g).d; then the totals every test up to the next line search may
// need (NcgScan in common.h), for d as formed and for the two restart directions -z and
// -g1.  The element loop, block count and block reduction are k_reduce's over the same
// vectors, so each total is the value the separate reduction would give.
template <bool UPD>
__global__ __launch_bounds__(kBlock) void k_ncg_scan(int64_t n, GridDev G, const double *__restrict__ xo,
                                                     const double *__restrict__ xn, const double *__restrict__ go,
                                                     const double *__restrict__ gn, const double *__restrict__ z,
                                                     double *__restrict__ d, const double *__restrict__ yzd,
                                                     double *partials) {
    double s[NCG_NV];
#pragma unroll
    for (int k = 0; k < NCG_NV; ++k) s[k] = 0.0;
    double beta = 0.0;
    if (UPD) {
        const double q = yzd[0] / yzd[1];
        beta = q < 0.0 ? 0.0 : q;
    }
    const int n1 = G.mn + 1;
    const int64_t ne = G.nx > 0 ? (int64_t)n1 * n1 * n1 * 3 : n;
    for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < ne; e += (int64_t)gridDim.x * kBlock) {
        int64_t i = e;
        if (G.nx > 0) {
            const uint32_t eu = (uint32_t)e, n1u = (uint32_t)n1;
            const uint32_t t = eu / 3u, r = eu - t * 3u;
            const uint32_t tj = t / n1u, k = tj / n1u;
            i = 3 * gnode(G, (int)(t - tj * n1u), (int)(tj - k * n1u), (int)k) + r;
        }
        const double zv = z[i], gov = go[i], gnv = gn[i], xnv = xn[i];
        double dv = d[i];
        if (UPD) {
            dv = fma(beta, dv, -zv);
            d[i] = dv;
        }
        s[NCG_DG_OLD] += dv * gov;
        s[NCG_DD] += dv * dv;
        s[NCG_GG_OLD] += gov * gov;
        s[NCG_DX_MAX] = fmax(s[NCG_DX_MAX], fabs(xo[i] - xnv));
        s[NCG_X_MAX] = fmax(s[NCG_X_MAX], fabs(xnv));
        s[NCG_G_MAX] = fmax(s[NCG_G_MAX], fabs(gnv));
        s[NCG_DG] += dv * gnv;
        s[NCG_GG] += gnv * gnv;
        s[NCG_ZG] += zv * gnv;
        s[NCG_ZZ] += zv * zv;
        s[NCG_XX] += xnv * xnv;
        s[NCG_D_MAX] = fmax(s[NCG_D_MAX], fabs(dv));
        s[NCG_Z_MAX] = fmax(s[NCG_Z_MAX], fabs(zv));
    }
    block_partials<NCG_NV>(s, kNcgMaxMask, partials);
}

int launch_ncg_scan(const Launch &L, int64_t n3, const GridDev *G, const double *xo, const double *xn,
                    const double *go, const double *gn, const double *z, double *d, const double *yzd,
                    double *partials, double *out) {
    const GridDev gd = G ? *G : GridDev{};
    const int64_t n1 = gd.mn + 1;
    const int nb = grid_for(G ? n1 * n1 * n1 * 3 : n3, 8);
    if (yzd)
        hipLaunchKernelGGL(k_ncg_scan<true>, dim3(nb), dim3(kBlock), 0, L.s, n3, gd, xo, xn, go, gn, z, d, yzd,
                           partials);
    else
        hipLaunchKernelGGL(k_ncg_scan<false>, dim3(nb), dim3(kBlock), 0, L.s, n3, gd, xo, xn, go, gn, z, d, yzd,
                           partials);
    hipLaunchKernelGGL(k_fin_store<NCG_NV>, dim3(1), dim3(kFinThreads), 0, L.s, partials, nb, kNcgMaxMask, out);
    return nb;
}

__global__ void k_fill_zero(int64_t n, double *__restrict__ x, const int *skip) {
    if (skip && *skip) return;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        x[i] = 0.0;
}
void launch_fill_zero(const Launch &L, int64_t n, double *x, const int *skip_flag) {
    hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(n, 8)), dim3(kBlock), 0, L.s, n, x, skip_flag);
}

__global__ void k_fill_zero32(int64_t n, float *__restrict__ x, const int *skip) {
    if (skip && *skip) return;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        x[i] = 0.0f;
}
void launch_fill_zero32(const Launch &L, int64_t n, float *x, const int *skip_flag) {
    hipLaunchKernelGGL(k_fill_zero32, dim3(grid_for(n, 8)), dim3(kBlock), 0, L.s, n, x, skip_flag);
}

}  // namespace mm
