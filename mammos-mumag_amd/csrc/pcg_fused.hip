// Fused Jacobi-PCG iteration: two kernels per iteration, no finisher launches.
//
//   KA(i): every block sums the partials of the previous KB (r.z, r.r) in a fixed order
//          (identical in every block), derives beta_i = rz / gamma_{i-1}, gamma_i = rz
//          and the stop flag (r.r > bound and i < maxiter, jax CG), then forms
//          p_i = z + beta_i p_{i-1} on the fly (own node and neighbours), writes p_i,
//          q = K p_i and partials of p_i . q.  Block 0 publishes the scalars.
//   KB(i): every block sums KA's partials (p.q), alpha = gamma_i / pq;
//          x += alpha p_i, r -= alpha q, z = r / diag(K); partials (r.z, r.r).
//
// gamma is double-buffered by iteration parity (KA(i) reads gamma[(i-1)&1] while its
// block 0 writes gamma[i&1]).  Modes of KA: FIRST (after the initial residual of the
// FP64 path: p = z, the bound is fixed from b.b), NORMAL, RESTART (after the initial
// true residual of the mixed path: p = z, gamma and r.r from that residual) and KEEP
// (after a reliable update of the mixed path: the search direction is kept, i.e.
// p = z + beta p with gamma, r.r and z from the replaced residual).
// T = double (FP64 path) or float (mixed path: FP32 vectors, FP64 sums).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.h"

namespace mm {

namespace {

constexpr int kFB = 1024;  // threads per block of the fused kernels (16 wavefronts)

__device__ __forceinline__ double fwave_sum(double v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Sum of n partials (fixed order: per-thread strided sums, then a fixed tree).
__device__ double block_sum_parts(const double *__restrict__ parts, int n, double *sh) {
    double s = 0.0;
    for (int j = threadIdx.x; j < n; j += kFB) s += parts[j];
    s = fwave_sum(s);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) sh[w] = s;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < kFB / kWave; ++k) t += sh[k];
    return t;
}

// Block reduction of NV values into out[i * kMaxPartials + block].
template <int NV>
__device__ void fblock_partials(double (&v)[NV], double *sh, double *out) {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double s = fwave_sum(v[i]);
        __syncthreads();
        if (lane == 0) sh[w] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int k = 0; k < kFB / kWave; ++k) t += sh[k];
            out[i * kMaxPartials + bid] = t;
        }
    }
}

// Block-uniform read of the stop flag: block 0 of the same KA kernel may write it while
// other blocks start, so one thread reads it and broadcasts it through LDS.
__device__ bool block_done(const DevScalars *dsc, int *flag) {
    if (threadIdx.x == 0) *flag = dsc->pcg_done;
    __syncthreads();
    return *flag != 0;
}

// Scalars of iteration i from the partials of the previous KB (or the init / replaced
// residual).  Returns beta; sets *stop.  Block 0 thread 0 publishes.
__device__ double ka_prologue(const PcgIterArgs &a, DevScalars *dsc, double *sh, bool *stop) {
    double beta = 0.0, gamma, rr;
    if (a.mode == PCG_KEEP || a.mode == PCG_RESTART) {
        gamma = dsc->ru_gamma;
        rr = dsc->ir_rr;
        if (a.mode == PCG_KEEP) beta = gamma / dsc->pcg_gam[(a.iter + 1) & 1];
    } else {
        gamma = block_sum_parts(a.parts_in, a.nparts_in, sh);
        rr = block_sum_parts(a.parts_in + kMaxPartials, a.nparts_in, sh);
        if (a.mode == PCG_NORMAL) beta = gamma / dsc->pcg_gam[(a.iter + 1) & 1];
    }
    double bound = dsc->pcg_bound;
    if (a.mode == PCG_FIRST) {
        const double bb = block_sum_parts(a.parts_in + 2 * kMaxPartials, a.nparts_in, sh);
        bound = fmax(a.tol * a.tol * bb, a.atol * a.atol);
    }
    const bool done = !((rr > bound) && (a.iter < a.maxiter));
    if (threadIdx.x == 0 && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0) {
        dsc->pcg_gam[a.iter & 1] = gamma;
        dsc->pcg_beta = beta;
        dsc->pcg_rr = rr;
        dsc->pcg_iter = a.iter;
        if (a.mode == PCG_FIRST) {
            dsc->pcg_bound = bound;
            dsc->pcg_bs = fmax(rr, 0.0);
        }
        dsc->pcg_done = done;
    }
    *stop = done;
    return beta;
}

template <typename T>
__device__ __forceinline__ T edge_sum_f(const T *__restrict__ h1p, const T *__restrict__ h2p, int a, int b) {
    const T l1 = h1p[a], u1 = h1p[a + 1], l2 = h2p[b], u2 = h2p[b + 1];
    return (T)2 * l1 * l2 + l1 * u2 + u1 * l2 + (T)2 * u1 * u2;
}

// KA on the structured grid: 64 x 16 tiles of an (x, y) plane marching through a range
// of z planes (see k_grid_poisson in kernels.hip for the stencil coefficients).
template <typename T>
__global__ __launch_bounds__(kFB) void k_ka_grid(PoissonGrid<T> P, PcgIterArgs a,
                                                 const T *__restrict__ z, const T *__restrict__ pold,
                                                 T *__restrict__ pnew, T *__restrict__ q,
                                                 DevScalars *dsc) {
    __shared__ double sh[kFB / kWave];
    __shared__ int flag;
    if (block_done(dsc, &flag)) return;
    bool stop;
    const double bd = ka_prologue(a, dsc, sh, &stop);
    if (stop) return;
    const T beta = (T)bd;
    // p = z (FIRST, RESTART), z + beta p (NORMAL, KEEP)
    const int pm = (a.mode == PCG_FIRST || a.mode == PCG_RESTART) ? 0 : 2;
    const int nx = P.nx, ny = P.ny, nz = P.nz;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 16 + (threadIdx.x >> 6);
    const int chunk = (nz + gridDim.z - 1) / gridDim.z;
    const int k0 = blockIdx.z * chunk, k1 = min(nz, k0 + chunk);
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    auto pv = [&](int64_t idx) -> T { return pm == 0 ? z[idx] : pm == 1 ? pold[idx] : z[idx] + beta * pold[idx]; };
    double dot = 0.0;
    // empty trailing chunks (k0 >= nz) load nothing
    if (i < nx && j < ny && k0 < k1) {
        const bool edge_ij = i == 0 || j == 0 || i == nx - 1 || j == ny - 1;
        const T sz_ij = edge_sum_f(P.hxp, P.hyp, i, j);
        const T mxm = (i - 1 > 0) ? (T)1 : (T)0, mxp = (i + 1 < nx - 1) ? (T)1 : (T)0;
        const T mym = (j - 1 > 0) ? (T)1 : (T)0, myp = (j + 1 < ny - 1) ? (T)1 : (T)0;
        const T ixm = P.ixp[i], ixp = P.ixp[i + 1], iym = P.iyp[j], iyp = P.iyp[j + 1];
        const int64_t base = i + sy * j;
        T xm = k0 > 0 ? pv(base + sz * (k0 - 1)) : (T)0;
        T xc = pv(base + sz * k0);
        for (int k = k0; k < k1; ++k) {
            const int64_t idx = base + sz * k;
            const T xp = k + 1 < nz ? pv(idx + sz) : (T)0;
            T out;
            if (edge_ij || k == 0 || k == nz - 1) {
                out = xc;
            } else {
                const T xw = pv(idx - 1), xe = pv(idx + 1), xs = pv(idx - sy), xn = pv(idx + sy);
                const T sx = edge_sum_f(P.hyp, P.hzp, j, k);
                const T syy = edge_sum_f(P.hzp, P.hxp, k, i);
                const T cxm = -sx * ixm, cxp = -sx * ixp;
                const T cym = -syy * iym, cyp = -syy * iyp;
                const T czm = -sz_ij * P.izp[k], czp = -sz_ij * P.izp[k + 1];
                const T diag = -(cxm + cxp + cym + cyp + czm + czp);
                const T mzm = (k - 1 > 0) ? (T)1 : (T)0, mzp = (k + 1 < nz - 1) ? (T)1 : (T)0;
                out = diag * xc + mxm * cxm * xw + mxp * cxp * xe + mym * cym * xs + myp * cyp * xn +
                      mzm * czm * xm + mzp * czp * xp;
            }
            pnew[idx] = xc;
            q[idx] = out;
            dot += (double)out * (double)xc;
            xm = xc;
            xc = xp;
        }
    }
    double v[1] = {dot};
    fblock_partials<1>(v, sh, a.parts_out);
}

// KA on a SELL-64 matrix (general meshes): 16 slices per block.
template <typename T>
__global__ __launch_bounds__(kFB) void k_ka_sell(SellDev<T> A, PcgIterArgs a, const T *__restrict__ z,
                                                 const T *__restrict__ pold, T *__restrict__ pnew,
                                                 T *__restrict__ q, DevScalars *dsc) {
    __shared__ double sh[kFB / kWave];
    __shared__ int flag;
    if (block_done(dsc, &flag)) return;
    bool stop;
    const double bd = ka_prologue(a, dsc, sh, &stop);
    if (stop) return;
    const T beta = (T)bd;
    const int pm = (a.mode == PCG_FIRST || a.mode == PCG_RESTART) ? 0 : 2;
    auto pv = [&](int c) -> T { return pm == 0 ? z[c] : pm == 1 ? pold[c] : z[c] + beta * pold[c]; };
    const int lane = threadIdx.x & (kWave - 1);
    double dot = 0.0;
    for (int slice = blockIdx.x * (kFB / kWave) + threadIdx.x / kWave; slice < A.n_slices;
         slice += gridDim.x * (kFB / kWave)) {
        const int beg = A.slice_ptr[slice], end = A.slice_ptr[slice + 1];
        const int row = slice * kWave + lane;
        T acc = 0;
        for (int k = beg + lane; k < end; k += kWave) {
            const int c = A.col[k];
            acc += A.val[k] * pv(c);
        }
        if (row < A.n_rows) {
            const T pr = pv(row);
            pnew[row] = pr;
            q[row] = acc;
            dot += (double)acc * (double)pr;
        }
    }
    double v[1] = {dot};
    fblock_partials<1>(v, sh, a.parts_out);
}

// KB: alpha = gamma_i / (p.q); x += alpha p; r -= alpha q; z = r * invD; (r.z, r.r).
template <typename T>
__global__ __launch_bounds__(kFB) void k_kb(int64_t n, PcgIterArgs a, T *__restrict__ x,
                                            T *__restrict__ r, T *__restrict__ z,
                                            const T *__restrict__ p, const T *__restrict__ q,
                                            const T *__restrict__ invD, DevScalars *dsc) {
    __shared__ double sh[kFB / kWave];
    if (dsc->pcg_done) return;
    const double pq = block_sum_parts(a.parts_in, a.nparts_in, sh);
    const T alpha = (T)(dsc->pcg_gam[a.iter & 1] / pq);
    double s[2] = {0.0, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)kFB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kFB) {
        x[i] = x[i] + alpha * p[i];
        const T ri = r[i] - alpha * q[i];
        const T zi = ri * invD[i];
        r[i] = ri;
        z[i] = zi;
        s[0] += (double)ri * (double)zi;
        s[1] += (double)ri * (double)ri;
    }
    fblock_partials<2>(s, sh, a.parts_out);
}

// Init: r = b - A x0, z = r * invD; partials (r.z, r.r, b.b) for KA(0) in FIRST mode.
// Clears the stop flag left by the previous solve (KA(0) reads it on entry).
template <typename T>
__global__ __launch_bounds__(kFB) void k_pcg0(int64_t n, const T *__restrict__ b, const T *__restrict__ Ax0,
                                              T *__restrict__ r, T *__restrict__ z, const T *__restrict__ invD,
                                              double *parts_out, DevScalars *dsc) {
    __shared__ double sh[kFB / kWave];
    if (blockIdx.x == 0 && threadIdx.x == 0) dsc->pcg_done = 0;
    double s[3] = {0.0, 0.0, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)kFB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kFB) {
        const T bi = b[i];
        const T ri = bi - Ax0[i];
        const T zi = ri * invD[i];
        r[i] = ri;
        z[i] = zi;
        s[0] += (double)ri * (double)zi;
        s[1] += (double)ri * (double)ri;
        s[2] += (double)bi * (double)bi;
    }
    fblock_partials<3>(s, sh, parts_out);
}

template <typename T>
dim3 ka_grid_dims(const PoissonGrid<T> &P) {
    const int gx = (P.nx + 63) / 64, gy = (P.ny + 15) / 16;
    const int per_plane = gx * gy;
    int gz = std::max(1, std::min(P.nz, (kFusedBlocks + per_plane - 1) / per_plane));
    while (per_plane * gz > kMaxPartials && gz > 1) --gz;
    return dim3(gx, gy, gz);
}

int streaming_blocks(int64_t n) {
    return (int)std::max<int64_t>(1, std::min<int64_t>(kFusedBlocks, (n + kFB - 1) / kFB));
}

}  // namespace

template <typename T>
int launch_pcg0(const Launch &L, int64_t n, const T *b, const T *Ax0, T *r, T *z, const T *invD,
                double *parts_out) {
    const int nb = streaming_blocks(n);
    hipLaunchKernelGGL(k_pcg0<T>, dim3(nb), dim3(kFB), 0, L.s, n, b, Ax0, r, z, invD, parts_out, L.dsc);
    return nb;
}

template <typename T>
int launch_ka_grid(const Launch &L, const PoissonGrid<T> &P, const PcgIterArgs &a, const T *z,
                   const T *pold, T *pnew, T *q) {
    const dim3 g = ka_grid_dims(P);
    hipLaunchKernelGGL(k_ka_grid<T>, g, dim3(kFB), 0, L.s, P, a, z, pold, pnew, q, L.dsc);
    return (int)(g.x * g.y * g.z);
}

template <typename T>
int launch_ka_sell(const Launch &L, const SellDev<T> &A, const PcgIterArgs &a, const T *z,
                   const T *pold, T *pnew, T *q) {
    const int nb = std::max(1, std::min((A.n_slices + kFB / kWave - 1) / (kFB / kWave), kFusedBlocks));
    hipLaunchKernelGGL(k_ka_sell<T>, dim3(nb), dim3(kFB), 0, L.s, A, a, z, pold, pnew, q, L.dsc);
    return nb;
}

template <typename T>
int launch_kb(const Launch &L, int64_t n, const PcgIterArgs &a, T *x, T *r, T *z, const T *p,
              const T *q, const T *invD) {
    const int nb = streaming_blocks(n);
    hipLaunchKernelGGL(k_kb<T>, dim3(nb), dim3(kFB), 0, L.s, n, a, x, r, z, p, q, invD, L.dsc);
    return nb;
}

template int launch_pcg0<double>(const Launch &, int64_t, const double *, const double *, double *, double *,
                                  const double *, double *);
template int launch_pcg0<float>(const Launch &, int64_t, const float *, const float *, float *, float *,
                                 const float *, double *);
template int launch_ka_grid<double>(const Launch &, const PoissonGrid<double> &, const PcgIterArgs &,
                                    const double *, const double *, double *, double *);
template int launch_ka_grid<float>(const Launch &, const PoissonGrid<float> &, const PcgIterArgs &,
                                   const float *, const float *, float *, float *);
template int launch_ka_sell<double>(const Launch &, const SellDev<double> &, const PcgIterArgs &,
                                    const double *, const double *, double *, double *);
template int launch_ka_sell<float>(const Launch &, const SellDev<float> &, const PcgIterArgs &,
                                   const float *, const float *, float *, float *);
template int launch_kb<double>(const Launch &, int64_t, const PcgIterArgs &, double *, double *, double *,
                               const double *, const double *, const double *);
template int launch_kb<float>(const Launch &, int64_t, const PcgIterArgs &, float *, float *, float *,
                              const float *, const float *, const float *);

}  // namespace mm
