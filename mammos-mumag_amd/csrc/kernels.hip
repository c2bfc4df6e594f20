// HIP kernels for gfx950 (CDNA4): the per-field-step hot path of mammos-mumag.
//
// Streaming kernels use 256-thread blocks (4 wavefronts of 64).  Every scalar that
// an iteration needs (alpha, beta, convergence flags) is produced on the device by
// the "last block" of the preceding reduction, so solver iterations are launched
// back to back without a host round trip.  Reductions are deterministic: each block
// writes one partial per output, and the last block to finish sums the partials in
// a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.h"

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

__device__ __forceinline__ double load_agent(const double *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Reduce NV values per thread over the grid.  Bit i of maxmask selects max instead
// of sum for slot i.  The last block calls fin(vals) on thread 0 with the totals.
// Must be reached by every thread of every block (blockDim.x == kBlock).
template <int NV, typename Fin>
__device__ void grid_reduce(double (&v)[NV], unsigned maxmask, const Reducer &R, Fin fin) {
    __shared__ double sh[NV][kBlock / kWave];
    __shared__ int is_last;
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
            for (int k = 1; k < kBlock / kWave; ++k)
                a = ((maxmask >> i) & 1u) ? fmax(a, sh[i][k]) : a + sh[i][k];
            R.partials[i * kMaxPartials + blockIdx.x] = a;
        }
        __threadfence();
        unsigned prev = atomicAdd(R.counter, 1u);
        is_last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    double tot[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const bool mx = (maxmask >> i) & 1u;
        double a = mx ? 0.0 : 0.0;
        for (int j = threadIdx.x; j < (int)gridDim.x; j += kBlock) {
            double p = load_agent(&R.partials[i * kMaxPartials + j]);
            a = mx ? fmax(a, p) : a + p;
        }
        a = mx ? wave_max(a) : wave_sum(a);
        __syncthreads();
        if (lane == 0) sh[i][w] = a;
        __syncthreads();
        a = sh[i][0];
        for (int k = 1; k < kBlock / kWave; ++k) a = mx ? fmax(a, sh[i][k]) : a + sh[i][k];
        tot[i] = a;
    }
    if (threadIdx.x == 0) {
        fin(tot);
        *R.counter = 0u;
        __threadfence();
    }
}

static inline int grid_for(int64_t n, int per_thread = 1) {
    int64_t b = (n + (int64_t)kBlock * per_thread - 1) / ((int64_t)kBlock * per_thread);
    if (b < 1) b = 1;
    if (b > kMaxPartials) b = kMaxPartials;
    return (int)b;
}

// ------------------------------------------------------------------ SELL SpMV

template <typename T, typename TX>
__global__ __launch_bounds__(kBlock) void k_sell_spmv(SellDev<T> A, const TX *__restrict__ x,
                                                      TX *__restrict__ y, double *dot_out,
                                                      const int *skip, Reducer R) {
    if (skip && *skip) return;
    const int lane = threadIdx.x & (kWave - 1);
    double dot = 0.0;
    for (int slice = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; slice < A.n_slices;
         slice += gridDim.x * (kBlock / kWave)) {
        const int beg = A.slice_ptr[slice], end = A.slice_ptr[slice + 1];
        const int row = slice * kWave + lane;
        T acc = 0;
        for (int k = beg + lane; k < end; k += kWave) acc += A.val[k] * (T)x[A.col[k]];
        if (row < A.n_rows) {
            y[row] = (TX)acc;
            dot += (double)acc * (double)x[row];
        }
    }
    if (dot_out) {
        double v[1] = {dot};
        grid_reduce<1>(v, 0u, R, [=](double *t) { *dot_out = t[0]; });
    }
}

void launch_sell_spmv(const Launch &L, const SellDev<double> &A, const double *x, double *y,
                      double *dot_out, const int *skip_flag) {
    int nb = std::min((A.n_slices + kBlock / kWave - 1) / (kBlock / kWave), kMaxPartials);
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL((k_sell_spmv<double, double>), dim3(nb), dim3(kBlock), 0, L.s, A, x, y,
                       dot_out, skip_flag, L.red);
}
void launch_sell_spmv32(const Launch &L, const SellDev<float> &A, const float *x, float *y,
                        double *dot_out, const int *skip_flag) {
    int nb = std::min((A.n_slices + kBlock / kWave - 1) / (kBlock / kWave), kMaxPartials);
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL((k_sell_spmv<float, float>), dim3(nb), dim3(kBlock), 0, L.s, A, x, y,
                       dot_out, skip_flag, L.red);
}

// ------------------------------------------------------------------ PCG, FP64
// jax.scipy.sparse.linalg.cg with M(x) = x * (1/D):
//   r = b - A x0; z = M r; p = z; gamma = r.z; bound = max(tol^2 b.b, atol^2)
//   while r.r > bound and k < maxiter: q = A p; alpha = gamma/(p.q); x += alpha p;
//       r -= alpha q; z = M r; gamma' = r.z; p = z + (gamma'/gamma) p; k += 1

__global__ __launch_bounds__(kBlock) void k_pcg_init(int n, const double *__restrict__ b,
                                                     const double *__restrict__ Ax0,
                                                     double *__restrict__ r, double *__restrict__ z,
                                                     double *__restrict__ p,
                                                     const double *__restrict__ invD, double tol,
                                                     double atol, int maxiter, DevScalars *dsc,
                                                     Reducer R) {
    double s[3] = {0.0, 0.0, 0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double bi = b[i];
        const double ri = bi - Ax0[i];
        const double zi = ri * invD[i];
        r[i] = ri;
        z[i] = zi;
        p[i] = zi;
        s[0] += ri * zi;
        s[1] += ri * ri;
        s[2] += bi * bi;
    }
    grid_reduce<3>(s, 0u, R, [=](double *t) {
        dsc->pcg_gamma = t[0];
        dsc->pcg_rr = t[1];
        dsc->pcg_bs = t[2];
        const double bound = fmax(tol * tol * t[2], atol * atol);
        dsc->pcg_bound = bound;
        dsc->pcg_iter = 0;
        dsc->pcg_maxiter = maxiter;
        dsc->pcg_done = !((t[1] > bound) && (0 < maxiter));
    });
}

__global__ __launch_bounds__(kBlock) void k_pcg_update(int n, double *__restrict__ x,
                                                       double *__restrict__ r,
                                                       double *__restrict__ z,
                                                       const double *__restrict__ p,
                                                       const double *__restrict__ q,
                                                       const double *__restrict__ invD,
                                                       DevScalars *dsc, Reducer R) {
    if (dsc->pcg_done) return;
    const double alpha = dsc->pcg_gamma / dsc->pcg_pq;
    double s[2] = {0.0, 0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double pi = p[i];
        x[i] = x[i] + alpha * pi;
        const double ri = r[i] - alpha * q[i];
        const double zi = ri * invD[i];
        r[i] = ri;
        z[i] = zi;
        s[0] += ri * zi;
        s[1] += ri * ri;
    }
    grid_reduce<2>(s, 0u, R, [=](double *t) {
        dsc->pcg_beta = t[0] / dsc->pcg_gamma;
        dsc->pcg_gamma = t[0];
        dsc->pcg_rr = t[1];
        const int it = dsc->pcg_iter + 1;
        dsc->pcg_iter = it;
        dsc->pcg_done = !((t[1] > dsc->pcg_bound) && (it < dsc->pcg_maxiter));
    });
}

__global__ __launch_bounds__(kBlock) void k_pcg_pupdate(int n, double *__restrict__ p,
                                                        const double *__restrict__ z,
                                                        const DevScalars *dsc) {
    if (dsc->pcg_done) return;
    const double beta = dsc->pcg_beta;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        p[i] = z[i] + beta * p[i];
}

void launch_pcg_init(const Launch &L, int n, const double *b, const double *Ax0, double *r,
                     double *z, double *p, const double *invD, double tol, double atol,
                     int maxiter) {
    hipLaunchKernelGGL(k_pcg_init, dim3(grid_for(n, 4)), dim3(kBlock), 0, L.s, n, b, Ax0, r, z, p,
                       invD, tol, atol, maxiter, L.dsc, L.red);
}
void launch_pcg_update(const Launch &L, int n, double *x, double *r, double *z, const double *p,
                       const double *q, const double *invD) {
    hipLaunchKernelGGL(k_pcg_update, dim3(grid_for(n, 4)), dim3(kBlock), 0, L.s, n, x, r, z, p, q,
                       invD, L.dsc, L.red);
}
void launch_pcg_pupdate(const Launch &L, int n, double *p, const double *z) {
    hipLaunchKernelGGL(k_pcg_pupdate, dim3(grid_for(n, 4)), dim3(kBlock), 0, L.s, n, p, z, L.dsc);
}

// ------------------------------------------------------------------ element kernels
// Element e with nodes t0..t3, barycentric gradients G_k, volume vol, material m:
//   exchange:    y_k,a += (2A/V) vol  G_k . grad(v_a)
//   anisotropy:  y_k   += (-2K/V) (vol/20) u (u.v_k + sum_l u.v_l)
//   divergence:  b_k   += Js vol G_k . mean(v)
//   gradient:    g_k   += Js vol/(4V) grad(u)        (same for the 4 nodes)

struct ElemLocal {
    double G[4][3];
};

__device__ __forceinline__ void load_G(const ElemDev &E, int e, ElemLocal &L) {
    const int n = E.n;
#pragma unroll
    for (int k = 1; k < 4; ++k)
#pragma unroll
        for (int a = 0; a < 3; ++a) L.G[k][a] = E.G[(3 * (k - 1) + a) * n + e];
#pragma unroll
    for (int a = 0; a < 3; ++a) L.G[0][a] = -(L.G[1][a] + L.G[2][a] + L.G[3][a]);
}

__global__ __launch_bounds__(kBlock) void k_ebe_exani(ElemDev E, const MatCoef *__restrict__ mats,
                                                      const double *__restrict__ v,
                                                      double *__restrict__ y,
                                                      double *__restrict__ b, const int *skip) {
    if (skip && *skip) return;
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= E.n) return;
    const int4 t = E.tets[e];
    const int tn[4] = {t.x, t.y, t.z, t.w};
    ElemLocal L;
    load_G(E, e, L);
    const double vol = E.vol[e];
    const MatCoef mc = mats[E.mat[e]];
    double vl[4][3];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int a = 0; a < 3; ++a) vl[k][a] = v[3 * (int64_t)tn[k] + a];
    double gv[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            gv[a][c] = L.G[0][c] * vl[0][a] + L.G[1][c] * vl[1][a] + L.G[2][c] * vl[2][a] +
                       L.G[3][c] * vl[3][a];
    const double cex = mc.cex * vol;
    const double ck = mc.cK * vol / 20.0;
    double s[4], S = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        s[k] = mc.u[0] * vl[k][0] + mc.u[1] * vl[k][1] + mc.u[2] * vl[k][2];
        S += s[k];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double yk = cex * (L.G[k][0] * gv[a][0] + L.G[k][1] * gv[a][1] + L.G[k][2] * gv[a][2]) +
                              ck * mc.u[a] * (s[k] + S);
            unsafeAtomicAdd(&y[3 * (int64_t)tn[k] + a], yk);
        }
    }
    if (b) {
        const double js = mc.Js * vol;
        double mb[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) mb[a] = 0.25 * (vl[0][a] + vl[1][a] + vl[2][a] + vl[3][a]);
#pragma unroll
        for (int k = 0; k < 4; ++k)
            unsafeAtomicAdd(&b[tn[k]], js * (L.G[k][0] * mb[0] + L.G[k][1] * mb[1] + L.G[k][2] * mb[2]));
    }
}

__global__ __launch_bounds__(kBlock) void k_ebe_grad(ElemDev E, const MatCoef *__restrict__ mats,
                                                     const double *__restrict__ u,
                                                     double *__restrict__ g, double invV) {
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
    for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int a = 0; a < 3; ++a) unsafeAtomicAdd(&g[3 * (int64_t)tn[k] + a], gu[a]);
}

void launch_ebe_exani(const Launch &L, const ElemDev &E, const double *v, double *y, double *b,
                      const int *skip_flag) {
    if (E.n == 0) return;
    hipLaunchKernelGGL(k_ebe_exani, dim3((E.n + kBlock - 1) / kBlock), dim3(kBlock), 0, L.s, E,
                       L.mats, v, y, b, skip_flag);
}
void launch_ebe_grad(const Launch &L, const ElemDev &E, const double *u, double *g, double invV) {
    if (E.n == 0) return;
    hipLaunchKernelGGL(k_ebe_grad, dim3((E.n + kBlock - 1) / kBlock), dim3(kBlock), 0, L.s, E,
                       L.mats, u, g, invV);
}

// ------------------------------------------------------------------ node kernels

__device__ __forceinline__ void cross3(const double *a, const double *b, double *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// g = gz + gC (+ gd); E = m.gz + m.gC/2 (+ m.gd/2); g_proj = m x (g x m)
__global__ __launch_bounds__(kBlock) void k_finish_eg(int n, const double *__restrict__ m,
                                                      const double *__restrict__ meas, double zc0,
                                                      double zc1, double zc2,
                                                      const double *__restrict__ gC,
                                                      const double *__restrict__ gd,
                                                      double *__restrict__ g_full,
                                                      double *__restrict__ g_proj, double *E_out,
                                                      int with_hmag, Reducer R) {
    double e = 0.0;
    const double zc[3] = {zc0, zc1, zc2};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double mi = meas[i];
        double mm[3], g[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            mm[a] = m[3 * (int64_t)i + a];
            const double gz = zc[a] * mi;
            const double gc = gC[3 * (int64_t)i + a];
            const double gdm = with_hmag ? gd[3 * (int64_t)i + a] : 0.0;
            g[a] = gz + gc + gdm;
            e += mm[a] * gz + 0.5 * mm[a] * gc + 0.5 * mm[a] * gdm;
            g_full[3 * (int64_t)i + a] = g[a];
        }
        double c[3], pr[3];
        cross3(g, mm, c);
        cross3(mm, c, pr);
#pragma unroll
        for (int a = 0; a < 3; ++a) g_proj[3 * (int64_t)i + a] = pr[a];
    }
    double v[1] = {e};
    grid_reduce<1>(v, 0u, R, [=](double *t) { *E_out = t[0]; });
}

void launch_finish_eg(const Launch &L, int n, const double *m, const double *meas, double zc0,
                      double zc1, double zc2, const double *gC, const double *gd, double *g_full,
                      double *g_proj, double *E_out, int with_hmag) {
    hipLaunchKernelGGL(k_finish_eg, dim3(grid_for(n, 2)), dim3(kBlock), 0, L.s, n, m, meas, zc0,
                       zc1, zc2, gC, gd, g_full, g_proj, E_out, with_hmag, L.red);
}

__global__ __launch_bounds__(kBlock) void k_external(int n, const double *__restrict__ m,
                                                     const double *__restrict__ meas, double zc0,
                                                     double zc1, double zc2,
                                                     double *__restrict__ g, double *E_out,
                                                     Reducer R) {
    double e = 0.0;
    const double zc[3] = {zc0, zc1, zc2};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double mi = meas[i];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double gz = zc[a] * mi;
            if (g) g[3 * (int64_t)i + a] = gz;
            e += m[3 * (int64_t)i + a] * gz;
        }
    }
    double v[1] = {e};
    grid_reduce<1>(v, 0u, R, [=](double *t) { *E_out = t[0]; });
}
void launch_external(const Launch &L, int n, const double *m, const double *meas, double zc0,
                     double zc1, double zc2, double *g, double *E_out) {
    hipLaunchKernelGGL(k_external, dim3(grid_for(n, 2)), dim3(kBlock), 0, L.s, n, m, meas, zc0,
                       zc1, zc2, g, E_out, L.red);
}

__global__ __launch_bounds__(kBlock) void k_project(int n, const double *__restrict__ m,
                                                    const double *__restrict__ g,
                                                    double *__restrict__ out) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double mm[3], gg[3], c[3], pr[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        mm[a] = m[3 * (int64_t)i + a];
        gg[a] = g[3 * (int64_t)i + a];
    }
    cross3(gg, mm, c);
    cross3(mm, c, pr);
#pragma unroll
    for (int a = 0; a < 3; ++a) out[3 * (int64_t)i + a] = pr[a];
}
void launch_project(const Launch &L, int n, const double *m, const double *g, double *out) {
    hipLaunchKernelGGL(k_project, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, L.s, n, m, g,
                       out);
}

// ------------------------------------------------------------------ truncated CG (M_inv)
// H v = m0 x ((C v) x m0) - (m0 . gF)_i v_i ;  den = d . H d

__global__ __launch_bounds__(kBlock) void k_hvec(int n, const double *__restrict__ m0,
                                                 const double *__restrict__ gF,
                                                 const double *__restrict__ Cd,
                                                 const double *__restrict__ d,
                                                 double *__restrict__ Hd, double *den_out,
                                                 const int *skip, Reducer R) {
    if (skip && *skip) return;
    double den = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        double mm[3], cd[3], dd[3], c[3], pr[3];
        double mg = 0.0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            mm[a] = m0[3 * (int64_t)i + a];
            cd[a] = Cd[3 * (int64_t)i + a];
            dd[a] = d[3 * (int64_t)i + a];
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) mg += mm[a] * gF[3 * (int64_t)i + a];
        cross3(cd, mm, c);
        cross3(mm, c, pr);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double h = pr[a] - mg * dd[a];
            Hd[3 * (int64_t)i + a] = h;
            den += dd[a] * h;
        }
    }
    double v[1] = {den};
    grid_reduce<1>(v, 0u, R, [=](double *t) { *den_out = t[0]; });
}
void launch_hvec(const Launch &L, int n, const double *m0, const double *gF, const double *Cd,
                 const double *d, double *Hd, double *den_out, const int *skip_flag) {
    hipLaunchKernelGGL(k_hvec, dim3(grid_for(n, 2)), dim3(kBlock), 0, L.s, n, m0, gF, Cd, d, Hd,
                       den_out, skip_flag, L.red);
}

// grad_fk = -g: p = 0, r = grad_fk, z = r / D, d = -z; eps = min(0.5, sqrt(|g|)|g|)
__global__ __launch_bounds__(kBlock) void k_tcg_init(int n3, const double *__restrict__ g,
                                                     const double *__restrict__ D,
                                                     double *__restrict__ p, double *__restrict__ r,
                                                     double *__restrict__ z, double *__restrict__ d,
                                                     DevScalars *dsc, Reducer R) {
    double s[2] = {0.0, 0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n3; i += gridDim.x * kBlock) {
        const double ri = -g[i];
        const double zi = ri / D[i];
        p[i] = 0.0;
        r[i] = ri;
        z[i] = zi;
        d[i] = -zi;
        s[0] += ri * ri;
        s[1] += ri * zi;
    }
    grid_reduce<2>(s, 0u, R, [=](double *t) {
        const double nrm = sqrt(t[0]);
        dsc->tcg_eps = fmin(0.5, sqrt(nrm) * nrm);
        dsc->tcg_rz = t[1];
        dsc->tcg_done = 0;
        dsc->tcg_neg = 0;
        dsc->tcg_iter = 0;
    });
}

__global__ __launch_bounds__(kBlock) void k_tcg_update(int n3, double *__restrict__ p,
                                                       double *__restrict__ r,
                                                       double *__restrict__ z,
                                                       const double *__restrict__ d,
                                                       const double *__restrict__ Hd,
                                                       const double *__restrict__ D,
                                                       DevScalars *dsc, Reducer R) {
    if (dsc->tcg_done) return;
    const double den = dsc->tcg_den;
    const bool neg = den <= 0.0;
    const double alpha = neg ? 0.0 : dsc->tcg_rz / den;
    double s[2] = {0.0, 0.0};
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n3; i += gridDim.x * kBlock) {
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
    grid_reduce<2>(s, 0u, R, [=](double *t) {
        const bool conv = sqrt(t[0]) <= dsc->tcg_eps;
        const double rz = dsc->tcg_rz;
        dsc->tcg_beta = (!neg && !conv) ? t[1] / (rz + 1e-16) : 0.0;
        if (!neg) dsc->tcg_rz = t[1];
        dsc->tcg_neg = neg;
        dsc->tcg_done = neg || conv;
        dsc->tcg_iter += 1;
        dsc->red[15] += 1.0;  // cumulative truncated-CG iterations
    });
}

__global__ __launch_bounds__(kBlock) void k_tcg_dupdate(int n3, double *__restrict__ d,
                                                        const double *__restrict__ z,
                                                        const DevScalars *dsc) {
    if (dsc->tcg_done) return;
    const double beta = dsc->tcg_beta;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n3; i += gridDim.x * kBlock)
        d[i] = -z[i] + beta * d[i];
}

// z_out = (all p == 0) ? -grad_fk (= g) : p
__global__ __launch_bounds__(kBlock) void k_tcg_final(int n3, const double *__restrict__ p,
                                                      const double *__restrict__ g,
                                                      double *__restrict__ zout,
                                                      const DevScalars *dsc) {
    const bool zero = dsc->tcg_pmax == 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n3; i += gridDim.x * kBlock)
        zout[i] = zero ? g[i] : p[i];
}

void launch_tcg_init(const Launch &L, int n3, const double *g, const double *D, double *p,
                     double *r, double *z, double *d) {
    hipLaunchKernelGGL(k_tcg_init, dim3(grid_for(n3, 4)), dim3(kBlock), 0, L.s, n3, g, D, p, r, z,
                       d, L.dsc, L.red);
}
void launch_tcg_update(const Launch &L, int n3, double *p, double *r, double *z, double *d,
                       const double *Hd, const double *D) {
    hipLaunchKernelGGL(k_tcg_update, dim3(grid_for(n3, 4)), dim3(kBlock), 0, L.s, n3, p, r, z, d,
                       Hd, D, L.dsc, L.red);
}
void launch_tcg_dupdate(const Launch &L, int n3, double *d, const double *z) {
    hipLaunchKernelGGL(k_tcg_dupdate, dim3(grid_for(n3, 4)), dim3(kBlock), 0, L.s, n3, d, z, L.dsc);
}
void launch_tcg_final(const Launch &L, int n3, const double *p, const double *g, double *zout) {
    hipLaunchKernelGGL(k_tcg_final, dim3(grid_for(n3, 4)), dim3(kBlock), 0, L.s, n3, p, g, zout,
                       L.dsc);
}

// ------------------------------------------------------------------ vector kernels

// out_i = normalize(x_i + alpha d_i), zero-length vectors divided by 1 (jax_tools.py:4-18)
__global__ __launch_bounds__(kBlock) void k_update_m(int n, const double *__restrict__ x,
                                                     double alpha, const double *__restrict__ d,
                                                     double *__restrict__ out) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double v[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) v[a] = x[3 * (int64_t)i + a] + alpha * d[3 * (int64_t)i + a];
    double nrm = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (nrm == 0.0) nrm = 1.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) out[3 * (int64_t)i + a] = v[a] / nrm;
}
void launch_update_m(const Launch &L, int n, const double *x, double alpha, const double *d,
                     double *out) {
    hipLaunchKernelGGL(k_update_m, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, L.s, n, x,
                       alpha, d, out);
}

// out may alias y (d = -z + beta d): each element is read before it is written.
__global__ __launch_bounds__(kBlock) void k_axpby(int64_t n, double a, const double *__restrict__ x,
                                                  double b, const double *y, double *out) {
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        out[i] = (y ? a * x[i] + b * y[i] : a * x[i]);
}
void launch_axpby(const Launch &L, int64_t n, double a, const double *x, double b,
                  const double *y, double *out) {
    hipLaunchKernelGGL(k_axpby, dim3(grid_for(n, 4)), dim3(kBlock), 0, L.s, n, a, x, b, y, out);
}

// ------------------------------------------------------------------ generic reductions

template <int OP>
__global__ __launch_bounds__(kBlock) void k_reduce(int64_t n, const double *__restrict__ a,
                                                   const double *__restrict__ b,
                                                   const double *__restrict__ c,
                                                   const double *__restrict__ d, double h0,
                                                   double h1, double h2, double *out, Reducer R) {
    constexpr int NV = (OP == RED_STOP3 || OP == RED_DG3 || OP == RED_H3) ? 3
                       : (OP == RED_GD2 || OP == RED_YZD)                 ? 2
                                                                          : 1;
    constexpr unsigned MASK = OP == RED_NORMINF ? 1u : OP == RED_STOP3 ? 7u : OP == RED_H3 ? 4u : 0u;
    double s[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) s[k] = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
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
    grid_reduce<NV>(s, MASK, R, [=](double *t) {
#pragma unroll
        for (int k = 0; k < NV; ++k) out[k] = t[k];
    });
}

void launch_reduce(const Launch &L, RedOp op, int64_t n, const double *a, const double *b,
                   const double *c, const double *d, double h0, double h1, double h2,
                   double *out) {
    const dim3 g(grid_for(n, 8)), blk(kBlock);
    switch (op) {
#define MM_RED_CASE(OPC)                                                                        \
    case OPC:                                                                                   \
        hipLaunchKernelGGL(k_reduce<OPC>, g, blk, 0, L.s, n, a, b, c, d, h0, h1, h2, out, L.red); \
        break;
        MM_RED_CASE(RED_DOT)
        MM_RED_CASE(RED_NORMINF)
        MM_RED_CASE(RED_STOP3)
        MM_RED_CASE(RED_DG3)
        MM_RED_CASE(RED_H3)
        MM_RED_CASE(RED_GD2)
        MM_RED_CASE(RED_YZD)
        MM_RED_CASE(RED_MH)
        MM_RED_CASE(RED_SS)
#undef MM_RED_CASE
    }
}

__global__ void k_fill_zero(int64_t n, double *__restrict__ x, const int *skip) {
    if (skip && *skip) return;
    for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        x[i] = 0.0;
}
void launch_fill_zero(const Launch &L, int64_t n, double *x, const int *skip_flag) {
    hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(n, 8)), dim3(kBlock), 0, L.s, n, x, skip_flag);
}

}  // namespace mm
