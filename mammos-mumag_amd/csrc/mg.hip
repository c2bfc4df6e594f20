// Geometric multigrid V-cycle for the Poisson stencil on the structured (rectilinear Kuhn)
// grid: the preconditioner of the FP32 inner solve of the mixed-precision PCG (no
// counterpart in the reference, whose potential solve is Jacobi-preconditioned).
//
// Levels: spacing-aware coarsening per axis (two adjacent intervals merge while both are
// at most 3 times the level's smallest spacing), down to at most 16 nodes per axis; each
// level has the 7-point operator of its own rectilinear grid (same coefficients as
// k_grid_poisson), Dirichlet (zero) on the six faces.  Transfers are per-axis linear
// interpolation (prolongation) and its transpose (restriction).  Smoothing: Chebyshev-
// Richardson steps x += w_k D^-1 (b - A x) with w_k the reciprocal Chebyshev nodes of
// [2/30, 2] (2 bounds the spectrum of D^-1 A by Gershgorin), the post-smoother in the
// reverse order, so that the V-cycle is a symmetric operator; the coarsest level is a
// fixed Chebyshev-Richardson polynomial.  FP32 vectors, FP32 coefficients.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.h"
#include "mg.h"

namespace mm {

namespace {

template <typename T>
__device__ __forceinline__ T edge_sum_mg(const T *__restrict__ h1p, const T *__restrict__ h2p, int a, int b) {
    const T l1 = h1p[a], u1 = h1p[a + 1], l2 = h2p[b], u2 = h2p[b + 1];
    return (T)2 * l1 * l2 + l1 * u2 + u1 * l2 + (T)2 * u1 * u2;
}

// OP_FIRST:  out = w b / diag                (x = 0)
// OP_SMOOTH: out = x + w (b - A x) / diag
// OP_RESID:  out = b - A x
// Zero on the faces.  Tiles of 64 x 4 (x, y) nodes, each thread marching a z chunk with a
// sliding window for the z neighbours.  DOT (OP_SMOOTH): partials of b.out and b.b per
// block into parts[0 / kMaxPartials + block] (the next PCG step's r.z and r.r).
enum { OP_FIRST = 0, OP_SMOOTH = 1, OP_RESID = 2 };
template <int OP, bool DOT>
__global__ __launch_bounds__(256) void k_mg_stencil(PoissonGrid<float> P, const float *__restrict__ b,
                                                    const float *__restrict__ x, float *__restrict__ out, float w,
                                                    const int *skip, double *parts) {
    if (skip && *skip) return;   // block-uniform
    const int nx = P.nx, ny = P.ny, nz = P.nz;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int chunk = (nz + gridDim.z - 1) / gridDim.z;
    const int k0 = blockIdx.z * chunk, k1 = min(nz, k0 + chunk);
    const bool active = i < nx && j < ny && k0 < k1;
    if (!DOT && !active) return;
    double sbz = 0.0, sbb = 0.0;
    if (active) {
        const int64_t sy = nx, sz = (int64_t)nx * ny;
        const int64_t base = i + sy * j;
        const bool edge_ij = i == 0 || j == 0 || i == nx - 1 || j == ny - 1;
        float szij = 0.f, ixm = 0.f, ixp = 0.f, iym = 0.f, iyp = 0.f;
        if (!edge_ij) {
            szij = edge_sum_mg(P.hxp, P.hyp, i, j);
            ixm = P.ixp[i];
            ixp = P.ixp[i + 1];
            iym = P.iyp[j];
            iyp = P.iyp[j + 1];
        }
        float xm = 0.f, xc = 0.f;
        if (OP != OP_FIRST) {
            xm = k0 > 0 ? x[base + sz * (k0 - 1)] : 0.f;
            xc = x[base + sz * k0];
        }
        for (int k = k0; k < k1; ++k) {
            const int64_t idx = base + sz * k;
            float xp = 0.f;
            if (OP != OP_FIRST) xp = k + 1 < nz ? x[idx + sz] : 0.f;
            float o = 0.f;
            if (!(edge_ij || k == 0 || k == nz - 1)) {
                const float sx = edge_sum_mg(P.hyp, P.hzp, j, k);
                const float syy = edge_sum_mg(P.hzp, P.hxp, k, i);
                const float izm = P.izp[k], izq = P.izp[k + 1];
                const float diag = sx * (ixm + ixp) + syy * (iym + iyp) + szij * (izm + izq);
                const float bi = b[idx];
                if (OP == OP_FIRST) {
                    o = w * bi / diag;
                } else {
                    // faces hold zero, so their (dropped) couplings contribute nothing
                    const float xw = x[idx - 1], xe = x[idx + 1], xs = x[idx - sy], xn = x[idx + sy];
                    const float ax = diag * xc - sx * (ixm * xw + ixp * xe) - syy * (iym * xs + iyp * xn) -
                                     szij * (izm * xm + izq * xp);
                    o = OP == OP_RESID ? bi - ax : xc + w * (bi - ax) / diag;
                }
                if (DOT) {
                    sbz += (double)bi * (double)o;
                    sbb += (double)bi * (double)bi;
                }
            } else if (DOT) {   // r.r over every node, as the plain PCG forms it
                const double bf = b[idx];
                sbb += bf * bf;
            }
            out[idx] = o;
            xm = xc;
            xc = xp;
        }
    }
    if (DOT) {
        __shared__ double sh[2][256 / kWave];
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) {
            sbz += __shfl_xor(sbz, o, kWave);
            sbb += __shfl_xor(sbb, o, kWave);
        }
        const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
        if (lane == 0) {
            sh[0][wv] = sbz;
            sh[1][wv] = sbb;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double a = 0.0, c = 0.0;
#pragma unroll
            for (int q = 0; q < 256 / kWave; ++q) {
                a += sh[0][q];
                c += sh[1][q];
            }
            const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
            parts[bid] = a;
            parts[kMaxPartials + bid] = c;
        }
    }
}

// b_c = R r_f (transpose of the per-axis linear interpolation); zero on the coarse faces.
// Per coarse index I of an axis, up to 3 (fine index, weight) pairs at [3 I + q] (unused
// pairs: index 0, weight 0).  Also x1c = invw_c b_c, the coarse level's first smoothing
// step from zero (invw_c: the level's first Chebyshev weight over the diagonal, 0 on the
// faces).  Tiles of 64 x 4 coarse (x, y) columns, each thread marching a z chunk; the
// z and y pairs are uniform over a wave, the x pairs per lane.
__global__ __launch_bounds__(256) void k_mg_restrict(int nxc, int nyc, int nzc, int nxf, int nyf,
                                                     const int *__restrict__ rxi, const float *__restrict__ rxw,
                                                     const int *__restrict__ ryi, const float *__restrict__ ryw,
                                                     const int *__restrict__ rzi, const float *__restrict__ rzw,
                                                     const float *__restrict__ rf, float *__restrict__ bc,
                                                     const float *__restrict__ invwc, float *__restrict__ x1c,
                                                     const int *skip) {
    if (skip && *skip) return;
    const int I = blockIdx.x * 64 + (threadIdx.x & 63);
    const int J = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int chunk = (nzc + gridDim.z - 1) / gridDim.z;
    const int k0 = blockIdx.z * chunk, k1 = min(nzc, k0 + chunk);
    if (I >= nxc || J >= nyc) return;
    const bool edge = I == 0 || J == 0 || I == nxc - 1 || J == nyc - 1;
    const int64_t sf = (int64_t)nxf * nyf, sc = (int64_t)nxc * nyc;
    int xi[3], yo[3];
    float xw[3], yw[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        xi[q] = rxi[3 * I + q];
        xw[q] = rxw[3 * I + q];
        yo[q] = nxf * ryi[3 * J + q];
        yw[q] = ryw[3 * J + q];
    }
    for (int K = k0; K < k1; ++K) {
        const int64_t idx = I + (int64_t)nxc * J + sc * K;
        float s = 0.f;
        if (!(edge || K == 0 || K == nzc - 1)) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float wz = rzw[3 * K + c];
                if (wz == 0.f) continue;
                const float *pz = rf + sf * rzi[3 * K + c];
                float sz = 0.f;
#pragma unroll
                for (int bq = 0; bq < 3; ++bq) {
                    if (yw[bq] == 0.f) continue;
                    const float *py = pz + yo[bq];
                    sz += yw[bq] * (xw[0] * py[xi[0]] + xw[1] * py[xi[1]] + xw[2] * py[xi[2]]);
                }
                s += wz * sz;
            }
        }
        bc[idx] = s;
        x1c[idx] = s * invwc[idx];
    }
}

// x_f += P x_c on the interior: per fine index i of an axis, coarse c0[i] with weight
// 1 - t[i] and c0[i] + 1 with weight t[i] (t = 0 on coarse nodes; an interior fine node
// has c0 + 1 inside the coarse axis).  Fine faces stay zero.  Tiles of 64 x 4 fine
// columns marching z chunks.
__global__ __launch_bounds__(256) void k_mg_prolong(int nxf, int nyf, int nzf, int nxc, int nyc,
                                                    const int *__restrict__ pxc, const float *__restrict__ pxt,
                                                    const int *__restrict__ pyc, const float *__restrict__ pyt,
                                                    const int *__restrict__ pzc, const float *__restrict__ pzt,
                                                    const float *__restrict__ xc, float *__restrict__ xf,
                                                    const int *skip) {
    if (skip && *skip) return;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int chunk = (nzf + gridDim.z - 1) / gridDim.z;
    const int k0 = max(1, blockIdx.z * chunk), k1 = min(nzf - 1, blockIdx.z * chunk + chunk);
    if (i <= 0 || j <= 0 || i >= nxf - 1 || j >= nyf - 1) return;
    const int64_t sf = (int64_t)nxf * nyf, sc = (int64_t)nxc * nyc;
    const float tx = pxt[i], ty = pyt[j];
    const int64_t oxy = pxc[i] + (int64_t)nxc * pyc[j];
    for (int k = k0; k < k1; ++k) {
        const float tz = pzt[k];
        const float *c0 = xc + oxy + sc * pzc[k];
        const float *c1 = c0 + sc;
        const float a0 = (1.f - ty) * ((1.f - tx) * c0[0] + tx * c0[1]) + ty * ((1.f - tx) * c0[nxc] + tx * c0[nxc + 1]);
        float s = (1.f - tz) * a0;
        if (tz != 0.f) {   // uniform over the block
            const float a1 =
                (1.f - ty) * ((1.f - tx) * c1[0] + tx * c1[1]) + ty * ((1.f - tx) * c1[nxc] + tx * c1[nxc + 1]);
            s += tz * a1;
        }
        const int64_t idx = i + (int64_t)nxf * j + sf * k;
        xf[idx] += s;
    }
}

// ---------------------------------------------------------------- coarse tail in LDS
// The last levels of the V-cycle in one workgroup: each level's vectors live in LDS
// (x and y ping-pong, the right-hand side for all but the first tail level, whose b is
// read from global memory), and every smoothing step, residual and transfer is one
// pass of the workgroup over the level's nodes, separated by barriers.  Same arithmetic
// as the per-level kernels above.
constexpr int kTailThreads = 1024;

__device__ __forceinline__ void tail_stencil(const PoissonGrid<float> &P, const float *b, const float *x,
                                             float *out, float w, int op) {
    const int nx = P.nx, ny = P.ny, nz = P.nz, nxy = nx * ny, n = nxy * nz;
    for (int t = threadIdx.x; t < n; t += kTailThreads) {
        const int i = t % nx, j = (t / nx) % ny, k = t / nxy;
        float o = 0.f;
        if (!(i == 0 || j == 0 || k == 0 || i == nx - 1 || j == ny - 1 || k == nz - 1)) {
            const float sx = edge_sum_mg(P.hyp, P.hzp, j, k);
            const float syy = edge_sum_mg(P.hzp, P.hxp, k, i);
            const float szz = edge_sum_mg(P.hxp, P.hyp, i, j);
            const float ixm = P.ixp[i], ixq = P.ixp[i + 1], iym = P.iyp[j], iyq = P.iyp[j + 1];
            const float izm = P.izp[k], izq = P.izp[k + 1];
            const float diag = sx * (ixm + ixq) + syy * (iym + iyq) + szz * (izm + izq);
            const float bi = b[t];
            if (op == OP_FIRST) {
                o = w * bi / diag;
            } else {
                const float ax = diag * x[t] - sx * (ixm * x[t - 1] + ixq * x[t + 1]) -
                                 syy * (iym * x[t - nx] + iyq * x[t + nx]) - szz * (izm * x[t - nxy] + izq * x[t + nxy]);
                o = op == OP_RESID ? bi - ax : x[t] + w * (bi - ax) / diag;
            }
        }
        out[t] = o;
    }
    __syncthreads();
}

// bc = R rf (f: the fine level with its transfer lists), zero on the coarse faces
__device__ __forceinline__ void tail_restrict(const MgLevelDev &f, const PoissonGrid<float> &C, const float *rf,
                                              float *bc) {
    const int nxc = C.nx, nyc = C.ny, nzc = C.nz, nxyc = nxc * nyc, nc = nxyc * nzc;
    const int nxf = f.P.nx, nxyf = f.P.nx * f.P.ny;
    for (int t = threadIdx.x; t < nc; t += kTailThreads) {
        const int I = t % nxc, J = (t / nxc) % nyc, K = t / nxyc;
        float s = 0.f;
        if (!(I == 0 || J == 0 || K == 0 || I == nxc - 1 || J == nyc - 1 || K == nzc - 1)) {
            for (int c = 0; c < 3; ++c) {
                const float wz = f.rzw[3 * K + c];
                if (wz == 0.f) continue;
                const float *pz = rf + nxyf * f.rzi[3 * K + c];
                float sz = 0.f;
                for (int q = 0; q < 3; ++q) {
                    const float wy = f.ryw[3 * J + q];
                    if (wy == 0.f) continue;
                    const float *py = pz + nxf * f.ryi[3 * J + q];
                    sz += wy * (f.rxw[3 * I] * py[f.rxi[3 * I]] + f.rxw[3 * I + 1] * py[f.rxi[3 * I + 1]] +
                                f.rxw[3 * I + 2] * py[f.rxi[3 * I + 2]]);
                }
                s += wz * sz;
            }
        }
        bc[t] = s;
    }
    __syncthreads();
}

// xf += P xc on the interior of the fine level f
__device__ __forceinline__ void tail_prolong(const MgLevelDev &f, const PoissonGrid<float> &C, const float *xc,
                                             float *xf) {
    const int nx = f.P.nx, ny = f.P.ny, nz = f.P.nz, nxy = nx * ny, n = nxy * nz;
    const int nxc = C.nx, nxyc = C.nx * C.ny;
    for (int t = threadIdx.x; t < n; t += kTailThreads) {
        const int i = t % nx, j = (t / nx) % ny, k = t / nxy;
        if (i == 0 || j == 0 || k == 0 || i == nx - 1 || j == ny - 1 || k == nz - 1) continue;
        const float tx = f.pxt[i], ty = f.pyt[j], tz = f.pzt[k];
        const float *c0 = xc + f.pxc[i] + nxc * f.pyc[j] + nxyc * f.pzc[k];
        const float *c1 = c0 + nxyc;
        const float a0 = (1.f - ty) * ((1.f - tx) * c0[0] + tx * c0[1]) + ty * ((1.f - tx) * c0[nxc] + tx * c0[nxc + 1]);
        const float a1 = (1.f - ty) * ((1.f - tx) * c1[0] + tx * c1[1]) + ty * ((1.f - tx) * c1[nxc] + tx * c1[nxc + 1]);
        xf[t] += (1.f - tz) * a0 + tz * a1;
    }
    __syncthreads();
}

__global__ __launch_bounds__(kTailThreads) void k_mg_tail(MgTailArgs a, const int *skip) {
    if (skip && *skip) return;
    extern __shared__ float lds[];
    float *cur[kMgTailMax], *oth[kMgTailMax];
    const float *B[kMgTailMax];
    for (int q = 0; q < a.nt; ++q) {
        cur[q] = lds + a.off[q];
        oth[q] = lds + a.off[q] + a.n[q];
        B[q] = q == 0 ? a.b0 : lds + a.off[q] + 2 * a.n[q];
    }
    for (int q = 0; q + 1 < a.nt; ++q) {
        const PoissonGrid<float> &P = a.lev[q].P;
        tail_stencil(P, B[q], nullptr, cur[q], a.om[0], OP_FIRST);
        for (int k = 1; k < a.s; ++k) {
            tail_stencil(P, B[q], cur[q], oth[q], a.om[k], OP_SMOOTH);
            float *t = cur[q]; cur[q] = oth[q]; oth[q] = t;
        }
        tail_stencil(P, B[q], cur[q], oth[q], 0.f, OP_RESID);
        tail_restrict(a.lev[q], a.lev[q + 1].P, oth[q], (float *)B[q + 1]);
    }
    {
        const int q = a.nt - 1;
        const PoissonGrid<float> &P = a.lev[q].P;
        tail_stencil(P, B[q], nullptr, cur[q], a.omc[0], OP_FIRST);
        for (int k = 1; k < a.cd; ++k) {
            tail_stencil(P, B[q], cur[q], oth[q], a.omc[k], OP_SMOOTH);
            float *t = cur[q]; cur[q] = oth[q]; oth[q] = t;
        }
    }
    for (int q = a.nt - 2; q >= 0; --q) {
        const PoissonGrid<float> &P = a.lev[q].P;
        tail_prolong(a.lev[q], a.lev[q + 1].P, cur[q + 1], cur[q]);
        for (int k = a.s - 1; k >= 0; --k) {
            tail_stencil(P, B[q], cur[q], oth[q], a.om[k], OP_SMOOTH);
            float *t = cur[q]; cur[q] = oth[q]; oth[q] = t;
        }
    }
    for (int t = threadIdx.x; t < a.n[0]; t += kTailThreads) a.x0[t] = cur[0][t];
}

// At most kMaxPartials blocks (the fused dot writes one partial per block); the setup
// checks that one z chunk of tiles fits.
dim3 stencil_grid(const PoissonGrid<float> &P) {
    const int gx = (P.nx + 63) / 64, gy = (P.ny + 3) / 4;
    const int64_t per_plane = (int64_t)gx * gy;
    int gz = (int)std::min<int64_t>(P.nz, std::max<int64_t>(1, kMaxPartials / per_plane));
    return dim3(gx, gy, gz);
}


}  // namespace

void launch_mg_first(const Launch &L, const PoissonGrid<float> &P, const float *b, float *out, float w,
                     const int *skip) {
    hipLaunchKernelGGL((k_mg_stencil<OP_FIRST, false>), stencil_grid(P), dim3(256), 0, L.s, P, b,
                       (const float *)nullptr, out, w, skip, (double *)nullptr);
}
int launch_mg_smooth(const Launch &L, const PoissonGrid<float> &P, const float *b, const float *x, float *out,
                     float w, const int *skip, double *dot_parts) {
    const dim3 g = stencil_grid(P);
    if (dot_parts)
        hipLaunchKernelGGL((k_mg_stencil<OP_SMOOTH, true>), g, dim3(256), 0, L.s, P, b, x, out, w, skip, dot_parts);
    else
        hipLaunchKernelGGL((k_mg_stencil<OP_SMOOTH, false>), g, dim3(256), 0, L.s, P, b, x, out, w, skip,
                           (double *)nullptr);
    return (int)(g.x * g.y * g.z);
}
void launch_mg_resid(const Launch &L, const PoissonGrid<float> &P, const float *b, const float *x, float *out,
                     const int *skip) {
    hipLaunchKernelGGL((k_mg_stencil<OP_RESID, false>), stencil_grid(P), dim3(256), 0, L.s, P, b, x, out, 0.f, skip,
                       (double *)nullptr);
}
int mg_tail_lds_bytes(const MgTailArgs &a) {
    int f = 0;
    for (int q = 0; q < a.nt; ++q) f += (q == 0 ? 2 : 3) * a.n[q];
    return f * (int)sizeof(float);
}
void launch_mg_tail(const Launch &L, const MgTailArgs &a, const int *skip) {
    hipLaunchKernelGGL(k_mg_tail, dim3(1), dim3(kTailThreads), mg_tail_lds_bytes(a), L.s, a, skip);
}
bool mg_stencil_fits(const PoissonGrid<float> &P) {
    return (int64_t)((P.nx + 63) / 64) * ((P.ny + 3) / 4) <= kMaxPartials;
}
void launch_mg_restrict(const Launch &L, const MgLevelDev &f, const MgLevelDev &c, const float *rf, float *bc,
                        float *x1c, const int *skip) {
    hipLaunchKernelGGL(k_mg_restrict, stencil_grid(c.P), dim3(256), 0, L.s, c.P.nx, c.P.ny, c.P.nz, f.P.nx,
                       f.P.ny, f.rxi, f.rxw, f.ryi, f.ryw, f.rzi, f.rzw, rf, bc, c.invw, x1c, skip);
}
void launch_mg_prolong(const Launch &L, const MgLevelDev &f, const MgLevelDev &c, const float *xc, float *xf,
                       const int *skip) {
    hipLaunchKernelGGL(k_mg_prolong, stencil_grid(f.P), dim3(256), 0, L.s, f.P.nx, f.P.ny, f.P.nz, c.P.nx,
                       c.P.ny, f.pxc, f.pxt, f.pyc, f.pyt, f.pzc, f.pzt, xc, xf, skip);
}

}  // namespace mm
