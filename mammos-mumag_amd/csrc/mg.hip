// Geometric multigrid V-cycle for the Poisson stencil on the structured (rectilinear Kuhn)
// grid: the preconditioner of the FP32 inner solve of the mixed-precision PCG (no
// counterpart in the reference, whose potential solve is Jacobi-preconditioned).
//
// Levels: the node set of every axis keeps its even indices and its last index; each
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
// Per coarse index I of an axis, up to 3 (fine index, weight) pairs at [3 I + q].
// Also x1c = invw_c b_c, the coarse level's first smoothing step from zero (invw_c: the
// level's first Chebyshev weight over the diagonal, 0 on the faces).
__global__ __launch_bounds__(256) void k_mg_restrict(int nxc, int nyc, int nzc, int nxf, int nyf,
                                                     const int *__restrict__ rxi, const float *__restrict__ rxw,
                                                     const int *__restrict__ ryi, const float *__restrict__ ryw,
                                                     const int *__restrict__ rzi, const float *__restrict__ rzw,
                                                     const float *__restrict__ rf, float *__restrict__ bc,
                                                     const float *__restrict__ invwc, float *__restrict__ x1c,
                                                     const int *skip) {
    if (skip && *skip) return;
    const int64_t nc = (int64_t)nxc * nyc * nzc;
    const int64_t sf = (int64_t)nxf * nyf;
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < nc;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int I = (int)(idx % nxc), J = (int)((idx / nxc) % nyc), K = (int)(idx / ((int64_t)nxc * nyc));
        float s = 0.f;
        if (!(I == 0 || J == 0 || K == 0 || I == nxc - 1 || J == nyc - 1 || K == nzc - 1)) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float wz = rzw[3 * K + c];
                if (wz == 0.f) continue;
                const int64_t oz = sf * rzi[3 * K + c];
#pragma unroll
                for (int bq = 0; bq < 3; ++bq) {
                    const float wy = ryw[3 * J + bq];
                    if (wy == 0.f) continue;
                    const int64_t oy = oz + (int64_t)nxf * ryi[3 * J + bq];
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        const float wx = rxw[3 * I + a];
                        if (wx != 0.f) s += wz * wy * wx * rf[oy + rxi[3 * I + a]];
                    }
                }
            }
        }
        bc[idx] = s;
        x1c[idx] = s * invwc[idx];
    }
}

// x_f += P x_c: per fine index i of an axis, coarse c0[i] with weight 1 - t[i] and
// c0[i] + 1 with weight t[i] (t = 0 on coarse nodes).  Fine faces stay zero.
__global__ __launch_bounds__(256) void k_mg_prolong(int nxf, int nyf, int nzf, int nxc, int nyc,
                                                    const int *__restrict__ pxc, const float *__restrict__ pxt,
                                                    const int *__restrict__ pyc, const float *__restrict__ pyt,
                                                    const int *__restrict__ pzc, const float *__restrict__ pzt,
                                                    const float *__restrict__ xc, float *__restrict__ xf,
                                                    const int *skip) {
    if (skip && *skip) return;
    const int64_t nf = (int64_t)nxf * nyf * nzf;
    const int64_t sc = (int64_t)nxc * nyc;
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < nf;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(idx % nxf), j = (int)((idx / nxf) % nyf), k = (int)(idx / ((int64_t)nxf * nyf));
        if (i == 0 || j == 0 || k == 0 || i == nxf - 1 || j == nyf - 1 || k == nzf - 1) continue;
        const int cx = pxc[i], cy = pyc[j], cz = pzc[k];
        const float tx = pxt[i], ty = pyt[j], tz = pzt[k];
        float s = 0.f;
#pragma unroll
        for (int dz = 0; dz < 2; ++dz) {
            const float wz = dz ? tz : 1.f - tz;
            if (wz == 0.f) continue;
#pragma unroll
            for (int dy = 0; dy < 2; ++dy) {
                const float wy = dy ? ty : 1.f - ty;
                if (wy == 0.f) continue;
#pragma unroll
                for (int dx = 0; dx < 2; ++dx) {
                    const float wx = dx ? tx : 1.f - tx;
                    if (wx == 0.f) continue;
                    s += wz * wy * wx * xc[(cx + dx) + (int64_t)nxc * (cy + dy) + sc * (cz + dz)];
                }
            }
        }
        xf[idx] += s;
    }
}

// At most kMaxPartials blocks (the fused dot writes one partial per block); the setup
// checks that one z chunk of tiles fits.
dim3 stencil_grid(const PoissonGrid<float> &P) {
    const int gx = (P.nx + 63) / 64, gy = (P.ny + 3) / 4;
    const int64_t per_plane = (int64_t)gx * gy;
    int gz = (int)std::min<int64_t>(P.nz, std::max<int64_t>(1, kMaxPartials / per_plane));
    return dim3(gx, gy, gz);
}

int blocks_for(int64_t n) {
    const int64_t b = (n + 255) / 256;
    return (int)(b < 1 ? 1 : b > 8192 ? 8192 : b);
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
bool mg_stencil_fits(const PoissonGrid<float> &P) {
    return (int64_t)((P.nx + 63) / 64) * ((P.ny + 3) / 4) <= kMaxPartials;
}
void launch_mg_restrict(const Launch &L, const MgLevelDev &f, const MgLevelDev &c, const float *rf, float *bc,
                        float *x1c, const int *skip) {
    const int64_t nc = (int64_t)c.P.nx * c.P.ny * c.P.nz;
    hipLaunchKernelGGL(k_mg_restrict, dim3(blocks_for(nc)), dim3(256), 0, L.s, c.P.nx, c.P.ny, c.P.nz, f.P.nx,
                       f.P.ny, f.rxi, f.rxw, f.ryi, f.ryw, f.rzi, f.rzw, rf, bc, c.invw, x1c, skip);
}
void launch_mg_prolong(const Launch &L, const MgLevelDev &f, const MgLevelDev &c, const float *xc, float *xf,
                       const int *skip) {
    const int64_t nf = (int64_t)f.P.nx * f.P.ny * f.P.nz;
    hipLaunchKernelGGL(k_mg_prolong, dim3(blocks_for(nf)), dim3(256), 0, L.s, f.P.nx, f.P.ny, f.P.nz, c.P.nx,
                       c.P.ny, f.pxc, f.pxt, f.pyc, f.pyt, f.pzc, f.pzt, xc, xf, skip);
}

}  // namespace mm
