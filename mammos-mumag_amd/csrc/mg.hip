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
// [2/15, 2] (2 bounds the spectrum of D^-1 A by Gershgorin), the post-smoother in the
// reverse order, so that the V-cycle is a symmetric operator; the coarsest level is a
// fixed Chebyshev-Richardson polynomial (degree 12 on [2/100, 2]).  FP32 vectors, FP32 coefficients.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.h"
#include "mg.h"

namespace mm {

namespace {

// The skip flag of launches that have none (never written): the flag is then one global
// load without a branch (a global, not a flat load, whose wait would also hold the
// kernel-argument loads), so that its test waits only once the kernel's first loads are
// issued.
__device__ int kNoSkip = 0;

// Column tiles of the restriction kernel: kMgTX x kMgTY (x, y) columns per 256-thread block
// (64 x 4 measured faster than 32 x 8 on the 154^3 grid)
constexpr int kMgTX = kMgRestrictTX, kMgTY = 256 / kMgTX;

// Logical tile of this block in a 1-D launch of gx * gy * gz blocks.  Blocks are
// dispatched round-robin over the 8 XCDs (block b on XCD b % 8), so each XCD is given a
// contiguous range of logical tiles, z chunk fastest, then x, then y: the halo rows and
// planes that neighbouring tiles share are then read through one L2.
struct Tile {
    int bx, by, bz;
};
__device__ __forceinline__ Tile mg_tile(const MgGrid &g) {
    const int nb = g.x * g.y * g.z, b = blockIdx.x;
    const int q = nb / 8, r = nb % 8, x = b % 8, l = b / 8;
    const int t = x < r ? x * (q + 1) + l : r * (q + 1) + (x - r) * q + l;
    return Tile{(t / g.z) % g.x, t / (g.z * g.x), t % g.z};
}

template <typename T>
__device__ __forceinline__ T edge_sum_mg(const T *__restrict__ h1p, const T *__restrict__ h2p, int a, int b) {
    const T l1 = h1p[a], u1 = h1p[a + 1], l2 = h2p[b], u2 = h2p[b + 1];
    return (T)2 * l1 * l2 + l1 * u2 + u1 * l2 + (T)2 * u1 * u2;
}

// OP_FIRST:  out = w b / diag                (x = 0)
// OP_SMOOTH: out = x + w (b - A x) / diag
// OP_RESID:  out = b - A x
// Zero on the faces.  Strips of 256 consecutive nodes of a plane, each thread marching a
// z chunk with a sliding window for the z neighbours; the loads of plane k + 1 are issued
// before the arithmetic of plane k.  DOT (OP_SMOOTH): partials of b.out and b.b per
// block into parts[0 / kMaxPartials + block] (the next PCG step's r.z and r.r).
// The skip flag is tested after the first loads are issued (the coarse levels' kernels
// are latency-bound: the flag's load then overlaps theirs).
enum { OP_FIRST = 0, OP_SMOOTH = 1, OP_RESID = 2 };
template <int OP, bool DOT>
__global__ __launch_bounds__(256) void k_mg_stencil(PoissonGrid<float> P, MgGrid G, const float *__restrict__ b,
                                                    const float *__restrict__ x, float *__restrict__ out, float w,
                                                    const int *skip, double *parts) {
    const int sk = *(skip ? skip : &kNoSkip);   // block-uniform
    const int nx = P.nx, ny = P.ny, nz = P.nz;
    const Tile T = mg_tile(G);
    // a strip of 256 consecutive nodes of a plane (G.x strips per plane, G.y = 1)
    const int e = T.bx * 256 + threadIdx.x;
    const int j = e / nx, i = e - j * nx;
    const int chunk = (nz + G.z - 1) / G.z;
    const int k0 = T.bz * chunk, k1 = min(nz, k0 + chunk);
    const bool active = i < nx && j < ny && k0 < k1;
    if (!DOT && !active) return;
    double sbz = 0.0, sbb = 0.0;
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    const int64_t base = i + sy * j;
    const bool edge_ij = i == 0 || j == 0 || i == nx - 1 || j == ny - 1;
    auto inner = [&](int k) { return !(edge_ij || k == 0 || k == nz - 1); };
    // spacings of the column; loads of plane k (and x of plane k + 1), issued one plane
    // ahead of their use
    float hxl = 0.f, hxu = 0.f, hyl = 0.f, hyu = 0.f, ixm = 0.f, ixq = 0.f, iym = 0.f, iyq = 0.f;
    float bk = 0.f, xw = 0.f, xe = 0.f, xs = 0.f, xn = 0.f, xm = 0.f, xc = 0.f, xp = 0.f;
    // z spacings of plane k (hz[k], hz[k + 1], iz[k], iz[k + 1]; n + 1 entries each, so
    // k + 1 <= nz is valid), also one plane ahead
    float hzl = 0.f, hzu = 0.f, izl = 0.f, izu = 0.f;
    auto load_plane = [&](int k, float &bv, float &w_, float &e_, float &s_, float &n_) {
        const int64_t idx = base + sz * k;
        const bool in = inner(k);
        bv = (in || DOT) ? b[idx] : 0.f;
        if (OP != OP_FIRST && in) {
            w_ = x[idx - 1];
            e_ = x[idx + 1];
            s_ = x[idx - sy];
            n_ = x[idx + sy];
        }
    };
    if (active) {
        // straight-line loads (clamped to valid indices, the unused values zeroed after),
        // so that they are all in flight together: the spacing arrays hold n + 1 entries
        // (pads at 0 and n), so i + 1 and j + 1 are valid; the column's factors are unused
        // on an edge column
        hxl = P.hxp[i];
        hxu = P.hxp[i + 1];
        hyl = P.hyp[j];
        hyu = P.hyp[j + 1];
        ixm = P.ixp[i];
        ixq = P.ixp[i + 1];
        iym = P.iyp[j];
        iyq = P.iyp[j + 1];
        hzl = P.hzp[k0];
        hzu = P.hzp[k0 + 1];
        izl = P.izp[k0];
        izu = P.izp[k0 + 1];
        const int64_t idx = base + sz * k0;
        const bool in = inner(k0);
        bk = b[idx];
        if (!(in || DOT)) bk = 0.f;
        if (OP != OP_FIRST) {
            const int64_t dx1 = in ? 1 : 0, dy1 = in ? sy : 0;   // in-bounds neighbours
            xw = x[idx - dx1];
            xe = x[idx + dx1];
            xs = x[idx - dy1];
            xn = x[idx + dy1];
            xm = x[k0 > 0 ? idx - sz : idx];
            xc = x[idx];
            xp = x[k0 + 1 < nz ? idx + sz : idx];
            if (!in) xw = xe = xs = xn = 0.f;
            if (k0 == 0) xm = 0.f;
            if (k0 + 1 >= nz) xp = 0.f;
        }
    }
    // the loads above complete before the flag's test (rather than being moved after it)
    asm volatile("" : "+v"(hxl), "+v"(hxu), "+v"(hyl), "+v"(hyu), "+v"(ixm), "+v"(ixq), "+v"(iym), "+v"(iyq));
    asm volatile("" : "+v"(xm), "+v"(xc), "+v"(xp), "+v"(bk), "+v"(xw), "+v"(xe), "+v"(xs), "+v"(xn));
    asm volatile("" : "+v"(hzl), "+v"(hzu), "+v"(izl), "+v"(izu));
    if (sk) return;
    if (active) {
        // per-column factors: edge sums regrouped per axis, e.g.
        // edge_sum(hz, hx; k, i) = hzl (2 hxl + hxu) + hzu (hxl + 2 hxu)
        const float cx = 2.f * hxl + hxu, dx = hxl + 2.f * hxu;
        const float szz = hxl * (2.f * hyl + hyu) + hxu * (hyl + 2.f * hyu);
        for (int k = k0; k < k1; ++k) {
            const int64_t idx = base + sz * k;
            float bn = 0.f, wn = 0.f, en = 0.f, sn = 0.f, nn = 0.f, xpp = 0.f, hzn = 0.f, izn = 0.f;
            if (k + 1 < k1) {
                load_plane(k + 1, bn, wn, en, sn, nn);
                if (OP != OP_FIRST && k + 2 < nz) xpp = x[idx + 2 * sz];
                hzn = P.hzp[k + 2];   // k + 2 <= k1 <= nz
                izn = P.izp[k + 2];
            }
            float o = 0.f;
            if (inner(k)) {
                const float izm = izl, izq = izu;
                const float sx = hyl * (2.f * hzl + hzu) + hyu * (hzl + 2.f * hzu);
                const float syy = hzl * cx + hzu * dx;
                const float diag = sx * (ixm + ixq) + syy * (iym + iyq) + szz * (izm + izq);
                if (OP == OP_FIRST) {
                    o = w * bk / diag;
                } else {
                    // faces hold zero, so their (dropped) couplings contribute nothing
                    const float ax = diag * xc - sx * (ixm * xw + ixq * xe) - syy * (iym * xs + iyq * xn) -
                                     szz * (izm * xm + izq * xp);
                    o = OP == OP_RESID ? bk - ax : xc + w * (bk - ax) / diag;
                }
                if (DOT) {
                    sbz += (double)bk * (double)o;
                    sbb += (double)bk * (double)bk;
                }
            } else if (DOT) {   // r.r over every node, as the plain PCG forms it
                sbb += (double)bk * (double)bk;
            }
            out[idx] = o;
            xm = xc;
            xc = xp;
            xp = xpp;
            bk = bn;
            xw = wn;
            xe = en;
            xs = sn;
            xn = nn;
            hzl = hzu;
            hzu = hzn;
            izl = izu;
            izu = izn;
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
            parts[blockIdx.x] = a;
            parts[kMaxPartials + blockIdx.x] = c;
        }
    }
}

// Two smoothing steps in one pass, for the coarse levels whose kernels are launch-bound:
// out = S_w2(S_w1(x)), S_w(x) = x + w (b - A x) / diag on the interior, 0 on the faces,
// with k_mg_stencil's expressions.  One node per thread: it forms S_w1(x) at its node and
// at its six neighbours (the interior ones; faces hold zero) from the 25 x values within
// two axis steps and the spacings of the three columns / rows / planes around it (read
// through the caches), then S_w2 at its node.
__device__ __forceinline__ float mg_smooth_value(float hxl, float hxu, float hyl, float hyu, float hzl, float hzu,
                                                 float ixm, float ixq, float iym, float iyq, float izm, float izq,
                                                 float w, float bk, float xc, float xw, float xe, float xs,
                                                 float xn, float xm, float xp) {
    const float cx = 2.f * hxl + hxu, dx = hxl + 2.f * hxu;
    const float szz = hxl * (2.f * hyl + hyu) + hxu * (hyl + 2.f * hyu);
    const float sx = hyl * (2.f * hzl + hzu) + hyu * (hzl + 2.f * hzu);
    const float syy = hzl * cx + hzu * dx;
    const float diag = sx * (ixm + ixq) + syy * (iym + iyq) + szz * (izm + izq);
    const float ax = diag * xc - sx * (ixm * xw + ixq * xe) - syy * (iym * xs + iyq * xn) - szz * (izm * xm + izq * xp);
    return xc + w * (bk - ax) / diag;
}
__global__ __launch_bounds__(256) void k_mg_smooth2(PoissonGrid<float> P, const float *__restrict__ b,
                                                    const float *__restrict__ x, float *__restrict__ out, float w1,
                                                    float w2, const int *skip) {
    const int nx = P.nx, ny = P.ny, nz = P.nz;
    const int64_t n = (int64_t)nx * ny * nz;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    if (skip && *skip) return;
    const uint32_t tu = (uint32_t)t, r = tu / (uint32_t)nx, kk = r / (uint32_t)ny;
    const int i = (int)(tu - r * (uint32_t)nx), j = (int)(r - kk * (uint32_t)ny), k = (int)kk;
    if (i == 0 || j == 0 || k == 0 || i == nx - 1 || j == ny - 1 || k == nz - 1) {
        out[t] = 0.f;
        return;
    }
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    // spacings of index - 1 .. index + 2 per axis (n + 1 entries each; 1 <= i <= nx - 2)
    float hx[4], ix[4], hy[4], iy[4], hz[4], iz[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        hx[q] = P.hxp[i - 1 + q];
        ix[q] = P.ixp[i - 1 + q];
        hy[q] = P.hyp[j - 1 + q];
        iy[q] = P.iyp[j - 1 + q];
        hz[q] = P.hzp[k - 1 + q];
        iz[q] = P.izp[k - 1 + q];
    }
    auto X = [&](int di, int dj, int dk) { return x[t + di + sy * dj + sz * dk]; };
    auto B = [&](int di, int dj, int dk) { return b[t + di + sy * dj + sz * dk]; };
    // step 1 at (i + di, j + dj, k + dk), one offset non-zero at most; 0 on the faces
    auto s1 = [&](int di, int dj, int dk) -> float {
        const int ii = i + di, jj = j + dj, kz = k + dk;
        if (ii == 0 || jj == 0 || kz == 0 || ii == nx - 1 || jj == ny - 1 || kz == nz - 1) return 0.f;
        const int a = 1 + di, c = 1 + dj, e = 1 + dk;
        return mg_smooth_value(hx[a], hx[a + 1], hy[c], hy[c + 1], hz[e], hz[e + 1], ix[a], ix[a + 1], iy[c],
                               iy[c + 1], iz[e], iz[e + 1], w1, B(di, dj, dk), X(di, dj, dk), X(di - 1, dj, dk),
                               X(di + 1, dj, dk), X(di, dj - 1, dk), X(di, dj + 1, dk), X(di, dj, dk - 1),
                               X(di, dj, dk + 1));
    };
    const float yc = s1(0, 0, 0), yw = s1(-1, 0, 0), ye = s1(1, 0, 0), ys = s1(0, -1, 0), yn = s1(0, 1, 0),
                ym = s1(0, 0, -1), yp = s1(0, 0, 1);
    out[t] = mg_smooth_value(hx[1], hx[2], hy[1], hy[2], hz[1], hz[2], ix[1], ix[2], iy[1], iy[2], iz[1], iz[2], w2,
                             B(0, 0, 0), yc, yw, ye, ys, yn, ym, yp);
}

// b_c = R r_f (transpose of the per-axis linear interpolation); zero on the coarse faces.
// Per coarse index I of an axis, up to 3 (fine index, weight) pairs at [3 I + q] (unused
// pairs: index 0, weight 0).  Also x1c = invw_c b_c, the coarse level's first smoothing
// step from zero (invw_c: the level's first Chebyshev weight over the diagonal, 0 on the
// faces).  Tiles of 64 x 4 coarse (x, y) columns, each thread marching a z chunk; the
// z and y pairs are uniform over a wave, the x pairs per lane.
// Each wave restricts 64 consecutive coarse nodes (I) of one coarse row (J, K): it first
// combines the 3 x 3 fine rows (z pairs x y pairs) with their weights over the fine span
// the 64 nodes use (at most 131 nodes from the first coarse node's first fine index: a
// coarse interval covers at most two fine ones), each lane loading 2 (+1) consecutive fine
// nodes per row, so every load is contiguous; the x pairs are then read from that
// combined row in LDS.  Loads are clamped into the fine row and LDS reads into the span,
// so every access stays in bounds; unused pairs (weight 0) add exactly zero.
constexpr int kRSpan = kMgRestrictSpan;
__global__ __launch_bounds__(256) void k_mg_restrict(MgGrid G, int nxc, int nyc, int nzc, int nxf, int nyf,
                                                     const int *__restrict__ rxi, const float *__restrict__ rxw,
                                                     const int *__restrict__ ryi, const float *__restrict__ ryw,
                                                     const int *__restrict__ rzi, const float *__restrict__ rzw,
                                                     const float *__restrict__ rf, float *__restrict__ bc,
                                                     const float *__restrict__ invwc, float *__restrict__ x1c,
                                                     const int *skip) {
    __shared__ float row[kMgTY][kRSpan + 4];
    const int sk = *(skip ? skip : &kNoSkip);   // tested once the transfer lists' loads are issued
    const Tile T = mg_tile(G);
    const int lane = threadIdx.x % kMgTX, w = threadIdx.x / kMgTX;
    const int I = T.bx * kMgTX + lane, J = T.by * kMgTY + w;
    const int Ic = min(I, nxc - 1), Jc = min(J, nyc - 1);   // list indices (clamped)
    const bool act = I < nxc && J < nyc;
    const bool edge = I == 0 || J == 0 || I >= nxc - 1 || J >= nyc - 1;
    const int chunk = (nzc + G.z - 1) / G.z;
    const int k0 = T.bz * chunk, k1 = min(nzc, k0 + chunk);
    const int64_t sf = (int64_t)nxf * nyf, sc = (int64_t)nxc * nyc;
    int base = rxi[3 * (T.bx * kMgTX)];   // the wave's first coarse node's first fine index
    int xi[3], yo[3];
    float xw[3], yw[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        xi[q] = rxi[3 * Ic + q];
        xw[q] = rxw[3 * Ic + q];
        yo[q] = nxf * ryi[3 * Jc + q];
        yw[q] = ryw[3 * Jc + q];
    }
    // the z pairs of plane K (loaded one plane ahead)
    int zi[3] = {0, 0, 0};
    float zw[3] = {0.f, 0.f, 0.f};
    if (k0 < k1) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            zi[c] = rzi[3 * k0 + c];
            zw[c] = rzw[3 * k0 + c];
        }
    }
    // the list loads complete before the flag's test (rather than being moved after it)
    asm volatile("" : "+v"(xi[0]), "+v"(xi[1]), "+v"(xi[2]), "+v"(xw[0]), "+v"(xw[1]), "+v"(xw[2]), "+v"(base));
    asm volatile("" : "+v"(yo[0]), "+v"(yo[1]), "+v"(yo[2]), "+v"(yw[0]), "+v"(yw[1]), "+v"(yw[2]));
    asm volatile("" : "+v"(zi[0]), "+v"(zi[1]), "+v"(zi[2]), "+v"(zw[0]), "+v"(zw[1]), "+v"(zw[2]));
    if (sk) return;   // the flag is the same for every block: no barrier is left waiting
    int xo[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) xo[q] = min(max(xi[q] - base, 0), kRSpan - 1);
    // this lane's fine nodes of the span: base + lane, base + 64 + lane, and for lanes
    // below kRSpan - 128 base + 128 + lane (clamped into the row)
    const int f0 = min(base + lane, nxf - 1), f1 = min(base + kMgTX + lane, nxf - 1);
    const int f2 = min(base + 2 * kMgTX + min(lane, kRSpan - 2 * kMgTX - 1), nxf - 1);
    for (int K = k0; K < k1; ++K) {
        const int64_t idx = Ic + (int64_t)nxc * Jc + sc * K;
        const float iwc = invwc[idx];
        int zin[3] = {0, 0, 0};
        float zwn[3] = {0.f, 0.f, 0.f};
        if (K + 1 < k1) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                zin[c] = rzi[3 * (K + 1) + c];
                zwn[c] = rzw[3 * (K + 1) + c];
            }
        }
        float v[3][3][3];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int bq = 0; bq < 3; ++bq) {
                const float *py = rf + sf * zi[c] + yo[bq];
                v[c][bq][0] = py[f0];
                v[c][bq][1] = py[f1];
                v[c][bq][2] = py[f2];
            }
        float g[3];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            float sz = 0.f;
#pragma unroll
            for (int c = 0; c < 3; ++c) sz += zw[c] * (yw[0] * v[c][0][m] + yw[1] * v[c][1][m] + yw[2] * v[c][2][m]);
            g[m] = sz;
        }
        row[w][lane] = g[0];
        row[w][kMgTX + lane] = g[1];
        if (lane < kRSpan - 2 * kMgTX) row[w][2 * kMgTX + lane] = g[2];
        __syncthreads();
        float s = 0.f;
        if (!(edge || K == 0 || K == nzc - 1))
            s = xw[0] * row[w][xo[0]] + xw[1] * row[w][xo[1]] + xw[2] * row[w][xo[2]];
        __syncthreads();   // the rows are rewritten for the next plane
        if (act) {
            bc[idx] = s;
            x1c[idx] = s * iwc;
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            zi[c] = zin[c];
            zw[c] = zwn[c];
        }
    }
}

// x_f += P x_c on the interior: per fine index i of an axis, coarse c0[i] with weight
// 1 - t[i] and c0[i] + 1 with weight t[i] (t = 0 on coarse nodes; an interior fine node
// has c0 + 1 inside the coarse axis).  Fine faces stay zero.  Strips of 256 consecutive
// fine plane nodes marching z chunks.
__global__ __launch_bounds__(256) void k_mg_prolong(MgGrid G, int nxf, int nyf, int nzf, int nxc, int nyc,
                                                    const int *__restrict__ pxc, const float *__restrict__ pxt,
                                                    const int *__restrict__ pyc, const float *__restrict__ pyt,
                                                    const int *__restrict__ pzc, const float *__restrict__ pzt,
                                                    const float *__restrict__ xc, float *__restrict__ xf,
                                                    const int *skip) {
    const int sk = *(skip ? skip : &kNoSkip);   // tested once the transfer lists' loads are issued
    const Tile T = mg_tile(G);
    const int e = T.bx * 256 + threadIdx.x;   // strips of 256 consecutive plane nodes
    const int j = e / nxf, i = e - j * nxf;
    const int chunk = (nzf + G.z - 1) / G.z;
    const int k0 = max(1, T.bz * chunk), k1 = min(nzf - 1, T.bz * chunk + chunk);
    if (i <= 0 || j <= 0 || i >= nxf - 1 || j >= nyf - 1) return;
    const int64_t sf = (int64_t)nxf * nyf, sc = (int64_t)nxc * nyc;
    float tx = pxt[i], ty = pyt[j];
    int cx = pxc[i], cy = pyc[j];
    float tzk = 0.f;   // the z pair of plane k, loaded one plane ahead
    int czk = 0;
    if (k0 < k1) {
        tzk = pzt[k0];
        czk = pzc[k0];
    }
    // loaded before the flag's test
    asm volatile("" : "+v"(tx), "+v"(ty), "+v"(cx), "+v"(cy), "+v"(tzk), "+v"(czk));
    if (sk) return;
    const int64_t oxy = cx + (int64_t)nxc * cy;
    for (int k = k0; k < k1; ++k) {
        const float tz = tzk;
        const float *c0 = xc + oxy + sc * czk;
        if (k + 1 < k1) {
            tzk = pzt[k + 1];
            czk = pzc[k + 1];
        }
        const float *c1 = c0 + sc;   // inside the coarse axis for an interior fine node
        // both coarse planes' loads at once; with tz = 0 the second adds exactly zero
        const float a0 = (1.f - ty) * ((1.f - tx) * c0[0] + tx * c0[1]) + ty * ((1.f - tx) * c0[nxc] + tx * c0[nxc + 1]);
        const float a1 =
            (1.f - ty) * ((1.f - tx) * c1[0] + tx * c1[1]) + ty * ((1.f - tx) * c1[nxc] + tx * c1[nxc + 1]);
        float s = (1.f - tz) * a0;
        s += tz * a1;
        const int64_t idx = i + (int64_t)nxf * j + sf * k;
        xf[idx] += s;
    }
}

// ---------------------------------------------------------------- coarse tail in LDS
// The last levels of the V-cycle in one workgroup: each level's vectors live in LDS
// (x and y ping-pong, the right-hand side for all but the first tail level, whose b is
// read from global memory), with the level's stencil factors (the three edge-sum planes
// and the 1/(6h) rows) and the transfer lists copied to LDS once; every smoothing step,
// residual and transfer is one pass of the workgroup over the level's nodes, separated
// by barriers.  Same arithmetic as the per-level kernels above.
constexpr int kTailThreads = 1024;

// The tail's dynamic LDS; every tail array is an offset into it (so that the accesses
// are LDS instructions; integers are stored bit-cast), 16-byte aligned for the float4
// copy of the tables.
extern __shared__ __align__(16) float tail_lds[];

// A level's stencil factors (offsets): SX[j + ny k] = edge_sum(hy, hz; j, k),
// SY[k + nz i] = edge_sum(hz, hx; k, i), SZ[i + nx j] = edge_sum(hx, hy; i, j),
// IX[0..nx] = 1/(6 hx) (0 on the pads), likewise IY, IZ.
struct TailLevel {
    int nx, ny, nz;
    int SX, SY, SZ, IX, IY, IZ;
};
// Transfer lists of a fine tail level (offsets; layout of MgLevelDev's lists)
struct TailTransfer {
    int rxi, ryi, rzi, rxw, ryw, rzw, pxc, pyc, pzc, pxt, pyt, pzt;
};

__host__ __device__ inline int tail_coef_floats(int nx, int ny, int nz) {
    return ny * nz + nz * nx + nx * ny + (nx + 1) + (ny + 1) + (nz + 1);
}
__host__ __device__ inline int tail_transfer_words(int nxf, int nyf, int nzf, int nxc, int nyc, int nzc) {
    return 6 * (nxc + nyc + nzc) + 2 * (nxf + nyf + nzf);
}

// A thread's first node t = threadIdx.x on an nx x ny x nz level and the (i, j, k) step of
// t += kTailThreads: computed once per level (they need divisions), then every pass over
// the level walks its nodes from them without divisions (TailIter).
struct TailPos {
    int i, j, k, si, sj, sk;
};
__device__ TailPos tail_pos(int nx, int ny) {
    TailPos p;
    const int r = threadIdx.x / nx;
    p.i = threadIdx.x - r * nx;
    p.j = r % ny;
    p.k = r / ny;
    const int s = kTailThreads / nx;
    p.si = kTailThreads - s * nx;
    p.sj = s % ny;
    p.sk = s / ny;
    return p;
}
// Node coordinates of t = threadIdx.x, threadIdx.x + kTailThreads, ... on an nx x ny x nz
// level, advanced without divisions.
struct TailIter {
    int i, j, k, si, sj, sk, nx, ny;
    __device__ TailIter(const TailPos &p, int nx_, int ny_)
        : i(p.i), j(p.j), k(p.k), si(p.si), sj(p.sj), sk(p.sk), nx(nx_), ny(ny_) {}
    __device__ void next() {
        i += si;
        j += sj;
        k += sk;
        if (i >= nx) {
            i -= nx;
            ++j;
        }
        if (j >= ny) {
            j -= ny;
            ++k;
        }
    }
};

__device__ __forceinline__ float lf(int o) { return tail_lds[o]; }
__device__ __forceinline__ int li(int o) { return __float_as_int(tail_lds[o]); }

__device__ TailLevel tail_level(int base, int nx, int ny, int nz) {
    TailLevel t;
    t.nx = nx;
    t.ny = ny;
    t.nz = nz;
    t.SX = base;
    t.SY = t.SX + ny * nz;
    t.SZ = t.SY + nz * nx;
    t.IX = t.SZ + nx * ny;
    t.IY = t.IX + nx + 1;
    t.IZ = t.IY + ny + 1;
    return t;
}
__device__ TailTransfer tail_transfer(int base, int nxf, int nyf, int nzf, int nxc, int nyc, int nzc) {
    TailTransfer t;
    t.rxi = base;
    t.ryi = t.rxi + 3 * nxc;
    t.rzi = t.ryi + 3 * nyc;
    t.rxw = t.rzi + 3 * nzc;
    t.ryw = t.rxw + 3 * nxc;
    t.rzw = t.ryw + 3 * nyc;
    t.pxc = t.rzw + 3 * nzc;
    t.pyc = t.pxc + nxf;
    t.pzc = t.pyc + nyf;
    t.pxt = t.pzc + nzf;
    t.pyt = t.pxt + nxf;
    t.pzt = t.pyt + nyf;
    return t;
}

__device__ void tail_fill_level(const PoissonGrid<float> &P, const TailLevel &T) {
    const int nx = P.nx, ny = P.ny, nz = P.nz;
    for (int t = threadIdx.x; t < ny * nz; t += kTailThreads) tail_lds[T.SX + t] = edge_sum_mg(P.hyp, P.hzp, t % ny, t / ny);
    for (int t = threadIdx.x; t < nz * nx; t += kTailThreads) tail_lds[T.SY + t] = edge_sum_mg(P.hzp, P.hxp, t % nz, t / nz);
    for (int t = threadIdx.x; t < nx * ny; t += kTailThreads) tail_lds[T.SZ + t] = edge_sum_mg(P.hxp, P.hyp, t % nx, t / nx);
    for (int t = threadIdx.x; t <= nx; t += kTailThreads) tail_lds[T.IX + t] = P.ixp[t];
    for (int t = threadIdx.x; t <= ny; t += kTailThreads) tail_lds[T.IY + t] = P.iyp[t];
    for (int t = threadIdx.x; t <= nz; t += kTailThreads) tail_lds[T.IZ + t] = P.izp[t];
}
__device__ void tail_fill_pairs(int n, const int *idx, const float *w, int oi, int ow) {
    for (int t = threadIdx.x; t < n; t += kTailThreads) {
        tail_lds[oi + t] = __int_as_float(idx[t]);
        tail_lds[ow + t] = w[t];
    }
}
__device__ void tail_fill_transfer(const MgLevelDev &f, const PoissonGrid<float> &C, const TailTransfer &T) {
    tail_fill_pairs(3 * C.nx, f.rxi, f.rxw, T.rxi, T.rxw);
    tail_fill_pairs(3 * C.ny, f.ryi, f.ryw, T.ryi, T.ryw);
    tail_fill_pairs(3 * C.nz, f.rzi, f.rzw, T.rzi, T.rzw);
    tail_fill_pairs(f.P.nx, f.pxc, f.pxt, T.pxc, T.pxt);
    tail_fill_pairs(f.P.ny, f.pyc, f.pyt, T.pyc, T.pyt);
    tail_fill_pairs(f.P.nz, f.pzc, f.pzt, T.pzc, T.pzt);
}

// The stencil passes walk z columns: thread u takes column c = u mod (nx ny) and the
// u div (nx ny)-th of the kTailThreads / (nx ny) equal k ranges (empty past nz), so each
// node is one thread's once; the column's x and y factors and its face test are loaded
// once per pass.  Needs nx ny <= kTailThreads (mg_tail_fits).
struct TailCol {
    int i, j, k0, k1;
};
__device__ TailCol tail_col(int nx, int ny, int nz) {
    const int ncol = nx * ny, ch = kTailThreads / ncol, kc = (nz + ch - 1) / ch;
    const int c = threadIdx.x % ncol, chunk = threadIdx.x / ncol;
    TailCol p;
    p.i = c % nx;
    p.j = c / nx;
    p.k0 = min(nz, chunk * kc);   // chunk >= ch gives k0 = nz: idle
    p.k1 = min(nz, p.k0 + kc);
    return p;
}

// out = op(b, x) on level P; b at LDS offset ob, or the global bg when ob < 0
__device__ __forceinline__ void tail_stencil(const TailLevel &P, const TailCol &col, int ob, const float *bg, int ox,
                                             int oo, float w, int op) {
    const int nx = P.nx, ny = P.ny, nz = P.nz, nxy = nx * ny;
    const int i = col.i, j = col.j;
    const bool in_ij = i > 0 && j > 0 && i < nx - 1 && j < ny - 1;
    const float szz = lf(P.SZ + i + nx * j);
    const float ixm = lf(P.IX + i), ixq = lf(P.IX + i + 1), iym = lf(P.IY + j), iyq = lf(P.IY + j + 1);
    int t = i + nx * j + nxy * col.k0;
    for (int k = col.k0; k < col.k1; ++k, t += nxy) {
        float o = 0.f;
        if (in_ij && k > 0 && k < nz - 1) {
            const float sx = lf(P.SX + j + ny * k), syy = lf(P.SY + k + nz * i);
            const float izm = lf(P.IZ + k), izq = lf(P.IZ + k + 1);
            const float diag = sx * (ixm + ixq) + syy * (iym + iyq) + szz * (izm + izq);
            const float bi = ob < 0 ? bg[t] : lf(ob + t);
            if (op == OP_FIRST) {
                o = w * bi / diag;
            } else {
                const int x = ox + t;
                const float ax = diag * lf(x) - sx * (ixm * lf(x - 1) + ixq * lf(x + 1)) -
                                 syy * (iym * lf(x - nx) + iyq * lf(x + nx)) -
                                 szz * (izm * lf(x - nxy) + izq * lf(x + nxy));
                o = op == OP_RESID ? bi - ax : lf(x) + w * (bi - ax) / diag;
            }
        }
        tail_lds[oo + t] = o;
    }
    __syncthreads();
}

// bc = R rf, zero on the coarse faces (fine level F with its lists T, coarse level C)
__device__ __forceinline__ void tail_restrict(const TailLevel &F, const TailTransfer &T, const TailLevel &C,
                                              const TailPos &cpos, int orf, int obc) {
    const int nxc = C.nx, nyc = C.ny, nzc = C.nz, nxyc = nxc * nyc, nc = nxyc * nzc;
    const int nxf = F.nx, nxyf = F.nx * F.ny;
    TailIter it(cpos, nxc, nyc);
    for (int t = threadIdx.x; t < nc; t += kTailThreads, it.next()) {
        const int I = it.i, J = it.j, K = it.k;
        float s = 0.f;
        if (!(I == 0 || J == 0 || K == 0 || I == nxc - 1 || J == nyc - 1 || K == nzc - 1)) {
            const int x0 = li(T.rxi + 3 * I), x1 = li(T.rxi + 3 * I + 1), x2 = li(T.rxi + 3 * I + 2);
            const float w0 = lf(T.rxw + 3 * I), w1 = lf(T.rxw + 3 * I + 1), w2 = lf(T.rxw + 3 * I + 2);
            for (int c = 0; c < 3; ++c) {
                const float wz = lf(T.rzw + 3 * K + c);
                if (wz == 0.f) continue;
                const int pz = orf + nxyf * li(T.rzi + 3 * K + c);
                float sz = 0.f;
                for (int q = 0; q < 3; ++q) {
                    const float wy = lf(T.ryw + 3 * J + q);
                    if (wy == 0.f) continue;
                    const int py = pz + nxf * li(T.ryi + 3 * J + q);
                    sz += wy * (w0 * lf(py + x0) + w1 * lf(py + x1) + w2 * lf(py + x2));
                }
                s += wz * sz;
            }
        }
        tail_lds[obc + t] = s;
    }
    __syncthreads();
}

// xf += P xc on the interior of the fine level F
__device__ __forceinline__ void tail_prolong(const TailLevel &F, const TailPos &fpos, const TailTransfer &T,
                                             const TailLevel &C, int oxc, int oxf) {
    const int nx = F.nx, ny = F.ny, nz = F.nz, nxy = nx * ny, n = nxy * nz;
    const int nxc = C.nx, nxyc = C.nx * C.ny;
    TailIter it(fpos, nx, ny);
    for (int t = threadIdx.x; t < n; t += kTailThreads, it.next()) {
        const int i = it.i, j = it.j, k = it.k;
        if (i == 0 || j == 0 || k == 0 || i == nx - 1 || j == ny - 1 || k == nz - 1) continue;
        const float tx = lf(T.pxt + i), ty = lf(T.pyt + j), tz = lf(T.pzt + k);
        const int c0 = oxc + li(T.pxc + i) + nxc * li(T.pyc + j) + nxyc * li(T.pzc + k), c1 = c0 + nxyc;
        const float a0 = (1.f - ty) * ((1.f - tx) * lf(c0) + tx * lf(c0 + 1)) +
                         ty * ((1.f - tx) * lf(c0 + nxc) + tx * lf(c0 + nxc + 1));
        const float a1 = (1.f - ty) * ((1.f - tx) * lf(c1) + tx * lf(c1 + 1)) +
                         ty * ((1.f - tx) * lf(c1 + nxc) + tx * lf(c1 + nxc + 1));
        tail_lds[oxf + t] += (1.f - tz) * a0 + tz * a1;
    }
    __syncthreads();
}

// The tail's stencil factors and transfer lists (its LDS words from coef_off[0] on) for NT
// levels, formed once into img: each tail launch then copies them in one pass instead of
// gathering them from the level arrays (one memory round trip per table).
template <int NT>
__global__ __launch_bounds__(kTailThreads) void k_mg_tail_image(MgTailArgs a, float *img) {
#pragma unroll
    for (int q = 0; q < NT; ++q) {
        const PoissonGrid<float> &P = a.lev[q].P;
        tail_fill_level(P, tail_level(a.coef_off[q], P.nx, P.ny, P.nz));
        if (q + 1 < NT) {
            const PoissonGrid<float> &C = a.lev[q + 1].P;
            tail_fill_transfer(a.lev[q], C, tail_transfer(a.tr_off[q], P.nx, P.ny, P.nz, C.nx, C.ny, C.nz));
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < a.img_n; t += kTailThreads) img[t] = tail_lds[a.coef_off[0] + t];
}

// NT tail levels (compile-time, so that the per-level state stays in registers)
template <int NT>
__global__ __launch_bounds__(kTailThreads) void k_mg_tail(MgTailArgs a, const int *skip) {
    // the image's first float4 per thread is loaded before the skip test (index clamped
    // into the image, so the load is unconditional); coef_off[0] and img_n are multiples
    // of 4 (mg_tail_layout)
    const int n4 = a.img_n >> 2;
    const float4 *src = reinterpret_cast<const float4 *>(a.img);
    float4 *dst = reinterpret_cast<float4 *>(tail_lds + a.coef_off[0]);
    float4 v = src[min((int)threadIdx.x, n4 - 1)];
    const int sk = *(skip ? skip : &kNoSkip);
    asm volatile("" : "+v"(v.x), "+v"(v.y), "+v"(v.z), "+v"(v.w));
    if (sk) return;
    if (threadIdx.x == 0 && a.count) *a.count += 1;
    if ((int)threadIdx.x < n4) dst[threadIdx.x] = v;
    for (int t = threadIdx.x + kTailThreads; t < n4; t += kTailThreads) dst[t] = src[t];
    int cur[NT], oth[NT], ob[NT];
    TailLevel TL[NT];
    TailTransfer TT[NT];
    TailPos pos[NT];
    TailCol col[NT];
#pragma unroll
    for (int q = 0; q < NT; ++q) {
        const PoissonGrid<float> &P = a.lev[q].P;
        pos[q] = tail_pos(P.nx, P.ny);
        col[q] = tail_col(P.nx, P.ny, P.nz);
        cur[q] = a.off[q];
        oth[q] = a.off[q] + a.n[q];
        ob[q] = q == 0 ? -1 : a.off[q] + 2 * a.n[q];
        TL[q] = tail_level(a.coef_off[q], P.nx, P.ny, P.nz);
        if (q + 1 < NT) {
            const PoissonGrid<float> &C = a.lev[q + 1].P;
            TT[q] = tail_transfer(a.tr_off[q], P.nx, P.ny, P.nz, C.nx, C.ny, C.nz);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q + 1 < NT; ++q) {
        tail_stencil(TL[q], col[q], ob[q], a.b0, 0, cur[q], a.om[0], OP_FIRST);
        for (int k = 1; k < a.s; ++k) {
            tail_stencil(TL[q], col[q], ob[q], a.b0, cur[q], oth[q], a.om[k], OP_SMOOTH);
            const int t = cur[q]; cur[q] = oth[q]; oth[q] = t;
        }
        tail_stencil(TL[q], col[q], ob[q], a.b0, cur[q], oth[q], 0.f, OP_RESID);
        tail_restrict(TL[q], TT[q], TL[q + 1], pos[q + 1], oth[q], ob[q + 1]);
    }
    {
        constexpr int q = NT - 1;
        tail_stencil(TL[q], col[q], ob[q], a.b0, 0, cur[q], a.omc[0], OP_FIRST);
        for (int k = 1; k < a.cd; ++k) {
            tail_stencil(TL[q], col[q], ob[q], a.b0, cur[q], oth[q], a.omc[k], OP_SMOOTH);
            const int t = cur[q]; cur[q] = oth[q]; oth[q] = t;
        }
    }
#pragma unroll
    for (int q = NT - 2; q >= 0; --q) {
        tail_prolong(TL[q], pos[q], TT[q], TL[q + 1], cur[q + 1], cur[q]);
        for (int k = a.s - 1; k >= 0; --k) {
            tail_stencil(TL[q], col[q], ob[q], a.b0, cur[q], oth[q], a.om[k], OP_SMOOTH);
            const int t = cur[q]; cur[q] = oth[q]; oth[q] = t;
        }
    }
    for (int t = threadIdx.x; t < a.n[0]; t += kTailThreads) a.x0[t] = tail_lds[cur[0] + t];
}

// At most kMaxPartials blocks (the fused dot writes one partial per block); the setup
// checks that one z chunk of tiles fits.
MgGrid tile_grid(const PoissonGrid<float> &P) {
    const int gx = (P.nx + kMgTX - 1) / kMgTX, gy = (P.ny + kMgTY - 1) / kMgTY;
    const int64_t per_plane = (int64_t)gx * gy;
    const int gz = (int)std::min<int64_t>(P.nz, std::max<int64_t>(1, kMaxPartials / per_plane));
    return MgGrid{gx, gy, gz};
}
// The stencil kernels march strips of 256 consecutive nodes of a plane (long contiguous
// runs per block and plane: 3% faster than 64 x 4 tiles on the 154^3 grid).
MgGrid stencil_grid(const PoissonGrid<float> &P) {
    const int gx = (P.nx * P.ny + 255) / 256, gy = 1;
    const int64_t per_plane = (int64_t)gx * gy;
    const int gz = (int)std::min<int64_t>(P.nz, std::max<int64_t>(1, kMaxPartials / per_plane));
    return MgGrid{gx, gy, gz};
}
inline dim3 blocks(const MgGrid &g) { return dim3(g.x * g.y * g.z); }


}  // namespace

void launch_mg_first(const Launch &L, const PoissonGrid<float> &P, const float *b, float *out, float w,
                     const int *skip) {
    const MgGrid g = stencil_grid(P);
    hipLaunchKernelGGL((k_mg_stencil<OP_FIRST, false>), blocks(g), dim3(256), 0, L.s, P, g, b,
                       (const float *)nullptr, out, w, skip, (double *)nullptr);
}
int launch_mg_smooth(const Launch &L, const PoissonGrid<float> &P, const float *b, const float *x, float *out,
                     float w, const int *skip, double *dot_parts) {
    const MgGrid g = stencil_grid(P);
    if (dot_parts)
        hipLaunchKernelGGL((k_mg_stencil<OP_SMOOTH, true>), blocks(g), dim3(256), 0, L.s, P, g, b, x, out, w, skip,
                           dot_parts);
    else if (L.t0)   // timed: events recorded on the dispatch itself
        hipExtLaunchKernelGGL((k_mg_stencil<OP_SMOOTH, false>), blocks(g), dim3(256), 0, L.s, L.t0, L.t1, 0, P, g,
                              b, x, out, w, skip, (double *)nullptr);
    else
        hipLaunchKernelGGL((k_mg_stencil<OP_SMOOTH, false>), blocks(g), dim3(256), 0, L.s, P, g, b, x, out, w, skip,
                           (double *)nullptr);
    return g.x * g.y * g.z;
}
void launch_mg_smooth2(const Launch &L, const PoissonGrid<float> &P, const float *b, const float *x, float *out,
                       float w1, float w2, const int *skip) {
    const int64_t n = (int64_t)P.nx * P.ny * P.nz;
    hipLaunchKernelGGL(k_mg_smooth2, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, L.s, P, b, x, out, w1, w2,
                       skip);
}
void launch_mg_resid(const Launch &L, const PoissonGrid<float> &P, const float *b, const float *x, float *out,
                     const int *skip) {
    const MgGrid g = stencil_grid(P);
    hipLaunchKernelGGL((k_mg_stencil<OP_RESID, false>), blocks(g), dim3(256), 0, L.s, P, g, b, x, out, 0.f, skip,
                       (double *)nullptr);
}
int mg_tail_layout(MgTailArgs &a) {
    int f = 0;
    for (int q = 0; q < a.nt; ++q) {
        a.off[q] = f;
        f += (q == 0 ? 2 : 3) * a.n[q];
    }
    f = (f + 3) & ~3;   // the tables (the image) start on a float4
    for (int q = 0; q < a.nt; ++q) {
        const PoissonGrid<float> &P = a.lev[q].P;
        a.coef_off[q] = f;
        f += tail_coef_floats(P.nx, P.ny, P.nz);
        if (q + 1 < a.nt) {
            const PoissonGrid<float> &C = a.lev[q + 1].P;
            a.tr_off[q] = f;
            f += tail_transfer_words(P.nx, P.ny, P.nz, C.nx, C.ny, C.nz);
        }
    }
    f = (f + 3) & ~3;
    a.img_n = f - a.coef_off[0];
    a.lds_bytes = f * (int)sizeof(float);
    return a.lds_bytes;
}
bool mg_tail_fits(const MgTailArgs &a) {
    for (int q = 0; q < a.nt; ++q)
        if (a.lev[q].P.nx * a.lev[q].P.ny > kTailThreads) return false;
    return a.lds_bytes <= kMgTailLdsBytes;
}
void launch_mg_tail_image(const Launch &L, const MgTailArgs &a, float *img) {
    switch (a.nt) {
        case 1: hipLaunchKernelGGL(k_mg_tail_image<1>, dim3(1), dim3(kTailThreads), a.lds_bytes, L.s, a, img); break;
        case 2: hipLaunchKernelGGL(k_mg_tail_image<2>, dim3(1), dim3(kTailThreads), a.lds_bytes, L.s, a, img); break;
        case 3: hipLaunchKernelGGL(k_mg_tail_image<3>, dim3(1), dim3(kTailThreads), a.lds_bytes, L.s, a, img); break;
        default: hipLaunchKernelGGL(k_mg_tail_image<4>, dim3(1), dim3(kTailThreads), a.lds_bytes, L.s, a, img); break;
    }
}
void launch_mg_tail(const Launch &L, const MgTailArgs &a, const int *skip) {
    switch (a.nt) {
        case 1: hipLaunchKernelGGL(k_mg_tail<1>, dim3(1), dim3(kTailThreads), a.lds_bytes, L.s, a, skip); break;
        case 2: hipLaunchKernelGGL(k_mg_tail<2>, dim3(1), dim3(kTailThreads), a.lds_bytes, L.s, a, skip); break;
        case 3: hipLaunchKernelGGL(k_mg_tail<3>, dim3(1), dim3(kTailThreads), a.lds_bytes, L.s, a, skip); break;
        default: hipLaunchKernelGGL(k_mg_tail<4>, dim3(1), dim3(kTailThreads), a.lds_bytes, L.s, a, skip); break;
    }
}
bool mg_stencil_fits(const PoissonGrid<float> &P) {
    const MgGrid g = stencil_grid(P);
    return (int64_t)g.x * g.y <= kMaxPartials;
}
void launch_mg_restrict(const Launch &L, const MgLevelDev &f, const MgLevelDev &c, const float *rf, float *bc,
                        float *x1c, const int *skip) {
    const MgGrid g = tile_grid(c.P);
    hipLaunchKernelGGL(k_mg_restrict, blocks(g), dim3(256), 0, L.s, g, c.P.nx, c.P.ny, c.P.nz, f.P.nx,
                       f.P.ny, f.rxi, f.rxw, f.ryi, f.ryw, f.rzi, f.rzw, rf, bc, c.invw, x1c, skip);
}
void launch_mg_prolong(const Launch &L, const MgLevelDev &f, const MgLevelDev &c, const float *xc, float *xf,
                       const int *skip) {
    const MgGrid g = stencil_grid(f.P);
    hipLaunchKernelGGL(k_mg_prolong, blocks(g), dim3(256), 0, L.s, g, f.P.nx, f.P.ny, f.P.nz, c.P.nx,
                       c.P.ny, f.pxc, f.pxt, f.pyc, f.pyt, f.pzc, f.pzt, xc, xf, skip);
}

}  // namespace mm
