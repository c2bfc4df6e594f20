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
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>

#include "common.h"

namespace mm {

namespace {

constexpr int kFB = 1024;  // threads per block of the streaming fused kernels (16 wavefronts)
// Tile rows, target block count and waves per SIMD of the structured KA (compile-time
// tunables; the defaults are the measured best on MI355X).
// (per precision: FP32 loads move half the bytes per request, so the FP32 kernel wants
// more, smaller tiles; swept on the 154^3 benchmark grid: FP32 64 x 4 tiles with 768
// blocks 15.3 us against 15.9 us for 64 x 8 / 512; FP64 64 x 8 with 256 blocks 24.2 us
// against 26.5 us)
#ifndef MM_KA_TY
#define MM_KA_TY 8
#endif
#ifndef MM_KA_BLOCKS
#define MM_KA_BLOCKS 256
#endif
#ifndef MM_KA_TY32
#define MM_KA_TY32 4
#endif
#ifndef MM_KA_TX
#define MM_KA_TX 64
#endif
#ifndef MM_KA_TX32
#define MM_KA_TX32 64
#endif
#ifndef MM_KA_BLOCKS32
#define MM_KA_BLOCKS32 768
#endif
#ifndef MM_KA_PF
#define MM_KA_PF 2
#endif
#ifndef MM_KA_WAVES
#define MM_KA_WAVES 0
#endif
#ifndef MM_KB_BLOCKS
#define MM_KB_BLOCKS 256
#endif
// rows of a KA tile on the structured grid (64 x TY nodes) and the target block count
template <typename T>
struct KaCfg {
    static constexpr int TX = MM_KA_TX, TY = MM_KA_TY, blocks = MM_KA_BLOCKS;
};
template <>
struct KaCfg<float> {
    static constexpr int TX = MM_KA_TX32, TY = MM_KA_TY32, blocks = MM_KA_BLOCKS32;
};

__device__ __forceinline__ double fwave_sum(double v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Sum of n partials (fixed order: per-thread strided sums, then a fixed tree).
template <int NT>
__device__ double block_sum_parts(const double *__restrict__ parts, int n, double *sh) {
    double s = 0.0;
    for (int j = threadIdx.x; j < n; j += NT) s += parts[j];
    s = fwave_sum(s);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) sh[w] = s;
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < NT / kWave; ++k) t += sh[k];
    return t;
}

// Block reduction of NV values into out[i * kMaxPartials + blockIdx.x].
template <int NV, int NT>
__device__ void fblock_partials(double (&v)[NV], double *sh, double *out) {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double s = fwave_sum(v[i]);
        __syncthreads();
        if (lane == 0) sh[w] = s;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int k = 0; k < NT / kWave; ++k) t += sh[k];
            out[i * kMaxPartials + blockIdx.x] = t;
        }
    }
}

// Scalars of iteration i from the partials of the previous KB (or the init / replaced
// residual).  Returns beta; sets *stop (block-uniform).  Every load is issued before the
// one barrier.  The stop flag may be written by block 0 of this same kernel while other
// blocks start (to the value they would compute themselves), so thread 0 reads it and
// broadcasts it through LDS.  Block 0 thread 0 publishes the scalars.  sh holds
// 3 * NT / kWave doubles.
template <int NT>
__device__ double ka_prologue(const PcgIterArgs &a, DevScalars *dsc, double *sh, int *flag, bool *stop) {
    constexpr int NW = NT / kWave;
    if (threadIdx.x == 0) *flag = dsc->pcg_done;
    const bool from_parts = !(a.mode == PCG_KEEP || a.mode == PCG_RESTART);   // FIRST, NORMAL, RESTART_PARTS
    const bool first = a.mode == PCG_FIRST;
    const double gprev = dsc->pcg_gam[(a.iter + 1) & 1];
    const double bound_prev = dsc->pcg_bound;
    const double ru_gamma = dsc->ru_gamma, ir_rr = dsc->ir_rr;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    if (from_parts) {
        for (int jj = threadIdx.x; jj < a.nparts_in; jj += NT) {
            s0 += a.parts_in[jj];
            s1 += a.parts_in[kMaxPartials + jj];
            if (first) s2 += a.parts_in[2 * kMaxPartials + jj];
        }
    }
    s0 = fwave_sum(s0);
    s1 = fwave_sum(s1);
    s2 = fwave_sum(s2);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    if (lane == 0) {
        sh[w] = s0;
        sh[NW + w] = s1;
        sh[2 * NW + w] = s2;
    }
    __syncthreads();
    if (*flag != 0) {
        *stop = true;
        return 0.0;
    }
    double gamma = 0.0, rr = 0.0, bb = 0.0;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        gamma += sh[k];
        rr += sh[NW + k];
        bb += sh[2 * NW + k];
    }
    double beta = 0.0;
    if (!from_parts) {
        gamma = ru_gamma;
        rr = ir_rr;
    }
    if (a.mode == PCG_NORMAL || a.mode == PCG_KEEP) beta = gamma / gprev;
    const double bound = first ? fmax(a.tol * a.tol * bb, a.atol * a.atol) : bound_prev;
    const bool done = !((rr > bound) && (a.iter < a.maxiter));
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        dsc->pcg_gam[a.iter & 1] = gamma;
        dsc->pcg_beta = beta;
        dsc->pcg_rr = rr;
        dsc->pcg_iter = a.iter;
        if (first) {
            dsc->pcg_bound = bound;
            dsc->pcg_bs = fmax(rr, 0.0);
        }
        dsc->pcg_done = done;
    }
    *stop = done;
    return beta;
}

// Logical tile of block b of nb: blocks are dispatched round-robin over the 8 XCDs
// (block b runs on XCD b % 8), so consecutive logical tiles are given to one XCD and
// the z-neighbouring chunks of one (x, y) column share that XCD's L2.  A bijection on
// [0, nb).
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    const int q = nb / 8, r = nb % 8, x = b % 8, l = b / 8;
    return x < r ? x * (q + 1) + l : r * (q + 1) + (x - r) * q + l;
}

// Uniform load through the scalar path (constant address space: the spacing arrays are
// read-only for the whole run).
template <typename T>
__device__ __forceinline__ T uload(const T *p, int k) {
    return ((const __attribute__((address_space(4))) T *)p)[k];
}

// Element access at a 32-bit byte offset (vectors of the structured path hold fewer
// than 2^29 entries, checked at set-up), so that addresses are a scalar base plus one
// 32-bit register.
template <typename T>
__device__ __forceinline__ T ld(const T *base, uint32_t off) {
    return *(const T *)((const char *)base + off);
}
template <typename T>
__device__ __forceinline__ void st(T *base, uint32_t off, T v) {
    *(T *)((char *)base + off) = v;
}

// Raw z / p_old pair of one node and plane.
template <typename T>
struct Raw {
    T z, p;
};

// KA on the structured grid: a block owns a 64 x TY tile of (x, y) and a chunk of at
// most kKaChunk z planes, marching upwards.  Per plane, p = z + beta p_old is formed
// once per node of the tile and of its one-node halo ring (south / north rows, west /
// east columns) into a double-buffered LDS plane; the in-plane neighbours come from
// LDS, the z neighbours from registers.  Raw z / p_old loads run MM_KA_PF planes ahead
// of their use (own node: plane k + 1 + R while plane k is computed; halo: plane k + R),
// in two register rings that alternate between rounds of R planes, so that no ring slot
// is copied (a copy waits for its load); the chunk's z spacings are staged in LDS.  The
// march has no control flow around its loads: where paths join the compiler waits for
// every outstanding memory operation, which had serialised the prefetch (19.9 -> 15.9 us
// on the 154^3 benchmark grid in FP32).
//
// Coefficients (as k_grid_poisson in kernels.hip): with A = 2 hz_l + hz_u and
// B = hz_l + 2 hz_u of plane k, the x-edge factor is sx = hy_l A + hy_u B, the y-edge
// factor syy = hx_l A + hx_u B, the z-edge factor sz = edge sum of (hx, hy); the
// off-diagonal entries are -sx / (6 h), ..., the diagonal their negated sum, and the
// entries towards Dirichlet nodes (first / last index on an axis) are dropped.
// Grid: 1D, gx * gy * gz blocks, logical tile = (z chunk fastest, then x tile, then y).
constexpr int kKaChunk = 256;
template <typename T, int TX, int TY, bool PZ>
__global__ __launch_bounds__(TX * TY)
#if MM_KA_WAVES > 0
__attribute__((amdgpu_waves_per_eu(MM_KA_WAVES)))
#endif
void k_ka_grid(PoissonGrid<T> P, PcgIterArgs a, int gx, int gz, const T *__restrict__ z,
               const T *__restrict__ pold, T *__restrict__ pnew, T *__restrict__ q, DevScalars *dsc) {
    static_assert(TX == 64 || TX == 32, "a wave holds one (TX = 64) or two (TX = 32) tile rows");
    static_assert(TY >= 3 && 2 * TY <= TX && TX * TY <= 1024 && (TX * TY) % kWave == 0,
                  "halo lanes need 3 <= TY <= TX / 2; blocks of whole waves, at most 1024 threads");
    constexpr int NT = TX * TY;
    __shared__ double sh[3 * NT / kWave];
    __shared__ T sp[2][TY + 2][TX + 2];
    __shared__ T szc[2][kKaChunk + 1];   // hz, 1 / (6 hz) of the chunk's planes k0 .. k1
    __shared__ int flag;
    const int nx = P.nx, ny = P.ny, nz = P.nz;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int cz = t % gz, bx = (t / gz) % gx, by = t / (gz * gx);
    const int tx = threadIdx.x % TX;
    // one row per wave (uniform) with 64-wide tiles, two rows per wave with 32-wide ones
    const int ty = TX == 64 ? __builtin_amdgcn_readfirstlane(threadIdx.x >> 6) : (int)(threadIdx.x / TX);
    const int i = bx * TX + tx, j = by * TY + ty;
    const int chunk = (nz + gz - 1) / gz;
    const int k0 = cz * chunk, k1 = min(nz, k0 + chunk);   // block-uniform; empty when k0 >= nz
    const uint32_t es = sizeof(T);
    const uint32_t sz = (uint32_t)nx * (uint32_t)ny * es;  // plane stride in bytes
    const bool in = i < nx && j < ny;
    const bool run = k0 < k1;
    constexpr bool pz = PZ;   // p = z (FIRST, RESTART, RESTART_PARTS), else z + beta p_old
    // halo node of this thread (at most one): south row (ty == 0), north row
    // (ty == TY - 1), west / east columns (ty == 1, lanes [0, TY) and [TX / 2, TX / 2 + TY))
    int hi = 0, hj = 0, hly = -1, hlx = 0;
    if (ty == 0) {
        hi = i; hj = by * TY - 1; hly = 0; hlx = tx + 1;
    } else if (ty == TY - 1) {
        hi = i; hj = by * TY + TY; hly = TY + 1; hlx = tx + 1;
    } else if (ty == 1 && tx < TY) {
        hi = bx * TX - 1; hj = by * TY + tx; hly = tx + 1; hlx = 0;
    } else if (ty == 1 && tx >= TX / 2 && tx < TX / 2 + TY) {
        hi = bx * TX + TX; hj = by * TY + tx - TX / 2; hly = tx - TX / 2 + 1; hlx = TX + 1;
    }
    const bool hin = hly >= 0 && hi >= 0 && hi < nx && hj >= 0 && hj < ny;
    const uint32_t ob = in ? (uint32_t)(i + nx * j) * es : 0u;
    const uint32_t hb_off = hin ? (uint32_t)(hi + nx * hj) * es : 0u;

    // Loads are unconditional and unmasked (control flow or a select at the load would make
    // the compiler wait for them where they are issued): an unused plane or halo slot
    // loads a valid address (plane clamped into the grid, node offset 0 outside the tile)
    // whose value is never used (own planes past the chunk or the grid are the upper
    // neighbours of Dirichlet planes or of no computed plane; halo values outside the
    // grid are replaced by 0 where they are written to LDS).
    auto raw = [&](uint32_t off, bool) -> Raw<T> {
        Raw<T> r;
        r.z = ld(z, off);
        r.p = pz ? (T)0 : ld(pold, off);
        return r;
    };
    auto plane = [&](int m) -> uint32_t { return (uint32_t)min(max(m, 0), nz - 1) * sz; };
    constexpr int R = MM_KA_PF;   // planes in flight per node (own and halo rings)
    // own ring slot u holds plane k0 + 1 + u (used as the upper neighbour at step k0 + u);
    // halo ring slot u holds halo plane k0 + u.  Own plane m is used iff m <= k1 and
    // m < nz; halo plane m iff m < k1.  Loads issued before the scalars are known.
    Raw<T> ra[R], rb[R], ha[R], hb[R];   // rounds alternate: A consumes ra / ha and loads rb / hb
    const Raw<T> om1 = raw(ob + plane(k0 - 1), true);
    const Raw<T> o0 = raw(ob + plane(k0), true);
#pragma unroll
    for (int u = 0; u < R; ++u) {
        ra[u] = raw(ob + plane(k0 + 1 + u), true);
        ha[u] = raw(hb_off + plane(k0 + u), true);
    }
    if (run) {
        for (int e = threadIdx.x; e <= k1 - k0; e += NT) {   // k1 - k0 <= kKaChunk
            szc[0][e] = P.hzp[k0 + e];
            szc[1][e] = P.izp[k0 + e];
        }
    }

    bool stop;
    const double bd = ka_prologue<NT>(a, dsc, sh, &flag, &stop);   // (its barrier publishes szc)
    if (stop) return;
    const T beta = (T)bd;
    auto comb = [&](const Raw<T> &r) -> T { return pz ? r.z : r.z + beta * r.p; };
    double dot = 0.0;
    if (run) {
        const bool edge_ij = !in || i == 0 || j == 0 || i == nx - 1 || j == ny - 1;
        // y factors: wave-uniform with 64-wide tiles (scalar path), x factors per lane
        const bool jin = j > 0 && j < ny - 1;
        auto yload = [&](const T *a, int k) -> T { return TX == 64 ? uload(a, k) : a[k]; };
        const T hyl = jin ? yload(P.hyp, j) : (T)0, hyu = jin ? yload(P.hyp, j + 1) : (T)0;
        const T b0 = jin ? yload(P.iyp, j) : (T)0, b1 = jin ? yload(P.iyp, j + 1) : (T)0;
        const T iys = b0 + b1, iym = j - 1 > 0 ? b0 : (T)0, iyp = j + 1 < ny - 1 ? b1 : (T)0;
        T hxl = 0, hxu = 0, ixs = 0, ixm = 0, ixp = 0;
        if (!edge_ij) {
            hxl = P.hxp[i]; hxu = P.hxp[i + 1];
            const T a0 = P.ixp[i], a1 = P.ixp[i + 1];
            ixs = a0 + a1;
            ixm = i - 1 > 0 ? a0 : (T)0; ixp = i + 1 < nx - 1 ? a1 : (T)0;
        }
        const T szij = (T)2 * hxl * hyl + hxl * hyu + hxu * hyl + (T)2 * hxu * hyu;
        T pm = comb(om1);   // plane k - 1 (0 below the grid or outside)
        T pc = comb(o0);    // plane k

        // One plane k: oc holds own plane k + 1, hc the halo of plane k; the loads of own
        // plane k + 1 + R and halo plane k + R go to on / hn (used R planes later).
        auto step = [&](int k, const Raw<T> &oc, Raw<T> &on, const Raw<T> &hc, Raw<T> &hn) {
            const T pp = comb(oc);
            const T hv = comb(hc);
            on = raw(ob + plane(k + 1 + R), true);
            hn = raw(hb_off + plane(k + R), true);
            T(*buf)[TX + 2] = sp[k & 1];
            buf[ty + 1][tx + 1] = pc;
            if (hly >= 0) buf[hly][hlx] = hin ? hv : (T)0;
            __syncthreads();
            if (in) {
                T out = pc;
                if (!(edge_ij || k == 0 || k == nz - 1)) {
                    const T xw = buf[ty + 1][tx], xe = buf[ty + 1][tx + 2];
                    const T xs = buf[ty][tx + 1], xn = buf[ty + 2][tx + 1];
                    const T hzl = szc[0][k - k0], hzu = szc[0][k - k0 + 1];
                    const T izl = szc[1][k - k0], izu = szc[1][k - k0 + 1];
                    const T A = (T)2 * hzl + hzu, B = hzl + (T)2 * hzu;
                    const T izm = k - 1 > 0 ? izl : (T)0, izq = k + 1 < nz - 1 ? izu : (T)0;
                    const T sx = hyl * A + hyu * B, syy = hxl * A + hxu * B;
                    const T diag = sx * ixs + syy * iys + szij * (izl + izu);
                    out = diag * pc - sx * (ixm * xw + ixp * xe) - syy * (iym * xs + iyp * xn) -
                          szij * (izm * pm + izq * pp);
                }
                const uint32_t off = ob + (uint32_t)k * sz;
                st(pnew, off, pc);
                st(q, off, out);
                dot += (double)out * (double)pc;
            }
            pm = pc;
            pc = pp;
        };
        // whole pairs of rounds without conditions (every slot stays in its registers),
        // then the remaining planes
        int kb = k0;
        for (; kb + 2 * R <= k1; kb += 2 * R) {
#pragma unroll
            for (int u = 0; u < R; ++u) step(kb + u, ra[u], rb[u], ha[u], hb[u]);
#pragma unroll
            for (int u = 0; u < R; ++u) step(kb + R + u, rb[u], ra[u], hb[u], ha[u]);
        }
#pragma unroll
        for (int u = 0; u < 2 * R - 1; ++u) {
            if (kb + u >= k1) break;
            if (u < R)
                step(kb + u, ra[u], rb[u], ha[u], hb[u]);
            else
                step(kb + u, rb[u - R], ra[u - R], hb[u - R], ha[u - R]);
        }
    }
    double v[1] = {dot};
    fblock_partials<1, NT>(v, sh, a.parts_out);
}

// KA on a SELL-64 matrix (general meshes): 16 slices per block.
template <typename T>
__global__ __launch_bounds__(kFB) void k_ka_sell(SellDev<T> A, PcgIterArgs a, const T *__restrict__ z,
                                                 const T *__restrict__ pold, T *__restrict__ pnew,
                                                 T *__restrict__ q, DevScalars *dsc) {
    __shared__ double sh[3 * kFB / kWave];
    __shared__ int flag;
    bool stop;
    const double bd = ka_prologue<kFB>(a, dsc, sh, &flag, &stop);
    if (stop) return;
    const T beta = (T)bd;
    const bool pz = a.mode == PCG_FIRST || a.mode == PCG_RESTART || a.mode == PCG_RESTART_PARTS;
    auto pv = [&](int c) -> T { return pz ? z[c] : z[c] + beta * pold[c]; };
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
    fblock_partials<1, kFB>(v, sh, a.parts_out);
}

// KB: alpha = gamma_i / (p.q); x += alpha p; r -= alpha q; WZ: z = r * invD and partials
// (r.z, r.r); !WZ: x and r only (another preconditioner forms z afterwards).
template <typename T, bool WZ>
__global__ __launch_bounds__(kFB) void k_kb(int64_t n, PcgIterArgs a, T *__restrict__ x,
                                            T *__restrict__ r, T *__restrict__ z,
                                            const T *__restrict__ p, const T *__restrict__ q,
                                            const T *__restrict__ invD, DevScalars *dsc) {
    __shared__ double sh[kFB / kWave];
    const int stop = dsc->pcg_done;   // tested after the partials' loads (block-uniform)
    const double pq = block_sum_parts<kFB>(a.parts_in, a.nparts_in, sh);
    if (stop) return;
    const T alpha = (T)(dsc->pcg_gam[a.iter & 1] / pq);
    double s[2] = {0.0, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)kFB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kFB) {
        x[i] = x[i] + alpha * p[i];
        const T ri = r[i] - alpha * q[i];
        r[i] = ri;
        if (WZ) {
            const T zi = ri * invD[i];
            z[i] = zi;
            s[0] += (double)ri * (double)zi;
            s[1] += (double)ri * (double)ri;
        }
    }
    if (WZ) fblock_partials<2, kFB>(s, sh, a.parts_out);
}

// Partials (r.z, r.r) for the next KA when z came from another preconditioner.
template <typename T>
__global__ __launch_bounds__(kFB) void k_dot2(int64_t n, const T *__restrict__ r, const T *__restrict__ z,
                                              double *parts_out, const DevScalars *dsc) {
    __shared__ double sh[kFB / kWave];
    if (dsc->pcg_done) return;
    double s[2] = {0.0, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)kFB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kFB) {
        const double ri = r[i];
        s[0] += ri * (double)z[i];
        s[1] += ri * ri;
    }
    fblock_partials<2, kFB>(s, sh, parts_out);
}

// Stop test between KB(i) and a preconditioner that runs after it (the multigrid
// V-cycle): r.r from KB's partials (slot 1); when it is at or below the stop bound, or
// iteration i + 1 would reach maxiter (KA(i + 1)'s own test), the iteration stops after
// i + 1 iterations, so that the preconditioner and every kernel queued after it are
// skipped.  One block.
constexpr int kCheckThreads = 256;
__global__ __launch_bounds__(kCheckThreads) void k_pcg_check(const double *__restrict__ parts, int nparts,
                                                            int iter, int maxiter, DevScalars *dsc) {
    __shared__ double sh[kCheckThreads / kWave];
    const int stop = dsc->pcg_done;   // tested after the partials' loads
    const double rr = block_sum_parts<kCheckThreads>(parts + kMaxPartials, nparts, sh);
    if (stop) return;
    if (threadIdx.x == 0 && !(rr > dsc->pcg_bound && iter + 1 < maxiter)) {
        dsc->pcg_rr = rr;
        dsc->pcg_iter = iter + 1;
        dsc->pcg_done = 1;
    }
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
    fblock_partials<3, kFB>(s, sh, parts_out);
}

// Blocks of the structured KA: 64 x TY tiles times gz z-chunks, gz chosen so that
// the launch has at most KaCfg<T>::blocks blocks, at least one chunk per tile column and at
// most kMaxPartials blocks.

struct KaDims {
    int gx, gy, gz;
};
template <typename T>
KaDims ka_grid_dims(const PoissonGrid<T> &P) {
    KaDims d;
    d.gx = (P.nx + KaCfg<T>::TX - 1) / KaCfg<T>::TX;
    d.gy = (P.ny + KaCfg<T>::TY - 1) / KaCfg<T>::TY;
    const int per_plane = d.gx * d.gy;
    d.gz = std::max(1, std::min(P.nz, KaCfg<T>::blocks / per_plane));
    d.gz = std::max(d.gz, (P.nz + kKaChunk - 1) / kKaChunk);   // chunks of at most kKaChunk planes
    return d;
}

int streaming_blocks(int64_t n) {
    return (int)std::max<int64_t>(1, std::min<int64_t>(MM_KB_BLOCKS, (n + kFB - 1) / kFB));
}

// Launch, timed by the dispatch's own event pair when the caller asks for it.
template <typename K, typename... Args>
void launch_timed(const Launch &L, K kernel, dim3 grid, dim3 block, Args... args) {
    if (L.t0)
        hipExtLaunchKernelGGL(kernel, grid, block, 0, L.s, L.t0, L.t1, 0, args...);
    else
        hipLaunchKernelGGL(kernel, grid, block, 0, L.s, args...);
}

}  // namespace

bool ka_grid_fits(int nx, int ny, int nz) {
    const int64_t gz = std::max<int64_t>(1, (nz + kKaChunk - 1) / kKaChunk);
    const int64_t n = (int64_t)nx * ny * nz;
    for (const int tt : {0, 1}) {
        const int tx = tt ? KaCfg<float>::TX : KaCfg<double>::TX, ty = tt ? KaCfg<float>::TY : KaCfg<double>::TY;
        const int64_t per_plane = (int64_t)((nx + tx - 1) / tx) * ((ny + ty - 1) / ty);
        if (per_plane * gz > kMaxPartials) return false;
    }
    return n * (int64_t)sizeof(double) < ((int64_t)1 << 32);
}

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
    const KaDims d = ka_grid_dims(P);
    const int nb = d.gx * d.gy * d.gz;
    if (a.mode == PCG_FIRST || a.mode == PCG_RESTART || a.mode == PCG_RESTART_PARTS)
        launch_timed(L, k_ka_grid<T, KaCfg<T>::TX, KaCfg<T>::TY, true>, dim3(nb), dim3(KaCfg<T>::TX * KaCfg<T>::TY), P, a, d.gx, d.gz, z,
                     pold, pnew, q,
                     L.dsc);
    else
        launch_timed(L, k_ka_grid<T, KaCfg<T>::TX, KaCfg<T>::TY, false>, dim3(nb), dim3(KaCfg<T>::TX * KaCfg<T>::TY), P, a, d.gx, d.gz, z,
                     pold, pnew, q,
                     L.dsc);
    return nb;
}

template <typename T>
int launch_ka_sell(const Launch &L, const SellDev<T> &A, const PcgIterArgs &a, const T *z,
                   const T *pold, T *pnew, T *q) {
    const int nb = std::max(1, std::min((A.n_slices + kFB / kWave - 1) / (kFB / kWave), kFusedBlocks));
    launch_timed(L, k_ka_sell<T>, dim3(nb), dim3(kFB), A, a, z, pold, pnew, q, L.dsc);
    return nb;
}

template <typename T>
int launch_kb(const Launch &L, int64_t n, const PcgIterArgs &a, T *x, T *r, T *z, const T *p,
              const T *q, const T *invD) {
    const int nb = streaming_blocks(n);
    if (z)
        launch_timed(L, k_kb<T, true>, dim3(nb), dim3(kFB), n, a, x, r, z, p, q, invD, L.dsc);
    else
        launch_timed(L, k_kb<T, false>, dim3(nb), dim3(kFB), n, a, x, r, z, p, q, invD, L.dsc);
    return nb;
}

void launch_pcg_check(const Launch &L, const double *parts, int nparts, int iter, int maxiter) {
    hipLaunchKernelGGL(k_pcg_check, dim3(1), dim3(kCheckThreads), 0, L.s, parts, nparts, iter, maxiter, L.dsc);
}

template <typename T>
int launch_dot2(const Launch &L, int64_t n, const T *r, const T *z, double *parts_out) {
    const int nb = streaming_blocks(n);
    hipLaunchKernelGGL(k_dot2<T>, dim3(nb), dim3(kFB), 0, L.s, n, r, z, parts_out, L.dsc);
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
template int launch_dot2<float>(const Launch &, int64_t, const float *, const float *, double *);
template int launch_dot2<double>(const Launch &, int64_t, const double *, const double *, double *);
template int launch_kb<double>(const Launch &, int64_t, const PcgIterArgs &, double *, double *, double *,
                               const double *, const double *, const double *);
template int launch_kb<float>(const Launch &, int64_t, const PcgIterArgs &, float *, float *, float *,
                              const float *, const float *, const float *);

}  // namespace mm
