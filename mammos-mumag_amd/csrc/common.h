// Internal declarations shared by the kernels (kernels.hip), the assembler
// (assemble.cpp) and the C ABI / solver drivers (capi.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mm {

constexpr int kWave = 64;          // CDNA wavefront width; SELL slice height
constexpr int kBlock = 256;        // threads per block of the streaming kernels
constexpr int kMaxPartials = 2048; // blocks of a reducing kernel (one partial each)
constexpr int kFusedBlocks = 256;  // blocks (of 1024 threads) of the fused PCG kernels

// Scalars that live on the device so that iterative solvers can run many
// iterations without a host round trip.  One instance per context.
struct DevScalars {
    // Jacobi PCG (jax CG semantics); the FP32 iteration of the mixed path reuses it.
    // gamma = r.z is double-buffered by iteration parity.
    double pcg_gam[2], pcg_rr, pcg_bound, pcg_bs, pcg_beta;
    int pcg_iter, pcg_done;
    // truncated CG (M_inv)
    double tcg_eps, tcg_rz, tcg_den, tcg_beta, tcg_pmax;
    int tcg_done, tcg_neg, tcg_iter, tcg_pad;
    long long tcg_total;
    double tcg_rzp[2];   // r.z by iteration parity (magnet-block kernels: read and written in one launch)
    // true (FP64) residual of the mixed-precision reliable-update PCG
    double ir_rr, ir_bound, ir_rmax2, ru_gamma;
    int ir_done, ir_outer, ru_maxiter, ru_pad;
    // generic reduction slots and the energy of the last evaluation
    double red[8];
    double energy;
    // multigrid V-cycles that ran (the coarse tail kernel counts itself when not skipped)
    long long mg_vcycles;
    // totals of the NCG scan between two iterations (NcgScan order)
    double ncg[16];
};

// Per-material coefficients, scaled as the reference scales them
// (scripts/materials.py:28-34: K = mu0 K1, A_eff = mu0 A / size^2) and divided by V.
struct MatCoef {
    double cex;   // 2 A_eff / V
    double cK;    // -2 K / V
    double Js;
    double u[3];
};

// SELL-C (C = 64) sparse matrix: slice s holds rows [64 s, 64 s + 64) in
// column-major order, padded to the slice's longest row.
template <typename T>
struct SellDev {
    int n_rows = 0, n_slices = 0;
    const int *slice_ptr = nullptr;   // [n_slices + 1], entry offsets
    const int *col = nullptr;
    const T *val = nullptr;
};

// Magnetic elements for the element-by-element kernels (general meshes).
struct ElemDev {
    int n = 0;
    const int4 *tets = nullptr;       // node indices
    const double *G = nullptr;        // SoA [9][n]: gradients of phi_1..phi_3 (phi_0 = -sum)
    const double *vol = nullptr;      // [n]
    const int *mat = nullptr;         // [n] material index
    // node-centric sums: nodes [nn]; per node its slots 4 e + k (element e, local node k)
    // in ascending order, CSR (inc_ptr [nn + 1], inc_slot [4 n]); contribution buffer
    // cy [3 n] (the gradient's per-element vector)
    int64_t nn = 0;
    const int *inc_ptr = nullptr, *inc_slot = nullptr;
    double *cy = nullptr;
};

// C and D of the general meshes assembled per node and material: node i has one row per
// material of its magnet elements (ascending), each row its couplings within that material
// (ascending column) as the exchange value S, the anisotropy's consistent-mass value M and
// the divergence's vector d; (C v)_i = sum over the rows (S v)_i + u_m (u_m . (M v)_i),
// u_m the material's axis, and (D v)_i = sum over the rows and entries d . v_j.
struct PairCDev {
    int64_t nn = 0;
    const int *node_ptr = nullptr;   // [nn + 1] rows of each node
    const int *row_mat = nullptr;    // [rows] material index
    const int *row_ptr = nullptr;    // [rows + 1] entry offsets
    const int *col = nullptr;        // [entries] node index
    const double2 *sm = nullptr;     // [entries] (S, M)
    // [3 entries]: the divergence's coupling d (b_i = sum over the rows d . v_j), read only
    // when b = D v is asked for
    const double *d3 = nullptr;
};

// Structured Kuhn grid: the magnet block of mn^3 uniform voxels (edge h) starting at
// grid node (m0x, m0y, m0z) of a rectilinear grid with nx * ny node rows.
struct GridDev {
    int nx = 0, ny = 0;
    int m0x = 0, m0y = 0, m0z = 0;
    int mn = 0;
    int n_mats = 0;                    // index of the all-zero material (outside voxels)
    const uint16_t *vmat = nullptr;    // [mn][mn][mn] material index per magnet voxel
    double fS = 0, fM = 0, fB = 0, fG = 0;  // h/6, h^3/120, h^2/24, h^2/(24 V)
};

// 7-point Poisson stencil on the rectilinear grid (voxel sizes per axis).
// hxp[i] = size of voxel i-1 along x for i in [1, nx-1], 0 at i = 0 and i = nx;
// ixp[i] = 1 / (6 hxp[i]) (0 for the pads).  Same for y and z.
template <typename T>
struct PoissonGrid {
    int nx = 0, ny = 0, nz = 0;        // nodes per axis
    const T *hxp = nullptr, *hyp = nullptr, *hzp = nullptr;
    const T *ixp = nullptr, *iyp = nullptr, *izp = nullptr;
};

struct Reducer {
    double *partials = nullptr;       // [4 * kMaxPartials]
};

enum FinOp {
    FIN_STORE = 0,
    FIN_TCG_INIT = 3,
    FIN_TCG_UPDATE = 4,
    FIN_RU_INIT = 5,
    FIN_RU_REPLACE = 6,
};

struct FinArgs {
    int op = FIN_STORE, nv = 1;
    unsigned maxmask = 0;
    int nparts = 1;
    const int *skip = nullptr;
    double *out = nullptr;
    double tol = 0, atol = 0;
    int maxiter = 0;
    double delta2 = 0;   // FIN_RU_*: > 0 sets the inner stop bound to max(outer, delta2 r.r)
    static FinArgs store(double *out, int nv, unsigned mask, int nparts, const int *skip) {
        FinArgs a;
        a.op = FIN_STORE;
        a.out = out;
        a.nv = nv;
        a.maxmask = mask;
        a.nparts = nparts;
        a.skip = skip;
        return a;
    }
};

// Arguments of one fused PCG iteration (pcg_fused.hip).
// RESTART_PARTS: p = z with gamma and r.r from the partials (a restart whose z came from a
// preconditioner other than Jacobi); the stop bound is kept.
enum PcgMode { PCG_FIRST = 0, PCG_NORMAL = 1, PCG_RESTART = 2, PCG_KEEP = 3, PCG_RESTART_PARTS = 4 };
struct PcgIterArgs {
    const double *parts_in = nullptr;  // partials of the previous kernel (slots of kMaxPartials)
    int nparts_in = 0;
    double *parts_out = nullptr;
    int iter = 0;                      // iteration index i (host-known)
    int mode = PCG_NORMAL;
    double tol = 0, atol = 0;          // FIRST: bound = max(tol^2 b.b, atol^2)
    int maxiter = 0;
};

struct Launch {
    hipStream_t s;
    Reducer red;
    DevScalars *dsc;
    const MatCoef *mats;
    // when set, the fused PCG launchers stamp this event pair from the dispatch itself
    // (hipExtLaunchKernelGGL), timing the kernel alone
    hipEvent_t t0 = nullptr, t1 = nullptr;
};

// ---------------- kernel launchers (kernels.hip) ----------------
void launch_fin(const Launch &L, const FinArgs &a);
// The device scalars d into the pinned host mirror h, then seq into *hseq (host memory)
void launch_publish(const Launch &L, const DevScalars *d, DevScalars *h, unsigned long long *hseq,
                    unsigned long long seq);

// y = A x, optional (x . y) into *dot_out (device); skipped when *skip_flag != 0.
void launch_sell_spmv(const Launch &L, const SellDev<double> &A, const double *x, double *y,
                      double *dot_out, const int *skip_flag);
void launch_sell_spmv32(const Launch &L, const SellDev<float> &A, const float *x, float *y,
                        double *dot_out, const int *skip_flag);
void launch_grid_poisson(const Launch &L, const PoissonGrid<double> &P, const double *x, double *y,
                         double *dot_out, const int *skip_flag);
void launch_grid_poisson32(const Launch &L, const PoissonGrid<float> &P, const float *x, float *y,
                           double *dot_out, const int *skip_flag);

// Fused PCG iteration (pcg_fused.hip); each returns the number of blocks (= partials).
template <typename T>
int launch_pcg0(const Launch &L, int64_t n, const T *b, const T *Ax0, T *r, T *z, const T *invD,
                double *parts_out);
// The structured PCG tiles fit: blocks within the partial slots, byte offsets in 32 bits.
bool ka_grid_fits(int nx, int ny, int nz);
template <typename T>
int launch_ka_grid(const Launch &L, const PoissonGrid<T> &P, const PcgIterArgs &a, const T *z,
                   const T *pold, T *pnew, T *q);
template <typename T>
int launch_ka_sell(const Launch &L, const SellDev<T> &A, const PcgIterArgs &a, const T *z,
                   const T *pold, T *pnew, T *q);
// z == nullptr: x and r only (the preconditioner runs after the kernel), no partials.
template <typename T>
int launch_kb(const Launch &L, int64_t n, const PcgIterArgs &a, T *x, T *r, T *z, const T *p,
              const T *q, const T *invD);
// Stop test after KB(i) from its r.r partials (parts slot 1, nparts of them), before a
// preconditioner that runs after KB: stops after i + 1 iterations when r.r <= the stop
// bound or i + 1 >= maxiter.
void launch_pcg_check(const Launch &L, const double *parts, int nparts, int iter, int maxiter);
// partials (r.z, r.r) of the preconditioned residual, for the next KA; skipped when done.
template <typename T>
int launch_dot2(const Launch &L, int64_t n, const T *r, const T *z, double *parts_out);
// delta2 > 0: the inner iteration stops by itself once its recursive r.r has dropped to
// delta2 times this true r.r (or to the outer bound)
void launch_ru_residual(const Launch &L, int64_t n, const double *b, const double *Ax, float *r32,
                        float *z32, float *p32, const float *invD32, double tol, double atol,
                        int maxiter, bool init, double delta2 = 0);
// The same from x on the structured grid, K x formed in the pass (as launch_grid_poisson
// forms it); grid_ru_residual_fits: its blocks fit the partial slots
bool grid_ru_residual_fits(const PoissonGrid<double> &P);
void launch_grid_ru_residual(const Launch &L, const PoissonGrid<double> &P, const double *x, const double *b,
                             float *r32, float *z32, float *p32, const float *invD32, double tol, double atol,
                             int maxiter, bool init, double delta2 = 0);
void launch_add_correction(const Launch &L, int64_t n, double *x, float *dx);

// Element-by-element operators (general meshes), accumulated into zeroed outputs.
// y = C v and, when b is given, b = D v for every node (zero rows: 0), from the assembled rows
void launch_pair_c(const Launch &L, const PairCDev &P, const double *v, double *y, double *b, const int *skip);
void launch_ebe_grad(const Launch &L, const ElemDev &E, const double *u, double *g, double invV);

// Structured magnet operators (node-centric, overwrite the magnet entries only).
void launch_grid_exani(const Launch &L, const GridDev &G, bool fp32, const double *v, double *y,
                       double *b, const int *skip_flag);
void launch_grid_grad(const Launch &L, const GridDev &G, bool fp32, const double *u, double *g);

// Node kernels.
// G: the structured grid (nodes off its magnet block are zero on input), or null;
// write_off = false: g_full and g_proj already hold zero off the block and are left there
void launch_finish_eg(const Launch &L, int64_t n, const GridDev *G, const double *m, const double *meas,
                      double zc0, double zc1, double zc2, const double *gC, const double *gd, double *g_full,
                      double *g_proj, double *E_out, int with_hmag, bool write_off = true);
void launch_external(const Launch &L, int64_t n, const double *m, const double *meas, double zc0,
                     double zc1, double zc2, double *g, double *E_out);
void launch_hvec(const Launch &L, int64_t n, const double *m0, const double *gF, const double *Cd,
                 const double *d, double *Hd, double *den_out, const int *skip_flag);
void launch_tcg_init(const Launch &L, int64_t n3, const double *g, const double *D, double *p,
                     double *r, double *z, double *d);
void launch_tcg_update(const Launch &L, int64_t n3, double *p, double *r, double *z, double *d,
                       const double *Hd, const double *D);
void launch_tcg_dupdate(const Launch &L, int64_t n3, double *d, const double *z);
void launch_tcg_final(const Launch &L, int64_t n3, const double *p, const double *g, double *zout);
// Truncated CG on the magnet block of a structured grid (vectors over the (mn+1)^3 block
// nodes, [3 t + a]); d ping-pongs between d[0] and d[1].
template <typename T>
struct TcgBlock {
    T *m0 = nullptr, *mg = nullptr, *D = nullptr, *p = nullptr, *r = nullptr, *z = nullptr, *Hd = nullptr;
    // the directions of one solve, n3 entries each: slot 0 the initial d, slot k + 1 the
    // k-th iteration's d'; p = sum_k alpha[k] slot[k + 1] is formed once at the end
    T *dring = nullptr;
    int slots = 0;
    double *alpha = nullptr;   // [slots - 1] step lengths (0: the iteration made no update)
};
// The magnet block's node tiles fit the partial slots (one partial per tile)
bool tcg_tiles_fit(int mn);
// write_off = false: zout already holds zero off the block and is written on it only;
// B.slots >= max_iter + 1
template <typename T>
void tcgc_run(const Launch &L, const GridDev &G, int64_t n, const TcgBlock<T> &B, const double *m0,
              const double *gF, const double *g, const double *D, int max_iter, double *zout,
              bool write_off = true);
// G (optional): a structured grid whose x and d are zero off the magnet block and whose
// out already holds zero there: the block's nodes only
void launch_update_m(const Launch &L, int64_t n, const double *x, double alpha, const double *d,
                     double *out, const GridDev *G = nullptr);
// G3 (optional): n = 3 N interleaved 3-vectors, x and y zero off the magnet block of G3 and
// out already zero there: the block's nodes only
void launch_axpby(const Launch &L, int64_t n, double a, const double *x, double b, const double *y,
                  double *out, const GridDev *G3 = nullptr);
void launch_fill_zero(const Launch &L, int64_t n, double *x, const int *skip_flag);
void launch_fill_zero32(const Launch &L, int64_t n, float *x, const int *skip_flag);

// Fused multi-reductions used by the NCG.
enum RedOp {
    RED_DOT = 0,          // out0 = a.b
    RED_NORMINF = 1,      // out0 = max |a|
    RED_STOP3 = 2,        // out0 = max|a-b|, out1 = max|b|, out2 = max|c|
    RED_DG3 = 3,          // out0 = a.b, out1 = a.a, out2 = b.b
    RED_H3 = 4,           // out0 = a.a, out1 = b.b, out2 = max|b|
    RED_GD2 = 5,          // out0 = a.c, out1 = b.c
    RED_YZD = 6,          // out0 = (b-a).c, out1 = (b-a).d
    RED_MH = 7,           // out0 = sum_i (h . m_i) meas_i  (a = m, b = meas)
    RED_SS = 8,           // out0 = a.a
};
// out = max |m| over the nodes off the magnet block of a structured grid
void launch_outside_max(const Launch &L, int64_t n, const GridDev &G, const double *m, double *out);
// G (optional): a structured grid whose vectors are zero off the magnet block, visited
// on the block only (comp = 3 for interleaved 3-vectors of n = 3 N, 1 for node vectors)
void launch_reduce(const Launch &L, RedOp op, int64_t n, const double *a, const double *b,
                   const double *c, const double *d, double h0, double h1, double h2, double *out,
                   const GridDev *G = nullptr, int comp = 1);

// Totals of the NCG scan between two iterations (k_ncg_scan): the old iterate xo and
// gradient go, the new xn and gn, the preconditioned gradient z, the direction d.
enum NcgScan {
    NCG_DG_OLD = 0,   // d.go   (the restart test after the HS+ update)
    NCG_DD = 1,       // d.d
    NCG_GG_OLD = 2,   // go.go
    NCG_DX_MAX = 3,   // max |xo - xn|   (the stop test)
    NCG_X_MAX = 4,    // max |xn|
    NCG_G_MAX = 5,    // max |gn|
    NCG_DG = 6,       // d.gn   (the restart test before the line search)
    NCG_GG = 7,       // gn.gn
    NCG_ZG = 8,       // z.gn
    NCG_ZZ = 9,       // z.z
    NCG_XX = 10,      // xn.xn  (the line search's first step)
    NCG_D_MAX = 11,   // max |d|
    NCG_Z_MAX = 12,   // max |z|
    NCG_NV = 13,
};
constexpr unsigned kNcgMaxMask = (1u << NCG_DX_MAX) | (1u << NCG_X_MAX) | (1u << NCG_G_MAX) |
                                 (1u << NCG_D_MAX) | (1u << NCG_Z_MAX);
// With yzd (device pointer to the totals (g1 - g).z, (g1 - g).d) first d = -z + beta d,
// beta = max(yzd[0] / yzd[1], 0); then the NCG_NV totals into out (device).  G as for
// launch_reduce over n3 = 3 N interleaved entries.  partials: NCG_NV * kMaxPartials
// doubles.  Returns the block count.
int launch_ncg_scan(const Launch &L, int64_t n3, const GridDev *G, const double *xo, const double *xn,
                    const double *go, const double *gn, const double *z, double *d, const double *yzd,
                    double *partials, double *out);

}  // namespace mm
