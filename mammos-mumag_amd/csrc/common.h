// Internal declarations shared by the kernels (kernels.hip), the assembler
// (assemble.cpp) and the C ABI / solver drivers (capi.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace mm {

constexpr int kWave = 64;        // CDNA wavefront width; SELL slice height
constexpr int kBlock = 256;      // threads per block of the streaming kernels
constexpr int kMaxPartials = 8192;

// Scalars that live on the device so that iterative solvers can run many
// iterations without a host round trip.  One instance per context.
struct DevScalars {
    // Jacobi PCG (jax CG semantics)
    double pcg_gamma, pcg_pq, pcg_rr, pcg_bound, pcg_bs, pcg_beta;
    int pcg_iter, pcg_maxiter, pcg_done, pcg_pad;
    // truncated CG (M_inv)
    double tcg_eps, tcg_rz, tcg_den, tcg_rn2, tcg_rznew, tcg_beta, tcg_alpha, tcg_pmax;
    int tcg_done, tcg_neg, tcg_iter, tcg_pad;
    // generic reduction slots
    double red[16];
    // energy of the last evaluation
    double energy;
    // mixed precision refinement
    double ir_rr, ir_bound;
};

// Per-material coefficients, already scaled as the reference scales them
// (scripts/materials.py:28-34: K = mu0 K1, A_eff = mu0 A / size^2; divided by V).
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
    const float *G32 = nullptr;       // same in FP32 (mixed precision)
    const double *vol = nullptr;      // [n]
    const int *mat = nullptr;         // [n] material index
};

// Structured Kuhn grid: magnet voxels on a uniform sub-grid.
struct GridDev {
    int nx = 0, ny = 0, nz = 0;       // nodes per axis of the full grid
    int mx0 = 0, my0 = 0, mz0 = 0;    // first voxel of the magnet block
    int mnx = 0, mny = 0, mnz = 0;    // magnet voxels per axis
    double h = 0;                     // magnet voxel edge
    const uint8_t *vmat = nullptr;    // [mnz][mny][mnx] material index per magnet voxel
};

struct Reducer {
    double *partials = nullptr;       // [kMaxPartials * 4]
    unsigned *counter = nullptr;
};

// ---------------- kernel launchers (kernels.hip) ----------------
struct Launch {
    hipStream_t s;
    Reducer red;
    DevScalars *dsc;
    const MatCoef *mats;
    int n_mats;
};

// SELL SpMV y = A x, optional fused dot (x . y) into *dot_out (device), skipped when
// *skip_flag != 0 (device).
void launch_sell_spmv(const Launch &L, const SellDev<double> &A, const double *x, double *y,
                      double *dot_out, const int *skip_flag);
void launch_sell_spmv32(const Launch &L, const SellDev<float> &A, const float *x, float *y,
                        double *dot_out, const int *skip_flag);

// PCG (jax CG semantics) building blocks, FP64.
void launch_pcg_init(const Launch &L, int n, const double *b, const double *Ax0, double *r,
                     double *z, double *p, const double *invD, double tol, double atol,
                     int maxiter);
void launch_pcg_update(const Launch &L, int n, double *x, double *r, double *z, const double *p,
                       const double *q, const double *invD);
void launch_pcg_pupdate(const Launch &L, int n, double *p, const double *z);

// Mixed-precision inner PCG (FP32 vectors, FP64 scalars).
void launch_pcg32_init(const Launch &L, int n, const float *r, float *z, float *p, float *x,
                       const float *invD, double tol_abs2, int maxiter);
void launch_pcg32_update(const Launch &L, int n, float *x, float *r, float *z, const float *p,
                         const float *q, const float *invD);
void launch_pcg32_pupdate(const Launch &L, int n, float *p, const float *z);
void launch_residual_f64(const Launch &L, int n, const double *b, const double *Ax, double *r,
                         float *r32, double *rr_out);
void launch_add_correction(const Launch &L, int n, double *x, const float *dx);

// Element-by-element operators (general meshes).  Outputs are accumulated with
// atomics: the caller zeroes them first.
void launch_ebe_exani(const Launch &L, const ElemDev &E, const double *v, double *y, double *b,
                      const int *skip_flag);
void launch_ebe_grad(const Launch &L, const ElemDev &E, const double *u, double *g, double invV);

// Structured-grid operators (node-centric stencils, no atomics).
void launch_grid_exani(const Launch &L, const GridDev &G, const double *v, double *y,
                       double *b, const int *skip_flag, const int *node_of_mag);
void launch_grid_grad(const Launch &L, const GridDev &G, const double *u, double *g,
                      double invV, const int *node_of_mag);

// Node kernels.
void launch_finish_eg(const Launch &L, int n, const double *m, const double *meas, double zc0,
                      double zc1, double zc2, const double *gC, const double *gd, double *g_full,
                      double *g_proj, double *E_out, int with_hmag);
void launch_external(const Launch &L, int n, const double *m, const double *meas, double zc0,
                     double zc1, double zc2, double *g, double *E_out);
void launch_project(const Launch &L, int n, const double *m, const double *g, double *out);
void launch_hvec(const Launch &L, int n, const double *m0, const double *gF, const double *Cd,
                 const double *d, double *Hd, double *den_out, const int *skip_flag);
void launch_tcg_init(const Launch &L, int n3, const double *g, const double *D, double *p,
                     double *r, double *z, double *d);
void launch_tcg_update(const Launch &L, int n3, double *p, double *r, double *z, double *d,
                       const double *Hd, const double *D);
void launch_tcg_dupdate(const Launch &L, int n3, double *d, const double *z);
void launch_tcg_final(const Launch &L, int n3, const double *p, const double *g, double *zout);
void launch_update_m(const Launch &L, int n, const double *x, double alpha, const double *d,
                     double *out);
void launch_axpby(const Launch &L, int64_t n, double a, const double *x, double b,
                  const double *y, double *out);
void launch_select_neg(const Launch &L, int64_t n, const double *src, double *dst);

// Generic reductions: op codes of the fused multi-reductions used by the NCG.
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
void launch_reduce(const Launch &L, RedOp op, int64_t n, const double *a, const double *b,
                   const double *c, const double *d, double h0, double h1, double h2,
                   double *out);

void launch_fill_zero(const Launch &L, int64_t n, double *x, const int *skip_flag);
void launch_f64_to_f32(const Launch &L, int64_t n, const double *x, float *y);
void launch_f32_to_f64(const Launch &L, int64_t n, const float *x, double *y);

}  // namespace mm
