// Smoothed-aggregation algebraic multigrid for the general (unstructured) meshes' inner
// Poisson solve: host set-up (amg.cpp) and the V-cycle kernels on SELL-64 matrices
// (amg.hip).  No counterpart in the reference, whose potential solve is
// Jacobi-preconditioned (scripts/solvers.py:222-231).
#pragma once
#include <string>
#include <vector>

#include "common.h"

namespace mm {

constexpr int kAmgMaxLevels = 10;

// CSR on the host: n rows, m columns, columns ascending within a row.
struct CsrHost {
    int n = 0, m = 0;
    std::vector<int> ptr, col;
    std::vector<double> val;
};

struct AmgLevelHost {
    CsrHost A;                  // the level's matrix (n x n)
    std::vector<double> dinv;   // 1 / diag(A)
    double rho = 0.0;           // Gershgorin bound of D^-1 A
    std::vector<double> om;     // Chebyshev-Richardson weights, largest node first
    CsrHost P, R;               // prolongation (n x nc) and R = P^T; not on the last level
};

struct AmgHost {
    std::vector<AmgLevelHost> levels;
    std::vector<double> inv;    // the last level's inverse, row-major
};

// The hierarchy of A (square CSR with Dirichlet rows as identity rows): strong couplings
// |a_ij| > theta sqrt(a_ii a_jj) between free nodes, three-pass aggregation in index order,
// P = (I - 4 / (3 rho) D^-1 A) T, A_c = P^T A P, down to at most `coarse` unknowns;
// smoothing of degree `deg0` on level 0 and `deg` on the coarser levels, on
// [rho / alpha, rho].  Returns "" or an error message.
std::string amg_setup(const CsrHost &A, int max_levels, int coarse, int deg0, int deg, double alpha, double theta,
                      AmgHost &out);

// CSR on the device (FP32) for rows with many entries (the restrictions, the matrices of
// the coarse levels): `group` lanes per row (a power of two, 4..64), chosen from the mean
// row length.
struct CsrDevF {
    int n_rows = 0, group = 4;
    const int *ptr = nullptr, *col = nullptr;
    const float *val = nullptr;
};

// Device side of one level (FP32): the matrix (level 0: the context's K in SELL-64, A;
// coarser levels: CSR, Ac), 1 / diag, the transfers to the next level (P in SELL-64, its
// rows hold a few entries; R = P^T in CSR) and the level's vectors.
struct AmgLevelDev {
    SellDev<float> A, P;
    CsrDevF Ac, R;
    int n = 0;
    const float *dinv = nullptr;
    float om[8] = {};
    int deg = 0;
    float *b = nullptr, *xa = nullptr, *xb = nullptr, *r = nullptr;
};

// out = x + w dinv (b - A x) (x == nullptr: out = w dinv b); resid: out = b - A x
void launch_amg_smooth(const Launch &L, const SellDev<float> &A, const float *dinv, const float *b,
                       const float *x, float *out, float w, const int *skip);
void launch_amg_resid(const Launch &L, const SellDev<float> &A, const float *b, const float *x, float *out,
                      const int *skip);
// the same on a coarse level's CSR matrix (x given)
void launch_amg_smooth(const Launch &L, const CsrDevF &A, const float *dinv, const float *b, const float *x,
                       float *out, float w, const int *skip);
void launch_amg_resid(const Launch &L, const CsrDevF &A, const float *b, const float *x, float *out,
                      const int *skip);
// bc = R r and, when x1c is given, x1c = w dinvc bc (the coarse level's first smoothing step)
void launch_amg_restrict(const Launch &L, const CsrDevF &R, const float *r, float *bc, float *x1c,
                         const float *dinvc, float w, const int *skip);
// x += P xc
void launch_amg_prolong(const Launch &L, const SellDev<float> &P, const float *xc, float *x, const int *skip);
// x = inv b, inv dense n x n row-major
void launch_amg_dense(const Launch &L, const float *inv, int n, const float *b, float *x, const int *skip);

}  // namespace mm
