// Geometric multigrid V-cycle for the structured Poisson stencil (mg.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.h"

namespace mm {

constexpr int kMgMaxLevels = 8;

// One level as the kernels see it, with the transfers to the next coarser level:
// restriction lists per coarse index (3 (fine index, weight) pairs, weight 0 = unused)
// and prolongation pairs per fine index (coarse c0, weight t of c0 + 1).
struct MgLevelDev {
    PoissonGrid<float> P;
    float *invw = nullptr;   // first smoothing weight / diag of the level's operator, 0 on the faces
    const int *rxi = nullptr, *ryi = nullptr, *rzi = nullptr;
    const float *rxw = nullptr, *ryw = nullptr, *rzw = nullptr;
    const int *pxc = nullptr, *pyc = nullptr, *pzc = nullptr;
    const float *pxt = nullptr, *pyt = nullptr, *pzt = nullptr;
};

void launch_mg_first(const Launch &L, const PoissonGrid<float> &P, const float *b, float *out, float w,
                     const int *skip);
// returns the number of blocks (= partials written when dot_parts != null)
int launch_mg_smooth(const Launch &L, const PoissonGrid<float> &P, const float *b, const float *x, float *out,
                     float w, const int *skip, double *dot_parts = nullptr);
void launch_mg_resid(const Launch &L, const PoissonGrid<float> &P, const float *b, const float *x, float *out,
                     const int *skip);
// b_c = R r_f and the coarse level's first smoothing step x1c = c.invw b_c
void launch_mg_restrict(const Launch &L, const MgLevelDev &f, const MgLevelDev &c, const float *rf, float *bc,
                        float *x1c, const int *skip);
bool mg_stencil_fits(const PoissonGrid<float> &P);
void launch_mg_prolong(const Launch &L, const MgLevelDev &f, const MgLevelDev &c, const float *xc, float *xf,
                       const int *skip);

}  // namespace mm
