// Geometric multigrid V-cycle for the structured Poisson stencil (mg.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.h"

namespace mm {

constexpr int kMgMaxLevels = 8;

// Tile grid of a per-level kernel: x tiles, y tiles, z chunks (launched as x * y * z blocks)
struct MgGrid {
    int x, y, z;
};

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

// The last levels of the V-cycle in one workgroup with their vectors in LDS: b0 is the
// right-hand side of the first tail level (global), x0 receives its result (all of it,
// zero on the faces); n[q] is tail level q's node count, off[q] the LDS offset (floats)
// of its vectors (x, y, and b for q > 0), coef_off[q] of its stencil factors and
// tr_off[q] of its transfer lists (mg_tail_layout fills the offsets, img_n and
// lds_bytes).  The tables (img_n floats from coef_off[0]) are formed once by
// launch_mg_tail_image into img (img_n floats, 16-byte aligned), which the tail copies.
constexpr int kMgTailMax = 4;
constexpr int kMgTailLdsBytes = 64 * 1024;
constexpr int kMgMaxDeg = 16;      // smoothing / coarsest-level polynomial degrees at most
struct MgTailArgs {
    int nt = 0, s = 0, cd = 0;
    int n[kMgTailMax] = {}, off[kMgTailMax] = {}, coef_off[kMgTailMax] = {}, tr_off[kMgTailMax] = {};
    int lds_bytes = 0;
    MgLevelDev lev[kMgTailMax];
    int img_n = 0;
    const float *img = nullptr;
    const float *b0 = nullptr;
    float *x0 = nullptr;
    float om[kMgMaxDeg] = {}, omc[kMgMaxDeg] = {};
    long long *count = nullptr;   // incremented once per launch that runs (V-cycles)
};
int mg_tail_layout(MgTailArgs &a);
// the laid-out tail fits the kernel: its LDS, and at most one z column per thread per level
bool mg_tail_fits(const MgTailArgs &a);
void launch_mg_tail_image(const Launch &L, const MgTailArgs &a, float *img);
void launch_mg_tail(const Launch &L, const MgTailArgs &a, const int *skip);

void launch_mg_first(const Launch &L, const PoissonGrid<float> &P, const float *b, float *out, float w,
                     const int *skip);
// returns the number of blocks (= partials written when dot_parts != null)
int launch_mg_smooth(const Launch &L, const PoissonGrid<float> &P, const float *b, const float *x, float *out,
                     float w, const int *skip, double *dot_parts = nullptr);
void launch_mg_resid(const Launch &L, const PoissonGrid<float> &P, const float *b, const float *x, float *out,
                     const int *skip);
// out = two smoothing steps (weights w1 then w2) from x, in one pass (for small levels)
void launch_mg_smooth2(const Launch &L, const PoissonGrid<float> &P, const float *b, const float *x, float *out,
                       float w1, float w2, const int *skip);
// b_c = R r_f and the coarse level's first smoothing step x1c = c.invw b_c
void launch_mg_restrict(const Launch &L, const MgLevelDev &f, const MgLevelDev &c, const float *rf, float *bc,
                        float *x1c, const int *skip);
bool mg_stencil_fits(const PoissonGrid<float> &P);
// The restriction kernel's wave covers kMgRestrictTX consecutive coarse x indices and
// stages the fine nodes they use in one row of kMgRestrictSpan nodes from the first coarse
// node's first fine index (2 * 63 + 3 fine indices are enough when every coarse interval
// covers at most two fine ones).  The x restriction lists (ri, rw: 3 pairs per coarse
// index, nc coarse indices) fit when every used fine index lies in that row.
constexpr int kMgRestrictTX = 64, kMgRestrictSpan = 132;
inline bool mg_restrict_span_fits(const int *ri, const float *rw, int nc) {
    for (int i0 = 0; i0 < nc; i0 += kMgRestrictTX) {
        const int base = ri[3 * i0];
        const int i1 = i0 + kMgRestrictTX < nc ? i0 + kMgRestrictTX : nc;
        for (int I = i0; I < i1; ++I)
            for (int q = 0; q < 3; ++q)
                if (rw[3 * I + q] != 0.f && (ri[3 * I + q] < base || ri[3 * I + q] - base >= kMgRestrictSpan))
                    return false;
    }
    return true;
}
void launch_mg_prolong(const Launch &L, const MgLevelDev &f, const MgLevelDev &c, const float *xc, float *xf,
                       const int *skip);

}  // namespace mm
