// Host-only check of mg_restrict_span_fits (mammos-mumag_amd/csrc/mg.h), the guard that
// mm_multigrid applies to every level's x restriction lists before the restriction kernel
// may stage them.  Prints one line per case: name, expected, result.
#include <cstdio>
#include <vector>

#include "mg.h"

using mm::mg_restrict_span_fits;

// Restriction pairs of linear interpolation from the kept fine nodes `keep` of nf fine
// nodes on a uniform axis: fine i feeds its coarse interval's ends with weights 1 - t, t.
static void pairs(const std::vector<int> &keep, int nf, std::vector<int> &ri, std::vector<float> &rw) {
    const int nc = (int)keep.size();
    ri.assign(3 * nc, 0);
    rw.assign(3 * nc, 0.f);
    std::vector<int> fill(nc, 0);
    int c0 = 0;
    for (int i = 0; i < nf; ++i) {
        while (c0 + 1 < nc && keep[c0 + 1] <= i) ++c0;
        const double t = keep[c0] == i ? 0.0 : double(i - keep[c0]) / (keep[c0 + 1] - keep[c0]);
        const double w[2] = {1.0 - t, t};
        for (int q = 0; q < 2; ++q) {
            if (w[q] == 0.0) continue;
            const int C = c0 + q;
            if (fill[C] < 3) {   // a coarse node takes at most 3 pairs (the kernel's layout)
                ri[3 * C + fill[C]] = i;
                rw[3 * C + fill[C]] = (float)w[q];
                ++fill[C];
            }
        }
    }
}

int main() {
    int bad = 0;
    auto check = [&](const char *name, const std::vector<int> &keep, int nf, bool want) {
        std::vector<int> ri;
        std::vector<float> rw;
        pairs(keep, nf, ri, rw);
        const bool got = mg_restrict_span_fits(ri.data(), rw.data(), (int)keep.size());
        std::printf("%s %d %d\n", name, (int)want, (int)got);
        bad += got != want;
    };
    // every interval merged (two fine per coarse): 64 coarse nodes use 2 * 63 + 2 fine ones
    std::vector<int> k2;
    for (int i = 0; i <= 300; i += 2) k2.push_back(i);
    check("two_fine_per_coarse", k2, 301, true);
    // nothing merged
    std::vector<int> k1;
    for (int i = 0; i <= 200; ++i) k1.push_back(i);
    check("identity", k1, 201, true);
    // one coarse interval of three fine ones among merged pairs: the row's slack takes it
    std::vector<int> k3;
    for (int i = 0; i <= 120; i += 2) k3.push_back(i);
    k3.push_back(123);
    for (int i = 125; i <= 301; i += 2) k3.push_back(i);
    check("one_three_fine_interval", k3, 302, true);
    // seven of them among the first 64 coarse nodes: 7 * 3 + 56 * 2 fine intervals plus the
    // next node's pairs reach past 132
    std::vector<int> k5{0};
    for (int q = 0; q < 150; ++q) k5.push_back(k5.back() + (q < 7 ? 3 : 2));
    check("seven_three_fine_intervals", k5, k5.back() + 1, false);
    // three-interval merges throughout
    std::vector<int> k4;
    for (int i = 0; i <= 300; i += 3) k4.push_back(i);
    check("three_fine_per_coarse", k4, 301, false);
    return bad;
}
