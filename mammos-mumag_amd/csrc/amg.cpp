// Smoothed-aggregation AMG set-up on the host (OpenMP for the products); the steps and
// their order follow the CPU restatement in oracle/amg.py, so that both build the same
// hierarchy (up to the summation order of the Galerkin products).
#include "amg.h"

#include <algorithm>
#include <cmath>
#include <cstdint>

namespace mm {
namespace {

// Rows whose stored off-diagonal entries are all zero (the Dirichlet identity rows)
std::vector<char> free_rows(const CsrHost &A) {
    std::vector<char> fr(A.n, 0);
    for (int i = 0; i < A.n; ++i)
        for (int q = A.ptr[i]; q < A.ptr[i + 1]; ++q)
            if (A.col[q] != i && A.val[q] != 0.0) {
                fr[i] = 1;
                break;
            }
    return fr;
}

std::vector<double> diag_of(const CsrHost &A) {
    std::vector<double> d(A.n, 0.0);
    for (int i = 0; i < A.n; ++i)
        for (int q = A.ptr[i]; q < A.ptr[i + 1]; ++q)
            if (A.col[q] == i) d[i] += A.val[q];
    return d;
}

double gershgorin(const CsrHost &A, const std::vector<double> &d) {
    double rho = 0.0;
    for (int i = 0; i < A.n; ++i) {
        double s = 0.0;
        for (int q = A.ptr[i]; q < A.ptr[i + 1]; ++q) s += std::fabs(A.val[q]);
        rho = std::max(rho, s / d[i]);
    }
    return rho;
}

std::vector<double> omegas(int deg, double lo, double hi) {
    std::vector<double> om(deg);
    for (int k = 0; k < deg; ++k)
        om[k] = 1.0 / ((hi + lo) / 2 + (hi - lo) / 2 * std::cos(M_PI * (2 * k + 1) / (2.0 * deg)));
    return om;
}

// Aggregate per node (-1 off the free set) and the aggregate count; strong neighbours of
// i are the free j != i with s_ij = |a_ij| / sqrt(a_ii a_jj) > max(theta, 1e-10) (couplings
// that cancel to rounding level never count), visited in ascending column order; in pass 2
// values within 1e-9 relative are ties, which go to the lower index.
int aggregate(const CsrHost &A, const std::vector<char> &fr, const std::vector<double> &d, double theta,
              std::vector<int> &agg) {
    const int n = A.n;
    agg.assign(n, -1);
    std::vector<double> sd(n);
    for (int i = 0; i < n; ++i) sd[i] = std::sqrt(std::fabs(d[i]));
    auto sval = [&](int i, int q) { return std::fabs(A.val[q]) / (sd[i] * sd[A.col[q]]); };
    const double th = std::max(theta, 1e-10);
    auto strong = [&](int i, int q) { const int j = A.col[q]; return j != i && fr[j] && sval(i, q) > th; };
    int na = 0;
    for (int i = 0; i < n; ++i) {   // pass 1
        if (!fr[i] || agg[i] >= 0) continue;
        bool all_free = true;
        for (int q = A.ptr[i]; q < A.ptr[i + 1] && all_free; ++q)
            if (strong(i, q) && agg[A.col[q]] >= 0) all_free = false;
        if (!all_free) continue;
        agg[i] = na;
        for (int q = A.ptr[i]; q < A.ptr[i + 1]; ++q)
            if (strong(i, q)) agg[A.col[q]] = na;
        ++na;
    }
    const std::vector<int> first = agg;
    for (int i = 0; i < n; ++i) {   // pass 2
        if (!fr[i] || agg[i] >= 0) continue;
        double best = -1.0;
        int pick = -1;
        for (int q = A.ptr[i]; q < A.ptr[i + 1]; ++q) {
            if (!strong(i, q)) continue;
            const int j = A.col[q];
            const double s = sval(i, q);
            if (first[j] >= 0 && s > best * (1 + 1e-9)) {
                best = s;
                pick = j;
            }
        }
        if (pick >= 0) agg[i] = first[pick];
    }
    for (int i = 0; i < n; ++i) {   // pass 3
        if (!fr[i] || agg[i] >= 0) continue;
        agg[i] = na;
        for (int q = A.ptr[i]; q < A.ptr[i + 1]; ++q)
            if (strong(i, q) && agg[A.col[q]] < 0) agg[A.col[q]] = na;
        ++na;
    }
    return na;
}

// Rows of a product built with a dense accumulator per thread; drops exact zeros.
template <typename RowFn>
CsrHost build_rows(int n, int m, RowFn row) {
    CsrHost C;
    C.n = n;
    C.m = m;
    std::vector<std::vector<int>> cols(n);
    std::vector<std::vector<double>> vals(n);
#pragma omp parallel
    {
        std::vector<double> acc(m, 0.0);
        std::vector<char> used(m, 0);
        std::vector<int> touched;
#pragma omp for schedule(dynamic, 256)
        for (int i = 0; i < n; ++i) {
            touched.clear();
            row(i, [&](int c, double v) {
                if (!used[c]) {
                    used[c] = 1;
                    touched.push_back(c);
                }
                acc[c] += v;
            });
            std::sort(touched.begin(), touched.end());
            for (int c : touched) {
                if (acc[c] != 0.0) {
                    cols[i].push_back(c);
                    vals[i].push_back(acc[c]);
                }
                acc[c] = 0.0;
                used[c] = 0;
            }
        }
    }
    C.ptr.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) C.ptr[i + 1] = C.ptr[i] + (int)cols[i].size();
    C.col.resize(C.ptr[n]);
    C.val.resize(C.ptr[n]);
    for (int i = 0; i < n; ++i) {
        std::copy(cols[i].begin(), cols[i].end(), C.col.begin() + C.ptr[i]);
        std::copy(vals[i].begin(), vals[i].end(), C.val.begin() + C.ptr[i]);
    }
    return C;
}

CsrHost transpose(const CsrHost &A) {
    CsrHost T;
    T.n = A.m;
    T.m = A.n;
    T.ptr.assign(A.m + 1, 0);
    for (int q = 0; q < A.ptr[A.n]; ++q) T.ptr[A.col[q] + 1]++;
    for (int c = 0; c < A.m; ++c) T.ptr[c + 1] += T.ptr[c];
    T.col.resize(A.ptr[A.n]);
    T.val.resize(A.ptr[A.n]);
    std::vector<int> fill(T.ptr.begin(), T.ptr.end() - 1);
    for (int i = 0; i < A.n; ++i)   // rows ascending: each row of T gets its columns in order
        for (int q = A.ptr[i]; q < A.ptr[i + 1]; ++q) {
            const int p = fill[A.col[q]]++;
            T.col[p] = i;
            T.val[p] = A.val[q];
        }
    return T;
}

// C = A B (CSR, A.m == B.n)
CsrHost multiply(const CsrHost &A, const CsrHost &B) {
    return build_rows(A.n, B.m, [&](int i, auto add) {
        for (int q = A.ptr[i]; q < A.ptr[i + 1]; ++q) {
            const int j = A.col[q];
            const double a = A.val[q];
            for (int t = B.ptr[j]; t < B.ptr[j + 1]; ++t) add(B.col[t], a * B.val[t]);
        }
    });
}

// Dense inverse of an SPD matrix by Cholesky; false if a pivot is not positive
bool spd_inverse(const CsrHost &A, std::vector<double> &inv) {
    const int n = A.n;
    std::vector<double> L((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int q = A.ptr[i]; q < A.ptr[i + 1]; ++q) L[(size_t)i * n + A.col[q]] = A.val[q];
    for (int j = 0; j < n; ++j) {
        double s = L[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) s -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
        if (!(s > 0.0)) return false;
        const double ljj = std::sqrt(s);
        L[(size_t)j * n + j] = ljj;
        for (int i = j + 1; i < n; ++i) {
            double t = L[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) t -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
            L[(size_t)i * n + j] = t / ljj;
        }
    }
    inv.assign((size_t)n * n, 0.0);
#pragma omp parallel for schedule(dynamic, 8)
    for (int c = 0; c < n; ++c) {   // column c of the inverse: L L^T x = e_c
        std::vector<double> y(n, 0.0);
        for (int i = 0; i < n; ++i) {
            double s = i == c ? 1.0 : 0.0;
            for (int k = 0; k < i; ++k) s -= L[(size_t)i * n + k] * y[k];
            y[i] = s / L[(size_t)i * n + i];
        }
        for (int i = n - 1; i >= 0; --i) {
            double s = y[i];
            for (int k = i + 1; k < n; ++k) s -= L[(size_t)k * n + i] * y[k];
            y[i] = s / L[(size_t)i * n + i];
        }
        for (int i = 0; i < n; ++i) inv[(size_t)i * n + c] = y[i];
    }
    return true;
}

}  // namespace

std::string amg_setup(const CsrHost &A0, int max_levels, int coarse, int deg0, int deg, double alpha, double theta,
                      AmgHost &out) {
    out.levels.clear();
    CsrHost A = A0;
    std::vector<char> fr = free_rows(A);
    for (;;) {
        AmgLevelHost lev;
        const std::vector<double> d = diag_of(A);
        for (int i = 0; i < A.n; ++i)
            if (!(d[i] > 0.0)) return "amg: non-positive diagonal";
        lev.dinv.resize(A.n);
        for (int i = 0; i < A.n; ++i) lev.dinv[i] = 1.0 / d[i];
        lev.rho = gershgorin(A, d);
        lev.om = omegas(out.levels.empty() ? deg0 : deg, lev.rho / alpha, lev.rho);
        const int n = A.n;
        if (n <= coarse || (int)out.levels.size() + 1 == max_levels) {
            lev.A = std::move(A);
            out.levels.push_back(std::move(lev));
            break;
        }
        std::vector<int> agg;
        const int na = aggregate(A, fr, d, theta, agg);
        if (na > 0.8 * n) {
            lev.A = std::move(A);
            out.levels.push_back(std::move(lev));
            break;
        }
        // P = T - w D^-1 (A T): row i of A T sums a_ij over the aggregate of each column j
        const double w = 4.0 / (3.0 * lev.rho);
        lev.P = build_rows(n, na, [&](int i, auto add) {
            if (agg[i] >= 0) add(agg[i], 1.0);
            const double s = w * lev.dinv[i];
            for (int q = A.ptr[i]; q < A.ptr[i + 1]; ++q) {
                const int a = agg[A.col[q]];
                if (a >= 0) add(a, -s * A.val[q]);
            }
        });
        lev.R = transpose(lev.P);
        CsrHost Ac = multiply(lev.R, multiply(A, lev.P));
        lev.A = std::move(A);
        out.levels.push_back(std::move(lev));
        A = std::move(Ac);
        fr.assign(A.n, 1);
    }
    const CsrHost &last = out.levels.back().A;
    if (last.n > 2000) return "amg: coarsening stalled above 2000 unknowns";
    if (!spd_inverse(last, out.inv)) return "amg: the coarsest matrix is not positive definite";
    return "";
}

}  // namespace mm
