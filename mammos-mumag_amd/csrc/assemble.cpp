// Host-side P1 assembly (FP64) of the operators of the hot path.
//
// Semantics follow the escript LinearPDE operators the reference converts to sparse
// matrices (scripts/mapping.py:171-236, scripts/matrix.py:9-74, scripts/trafo.py:4-45,
// scripts/escript_tools.py:13-17):
//   K_ij   = sum_e int grad phi_i . k_e grad phi_j   (k = I, or the Imhoff shell tensor
//            integrated with the 4-point degree-2 Keast rule); Dirichlet rows and
//            columns replaced by identity.
//   meas_i = sum_e Js_e vol_e / 4
//   Dex    = diagonal of the exchange operator (2 A_eff / V) int |grad phi_i|^2, 1 where
//            meas == 0 (scripts/mapping.py:207-209)
// Rows are assembled independently (node -> element incidence), so the loops run in
// parallel without atomics and the summation order per entry is the element order.
#include "assemble.h"

#include <algorithm>
#include <cmath>
#include <cstring>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace mm {

namespace {

constexpr double kMu0 = 4.0e-7 * M_PI;  // scripts/materials.py:21
constexpr double kKeastA = 0.5854101966249685;
constexpr double kKeastB = 0.1381966011250105;

struct Geo {
    double G[4][3];
    double vol;
};

inline void cross(const double *a, const double *b, double *c) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// Barycentric gradients: with edge rows e_k = x_k - x_0, column k of J^-1 is
// (e_{k+1} x e_{k+2}) / det, i.e. G_1 = e2 x e3 / det, G_2 = e3 x e1 / det,
// G_3 = e1 x e2 / det, G_0 = -(G_1 + G_2 + G_3).
inline bool geometry(const double *X, const int *t, Geo &g) {
    double e[3][3];
    for (int k = 0; k < 3; ++k)
        for (int a = 0; a < 3; ++a) e[k][a] = X[3 * (int64_t)t[k + 1] + a] - X[3 * (int64_t)t[0] + a];
    double c23[3], c31[3], c12[3];
    cross(e[1], e[2], c23);
    cross(e[2], e[0], c31);
    cross(e[0], e[1], c12);
    const double det = e[0][0] * c23[0] + e[0][1] * c23[1] + e[0][2] * c23[2];
    if (!(std::fabs(det) > 0.0)) return false;
    for (int a = 0; a < 3; ++a) {
        g.G[1][a] = c23[a] / det;
        g.G[2][a] = c31[a] / det;
        g.G[3][a] = c12[a] / det;
        g.G[0][a] = -(g.G[1][a] + g.G[2][a] + g.G[3][a]);
    }
    g.vol = std::fabs(det) / 6.0;
    return true;
}

// scripts/trafo.py:24-45 (IEEE Trans. Magn. 26 (1990), Imhoff et al.)
inline void t3d(const double *xq, double R, double Rinf, double k[3][3]) {
    const double x = std::sqrt(xq[0] * xq[0] + xq[1] * xq[1] + xq[2] * xq[2]);
    const double a2 = R * (Rinf - R);
    const double ax = Rinf / x;
    const double a2x2 = a2 / (x * x);
    const double f = ax * (2 - ax) / ((ax - 1) * (ax - 1));
    const double n[3] = {xq[0] / x, xq[1] / x, xq[2] / x};
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            k[a][b] = (a == b) ? a2x2 * (1 + f * (1 - n[a] * n[a])) : -a2x2 * (f * n[a] * n[b]);
}

// Row k of the local stiffness (4 entries) of element e.
inline void stiff_row(const mm_mesh_desc &d, const double *X, const int *t, bool shell,
                      const Geo &g, int k, double out[4]) {
    if (!shell) {
        for (int l = 0; l < 4; ++l)
            out[l] = g.vol * (g.G[k][0] * g.G[l][0] + g.G[k][1] * g.G[l][1] + g.G[k][2] * g.G[l][2]);
        return;
    }
    for (int l = 0; l < 4; ++l) out[l] = 0.0;
    for (int q = 0; q < 4; ++q) {
        double xq[3] = {0, 0, 0};
        for (int v = 0; v < 4; ++v) {
            const double lam = (v == q) ? kKeastA : kKeastB;
            for (int a = 0; a < 3; ++a) xq[a] += lam * X[3 * (int64_t)t[v] + a];
        }
        double kq[3][3];
        t3d(xq, d.shell_R, d.shell_Rinf, kq);
        double kg[3];
        for (int a = 0; a < 3; ++a)
            kg[a] = kq[a][0] * g.G[k][0] + kq[a][1] * g.G[k][1] + kq[a][2] * g.G[k][2];
        for (int l = 0; l < 4; ++l)
            out[l] += (g.vol / 4.0) * (kg[0] * g.G[l][0] + kg[1] * g.G[l][1] + kg[2] * g.G[l][2]);
    }
}

}  // namespace

std::string assemble(const mm_mesh_desc &d, double drop_tol, HostOps &o, bool want_elements) {
    const int N = d.n_nodes, NE = d.n_elems;
    o.N = N;
    o.NE = NE;
    const double *X = d.coords;
    const int *T = d.tets;
    const double size = d.length_scale > 0 ? d.length_scale : 1.0e-9;
    for (int64_t i = 0; i < 4 * (int64_t)NE; ++i)
        if (T[i] < 0 || T[i] >= N) return "element node index out of range";
    for (int e = 0; e < NE; ++e)
        if (d.elem_mat[e] < 0 || d.elem_mat[e] >= d.n_mats) return "element material index out of range";

    // Volume V (scripts/materials.py:36-37), a compensated (Neumaier) sum in element
    // order: a plain running sum of the 11.7 M equal volumes of a 125^3-voxel magnet is off
    // by 1.7e-10 relative, which every C and G entry (scaled by 1 / V) would carry.
    double V = 0.0, Vc = 0.0;
    std::vector<uint8_t> magnetic(NE);
    for (int e = 0; e < NE; ++e) {
        Geo g;
        if (!geometry(X, T + 4 * (int64_t)e, g)) return "degenerate element " + std::to_string(e);
        magnetic[e] = d.mat_Js[d.elem_mat[e]] > 0.0;
        if (magnetic[e]) {
            const double t = V + g.vol;
            Vc += std::fabs(V) >= std::fabs(g.vol) ? (V - t) + g.vol : (g.vol - t) + V;
            V = t;
        }
    }
    V += Vc;
    if (!(V > 0.0)) return "no magnetic element (all Js <= 0)";
    o.V = V;

    // node -> (element, local index) incidence
    std::vector<int> inc_ptr(N + 1, 0), inc(4 * (size_t)NE);
    for (int64_t i = 0; i < 4 * (int64_t)NE; ++i) inc_ptr[T[i] + 1]++;
    for (int i = 0; i < N; ++i) inc_ptr[i + 1] += inc_ptr[i];
    {
        std::vector<int> fill(inc_ptr.begin(), inc_ptr.end() - 1);
        for (int e = 0; e < NE; ++e)
            for (int k = 0; k < 4; ++k) inc[fill[T[4 * (int64_t)e + k]]++] = 4 * e + k;
    }

    std::vector<double> cex(d.n_mats), Js(d.n_mats);
    for (int m = 0; m < d.n_mats; ++m) {
        Js[m] = d.mat_Js[m] > 0 ? d.mat_Js[m] : 0.0;
        const double Aeff = d.mat_Js[m] > 0 ? kMu0 * d.mat_A[m] / (size * size) : 0.0;
        cex[m] = 2.0 * Aeff / V;
    }

    o.meas.assign(N, 0.0);
    o.Dex.assign(3 * (size_t)N, 0.0);
    const uint8_t *dir = d.dirichlet;
    std::vector<int> rowlen(N, 0);

    // pass 1: meas, Dex, row lengths
#pragma omp parallel
    {
        std::vector<int> cols;
#pragma omp for schedule(dynamic, 1024)
        for (int i = 0; i < N; ++i) {
            double meas = 0.0, dex = 0.0;
            cols.clear();
            for (int q = inc_ptr[i]; q < inc_ptr[i + 1]; ++q) {
                const int e = inc[q] >> 2, k = inc[q] & 3;
                const int m = d.elem_mat[e];
                if (magnetic[e]) {
                    Geo g;
                    geometry(X, T + 4 * (int64_t)e, g);
                    meas += Js[m] * g.vol / 4.0;
                    dex += cex[m] * g.vol *
                           (g.G[k][0] * g.G[k][0] + g.G[k][1] * g.G[k][1] + g.G[k][2] * g.G[k][2]);
                }
                if (!(dir && dir[i]))
                    for (int l = 0; l < 4; ++l) {
                        const int j = T[4 * (int64_t)e + l];
                        if (!(dir && dir[j])) cols.push_back(j);
                    }
            }
            o.meas[i] = meas;
            for (int a = 0; a < 3; ++a) o.Dex[3 * (size_t)i + a] = meas == 0.0 ? 1.0 : dex;
            if (dir && dir[i]) {
                rowlen[i] = 1;
            } else {
                std::sort(cols.begin(), cols.end());
                rowlen[i] = (int)(std::unique(cols.begin(), cols.end()) - cols.begin());
            }
        }
    }
    std::vector<int> rp(N + 1, 0);
    for (int i = 0; i < N; ++i) rp[i + 1] = rp[i] + rowlen[i];
    std::vector<int> col(rp[N]);
    std::vector<double> val(rp[N], 0.0);
    o.diagK.assign(N, 0.0);

    // pass 2: values
#pragma omp parallel
    {
        std::vector<int> cols;
#pragma omp for schedule(dynamic, 1024)
        for (int i = 0; i < N; ++i) {
            const int b = rp[i];
            if (dir && dir[i]) {
                col[b] = i;
                val[b] = 1.0;
                o.diagK[i] = 1.0;
                continue;
            }
            cols.clear();
            for (int q = inc_ptr[i]; q < inc_ptr[i + 1]; ++q) {
                const int e = inc[q] >> 2;
                for (int l = 0; l < 4; ++l) {
                    const int j = T[4 * (int64_t)e + l];
                    if (!(dir && dir[j])) cols.push_back(j);
                }
            }
            std::sort(cols.begin(), cols.end());
            cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
            std::copy(cols.begin(), cols.end(), col.begin() + b);
            for (int q = inc_ptr[i]; q < inc_ptr[i + 1]; ++q) {
                const int e = inc[q] >> 2, k = inc[q] & 3;
                const int *t = T + 4 * (int64_t)e;
                Geo g;
                geometry(X, t, g);
                double row[4];
                stiff_row(d, X, t, d.shell && d.shell[e], g, k, row);
                for (int l = 0; l < 4; ++l) {
                    const int j = t[l];
                    if (dir && dir[j]) continue;
                    const int pos = (int)(std::lower_bound(cols.begin(), cols.end(), j) - cols.begin());
                    val[b + pos] += row[l];
                }
            }
            const int pd = (int)(std::lower_bound(cols.begin(), cols.end(), i) - cols.begin());
            o.diagK[i] = val[b + pd];
        }
    }

    // drop numerically zero couplings (|K_ij| < drop_tol max(|K_ii|, |K_jj|))
    o.k_rowptr.assign(N + 1, 0);
    o.k_col.clear();
    o.k_val.clear();
    o.k_col.reserve(col.size());
    o.k_val.reserve(val.size());
    for (int i = 0; i < N; ++i) {
        for (int q = rp[i]; q < rp[i + 1]; ++q) {
            const int j = col[q];
            if (j != i && drop_tol > 0.0 &&
                std::fabs(val[q]) < drop_tol * std::max(std::fabs(o.diagK[i]), std::fabs(o.diagK[j])))
                continue;
            o.k_col.push_back(j);
            o.k_val.push_back(val[q]);
        }
        o.k_rowptr[i + 1] = (int)o.k_col.size();
    }

    o.n_mag_nodes = 0;
    for (int i = 0; i < N; ++i) o.n_mag_nodes += o.meas[i] > 0.0;

    o.mag_elems.clear();
    for (int e = 0; e < NE; ++e)
        if (magnetic[e]) o.mag_elems.push_back(e);
    if (want_elements) {
        const int nm = (int)o.mag_elems.size();
        o.mag_G.assign(9 * (size_t)nm, 0.0);
        o.mag_vol.assign(nm, 0.0);
        o.mag_mat.assign(nm, 0);
        o.mag_tets.assign(4 * (size_t)nm, 0);
#pragma omp parallel for schedule(static)
        for (int q = 0; q < nm; ++q) {
            const int e = o.mag_elems[q];
            Geo g;
            geometry(X, T + 4 * (int64_t)e, g);
            for (int k = 1; k < 4; ++k)
                for (int a = 0; a < 3; ++a) o.mag_G[(size_t)(3 * (k - 1) + a) * nm + q] = g.G[k][a];
            o.mag_vol[q] = g.vol;
            o.mag_mat[q] = d.elem_mat[e];
            for (int k = 0; k < 4; ++k) o.mag_tets[4 * (size_t)q + k] = T[4 * (int64_t)e + k];
        }
    }
    return "";
}

void csr_to_sell(int n, const std::vector<int> &rowptr, const std::vector<int> &col,
                 const std::vector<double> &val, std::vector<int> &slice_ptr,
                 std::vector<int> &scol, std::vector<double> &sval) {
    const int C = 64;
    const int ns = (n + C - 1) / C;
    slice_ptr.assign(ns + 1, 0);
    for (int s = 0; s < ns; ++s) {
        int w = 0;
        for (int r = s * C; r < std::min(n, (s + 1) * C); ++r) w = std::max(w, rowptr[r + 1] - rowptr[r]);
        slice_ptr[s + 1] = slice_ptr[s] + w * C;
    }
    scol.assign(slice_ptr[ns], 0);
    sval.assign(slice_ptr[ns], 0.0);
#pragma omp parallel for schedule(static)
    for (int s = 0; s < ns; ++s) {
        const int w = (slice_ptr[s + 1] - slice_ptr[s]) / C;
        for (int lane = 0; lane < C; ++lane) {
            const int r = s * C + lane;
            const int len = r < n ? rowptr[r + 1] - rowptr[r] : 0;
            const int pad_col = r < n ? r : 0;
            for (int j = 0; j < w; ++j) {
                const size_t idx = slice_ptr[s] + (size_t)j * C + lane;
                if (j < len) {
                    scol[idx] = col[rowptr[r] + j];
                    sval[idx] = val[rowptr[r] + j];
                } else {
                    scol[idx] = pad_col;
                    sval[idx] = 0.0;
                }
            }
        }
    }
}

}  // namespace mm
