// Host set-up of a structured (Kuhn-split rectilinear) mesh: validation of the
// description against the node and element lists, then meas, the exchange diagonal,
// diag(K) and V from the exact voxel tables (kuhn_tables.h) instead of a general
// assembly.  Semantics are those of assemble.cpp (scripts/mapping.py:171-236).
#include "structured.h"

#include <cmath>
#include <string>

#include "kuhn_tables.h"

namespace mm {

namespace {
constexpr double kMu0 = 4.0e-7 * M_PI;  // scripts/materials.py:21

// corner indices (c = cx + 2 cy + 4 cz) of the 6 Kuhn tetrahedra, generator order:
// axis orders (x,y,z), (y,z,x), (z,x,y), (x,z,y), (z,y,x), (y,x,z); the odd orders have
// nodes 1 and 2 swapped for positive orientation.
void kuhn_corners(int out[6][4]) {
    const int perms[6][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}, {0, 2, 1}, {2, 1, 0}, {1, 0, 2}};
    for (int t = 0; t < 6; ++t) {
        int c = 0;
        out[t][0] = 0;
        for (int s = 0; s < 3; ++s) {
            c |= 1 << perms[t][s];
            out[t][s + 1] = c;
        }
        if (t >= 3) std::swap(out[t][1], out[t][2]);
    }
}

bool close(double a, double b, double scale) { return std::fabs(a - b) <= 1e-12 * scale; }

// escript-equivalent edge coefficient of the 7-point stencil (see kernels.hip)
double edge_coef(const std::vector<double> &h1, const std::vector<double> &h2, int a, int b,
                 double inv6hd) {
    const int n1 = (int)h1.size(), n2 = (int)h2.size();
    double s = 0;
    const bool a0 = a - 1 >= 0, a1 = a < n1, b0 = b - 1 >= 0, b1 = b < n2;
    if (a0 && b0) s += 2.0 * h1[a - 1] * h2[b - 1];
    if (a0 && b1) s += h1[a - 1] * h2[b];
    if (a1 && b0) s += h1[a] * h2[b - 1];
    if (a1 && b1) s += 2.0 * h1[a] * h2[b];
    return -s * inv6hd;
}
}  // namespace

std::string structured_setup(const mm_mesh_desc &d, StructHost &o) {
    const int nvx = d.grid_n[0], nvy = d.grid_n[1], nvz = d.grid_n[2];
    if (nvx <= 0 || nvy <= 0 || nvz <= 0 || !d.grid_x || !d.grid_y || !d.grid_z)
        return "structured: grid_n / grid coordinates missing";
    const int nx = nvx + 1, ny = nvy + 1, nz = nvz + 1;
    if ((int64_t)nx * ny * nz != d.n_nodes) return "structured: node count != grid size";
    if ((int64_t)6 * nvx * nvy * nvz != d.n_elems) return "structured: element count != 6 x voxels";
    if (d.shell) return "structured: the shell transform is not a box grid";
    o.nx = nx;
    o.ny = ny;
    o.nz = nz;
    o.gx.assign(d.grid_x, d.grid_x + nx);
    o.gy.assign(d.grid_y, d.grid_y + ny);
    o.gz.assign(d.grid_z, d.grid_z + nz);
    double scale = 0;
    for (double v : o.gx) scale = std::max(scale, std::fabs(v));
    for (double v : o.gy) scale = std::max(scale, std::fabs(v));
    for (double v : o.gz) scale = std::max(scale, std::fabs(v));
    // node coordinates
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int64_t n = i + (int64_t)nx * (j + (int64_t)ny * k);
                if (!close(d.coords[3 * n], o.gx[i], scale) || !close(d.coords[3 * n + 1], o.gy[j], scale) ||
                    !close(d.coords[3 * n + 2], o.gz[k], scale))
                    return "structured: node " + std::to_string(n) + " is not at its grid position";
            }
    // elements: 6 Kuhn tetrahedra per voxel in generator order, one material per voxel
    int KC[6][4];
    kuhn_corners(KC);
    const int m0x = d.grid_mag0[0], m0y = d.grid_mag0[1], m0z = d.grid_mag0[2], mn = d.grid_magn;
    if (mn <= 0 || m0x < 0 || m0y < 0 || m0z < 0 || m0x + mn > nvx || m0y + mn > nvy || m0z + mn > nvz)
        return "structured: magnet block outside the grid";
    o.m0x = m0x;
    o.m0y = m0y;
    o.m0z = m0z;
    o.mn = mn;
    o.vmat.assign((size_t)mn * mn * mn, 0);
    if (d.n_mats >= 65535) return "structured: too many materials";
    for (int vz = 0; vz < nvz; ++vz)
        for (int vy = 0; vy < nvy; ++vy)
            for (int vx = 0; vx < nvx; ++vx) {
                const int64_t v = vx + (int64_t)nvx * (vy + (int64_t)nvy * vz);
                const int64_t base = vx + (int64_t)nx * (vy + (int64_t)ny * vz);
                const int mat = d.elem_mat[6 * v];
                for (int t = 0; t < 6; ++t) {
                    if (d.elem_mat[6 * v + t] != mat) return "structured: voxel with mixed materials";
                    for (int q = 0; q < 4; ++q) {
                        const int c = KC[t][q];
                        const int64_t want = base + (c & 1) + (int64_t)nx * (((c >> 1) & 1) + (int64_t)ny * ((c >> 2) & 1));
                        if (d.tets[4 * (6 * v + t) + q] != want)
                            return "structured: element " + std::to_string(6 * v + t) + " is not the Kuhn tetrahedron";
                    }
                }
                const bool mag = d.mat_Js[mat] > 0.0;
                const bool in = vx >= m0x && vx < m0x + mn && vy >= m0y && vy < m0y + mn && vz >= m0z &&
                                vz < m0z + mn;
                if (mag != in) return "structured: magnetic voxels are not exactly the magnet block";
                if (in) o.vmat[(vx - m0x) + (size_t)mn * ((vy - m0y) + (size_t)mn * (vz - m0z))] = (uint16_t)mat;
            }
    // uniform magnet spacing
    const double h = o.gx[m0x + 1] - o.gx[m0x];
    for (int s = 0; s < mn; ++s) {
        if (!close(o.gx[m0x + s + 1] - o.gx[m0x + s], h, scale) || !close(o.gy[m0y + s + 1] - o.gy[m0y + s], h, scale) ||
            !close(o.gz[m0z + s + 1] - o.gz[m0z + s], h, scale))
            return "structured: magnet voxels are not uniform cubes";
    }
    o.h = h;
    // Dirichlet = the six faces of the box (air-box branch of scripts/matrix.py:49-60)
    if (!d.dirichlet) return "structured: Dirichlet mask missing";
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int64_t n = i + (int64_t)nx * (j + (int64_t)ny * k);
                const bool face = i == 0 || j == 0 || k == 0 || i == nx - 1 || j == ny - 1 || k == nz - 1;
                if ((d.dirichlet[n] != 0) != face) return "structured: Dirichlet set is not the box faces";
            }
    // V, meas, Dex, diag K
    const double size = d.length_scale > 0 ? d.length_scale : 1.0e-9;
    const int64_t nmagv = (int64_t)mn * mn * mn;
    o.V = (double)nmagv * h * h * h;
    const int64_t N = d.n_nodes;
    o.meas.assign(N, 0.0);
    o.Dex.assign(3 * N, 1.0);
    std::vector<double> cex(d.n_mats, 0.0), Js(d.n_mats, 0.0);
    for (int m = 0; m < d.n_mats; ++m)
        if (d.mat_Js[m] > 0) {
            Js[m] = d.mat_Js[m];
            cex[m] = 2.0 * (kMu0 * d.mat_A[m] / (size * size)) / o.V;
        }
    int rowsumM[8];
    for (int p = 0; p < 8; ++p) {
        rowsumM[p] = 0;
        for (int c = 0; c < 8; ++c) rowsumM[p] += kKM[p][c];
    }
    const int n1 = mn + 1;
    o.n_mag_nodes = 0;
    for (int k = 0; k < n1; ++k)
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n1; ++i) {
                double meas = 0.0, dex = 0.0;
                for (int w = 0; w < 8; ++w) {
                    const int vx = w & 1, vy = (w >> 1) & 1, vz = (w >> 2) & 1;
                    const int xi = i - 1 + vx, yj = j - 1 + vy, zk = k - 1 + vz;
                    if (xi < 0 || yj < 0 || zk < 0 || xi >= mn || yj >= mn || zk >= mn) continue;
                    const int mat = o.vmat[xi + (size_t)mn * (yj + (size_t)mn * zk)];
                    const int p = (1 - vx) + 2 * (1 - vy) + 4 * (1 - vz);
                    meas += Js[mat] * h * h * h * rowsumM[p] / 120.0;
                    dex += cex[mat] * h * kKS[p][p] / 6.0;
                }
                const int64_t n = (m0x + i) + (int64_t)nx * ((m0y + j) + (int64_t)ny * (m0z + k));
                o.meas[n] = meas;
                if (meas != 0.0) {
                    for (int a = 0; a < 3; ++a) o.Dex[3 * n + a] = dex;
                    o.n_mag_nodes++;
                }
            }
    // voxel sizes and diag K
    o.hx.resize(nvx);
    o.hy.resize(nvy);
    o.hz.resize(nvz);
    for (int i = 0; i < nvx; ++i) o.hx[i] = o.gx[i + 1] - o.gx[i];
    for (int j = 0; j < nvy; ++j) o.hy[j] = o.gy[j + 1] - o.gy[j];
    for (int k = 0; k < nvz; ++k) o.hz[k] = o.gz[k + 1] - o.gz[k];
    o.diagK.assign(N, 1.0);
    for (int k = 1; k < nz - 1; ++k)
        for (int j = 1; j < ny - 1; ++j)
            for (int i = 1; i < nx - 1; ++i) {
                const double s = edge_coef(o.hy, o.hz, j, k, 1.0 / (6.0 * o.hx[i - 1])) +
                                 edge_coef(o.hy, o.hz, j, k, 1.0 / (6.0 * o.hx[i])) +
                                 edge_coef(o.hz, o.hx, k, i, 1.0 / (6.0 * o.hy[j - 1])) +
                                 edge_coef(o.hz, o.hx, k, i, 1.0 / (6.0 * o.hy[j])) +
                                 edge_coef(o.hx, o.hy, i, j, 1.0 / (6.0 * o.hz[k - 1])) +
                                 edge_coef(o.hx, o.hy, i, j, 1.0 / (6.0 * o.hz[k]));
                o.diagK[i + (int64_t)nx * (j + (int64_t)ny * k)] = -s;
            }
    return "";
}

}  // namespace mm
