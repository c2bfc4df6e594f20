// Host-side P1 assembly of the hot-path operators (runs once per context).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mumag_hip.h"

namespace mm {

struct HostOps {
    int N = 0, NE = 0;
    double V = 0.0;
    std::vector<double> meas;     // [N]
    std::vector<double> Dex;      // [3N] exchange diagonal, 1 where meas == 0
    std::vector<double> diagK;    // [N]
    // K in CSR (Dirichlet rows/cols replaced by identity, tiny entries dropped)
    std::vector<int> k_rowptr, k_col;
    std::vector<double> k_val;
    // magnetic elements (Js > 0)
    std::vector<int> mag_elems;   // indices into the element list
    std::vector<double> mag_G;    // SoA [9][n_mag]
    std::vector<double> mag_vol;  // [n_mag]
    std::vector<int> mag_mat;     // [n_mag]
    std::vector<int> mag_tets;    // [n_mag * 4]
    int n_mag_nodes = 0;
};

// Returns an empty string on success, else an error message.
std::string assemble(const mm_mesh_desc &d, double drop_tol, HostOps &out, bool want_elements);

// SELL-64 layout of a CSR matrix.
void csr_to_sell(int n, const std::vector<int> &rowptr, const std::vector<int> &col,
                 const std::vector<double> &val, std::vector<int> &slice_ptr,
                 std::vector<int> &scol, std::vector<double> &sval);

}  // namespace mm
