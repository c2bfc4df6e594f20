// Host set-up of structured Kuhn-grid meshes (structured.cpp).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mumag_hip.h"

namespace mm {

struct StructHost {
    int nx = 0, ny = 0, nz = 0;                 // nodes per axis
    std::vector<double> gx, gy, gz;             // node coordinates
    std::vector<double> hx, hy, hz;             // voxel sizes
    int m0x = 0, m0y = 0, m0z = 0, mn = 0;      // magnet block (voxels)
    double h = 0.0;                             // magnet voxel edge
    std::vector<uint16_t> vmat;                 // [mn^3] material per magnet voxel
    double V = 0.0;
    std::vector<double> meas, Dex, diagK;
    int64_t n_mag_nodes = 0;
};

// Validates the description against the mesh arrays; returns "" or the reason.
std::string structured_setup(const mm_mesh_desc &d, StructHost &o);

}  // namespace mm
