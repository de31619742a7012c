// Drop-in for FlowCraft's Voxelizer::Voxelize (FlowCraft/Voxelizer.h:40-41) on the
// device, header-only over the C ABI of libfc_lbm.so (fc_voxelize, fc_lbm.h).
//
// Same inputs (world-space vertex positions, 3 indices per triangle), parameters
// (Voxelizer::Params: targetResolution, paddingFactor) and result (the VoxelGrid the
// solver reads, solid mask and signed distance identical to the reference's); an empty
// mesh gives an empty grid as there.  A device error leaves the grid empty and, when
// `error` is given, stores the message (the reference has no failure path).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "VoxelGrid.h"
#include "fc_lbm.h"

inline VoxelGrid VoxelizeOnDevice(const std::vector<Vec3>& positions, const std::vector<uint32_t>& indices,
                                  int targetResolution = 640, Vec3 paddingFactor = Vec3{0.3f, 1.0f, 0.5f},
                                  std::string* error = nullptr) {
    VoxelGrid grid;
    std::vector<float> pos(positions.size() * 3);
    for (size_t i = 0; i < positions.size(); ++i) {
        pos[3 * i] = positions[i].x;
        pos[3 * i + 1] = positions[i].y;
        pos[3 * i + 2] = positions[i].z;
    }
    fc_vox_params p{targetResolution, {paddingFactor.x, paddingFactor.y, paddingFactor.z}};
    fc_vox_grid f{};
    const long long nv = static_cast<long long>(positions.size()), ni = static_cast<long long>(indices.size());
    int rc = fc_voxelize_frame(pos.data(), nv, indices.data(), ni, &p, &f);
    if (rc == FC_LBM_OK && f.nx > 0) {
        grid.nx = f.nx;
        grid.ny = f.ny;
        grid.nz = f.nz;
        grid.cellSize = f.cell_size;
        grid.origin = Vec3{f.origin[0], f.origin[1], f.origin[2]};
        grid.solid.resize(grid.CellCount());
        grid.signedDistance.resize(grid.CellCount());
        rc = fc_voxelize(pos.data(), nv, indices.data(), ni, &p, grid.solid.data(), grid.signedDistance.data());
    }
    if (rc != FC_LBM_OK) {
        if (error) *error = fc_voxelize_last_error();
        return VoxelGrid{};
    }
    return grid;
}
