// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A thin extern "C" driver around FlowCraft's own mesh voxelizer, Voxelizer::Voxelize
// (FlowCraft/Voxelizer.{h,cpp}), compiled from the reference sources where they lie
// by oracle/Makefile into oracle/_ref/libref_voxelizer.so.  It pins the restatement in
// vox_oracle.cpp and serves as the CPU baseline of the device voxelizer.  No reference
// source is copied here: this file only includes the reference header and calls its
// public API (Voxelizer.h:11-53).

#include <cstdint>
#include <cstring>
#include <vector>

#include "Voxelizer.h"

namespace {
struct RefVox {
    VoxelGrid grid;
};
}  // namespace

extern "C" {

// Runs Voxelize on positions (3 floats per vertex) / indices; returns a handle whose
// grid is read with ref_vox_info / ref_vox_copy and freed with ref_vox_free.
void* ref_voxelize(const float* pos, long long nv, const uint32_t* idx, long long ni, int target, const float* pad) {
    std::vector<Vec3> p(static_cast<size_t>(nv));
    for (long long i = 0; i < nv; ++i) p[i] = Vec3{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
    std::vector<uint32_t> ix(idx, idx + ni);
    Voxelizer::Params prm;
    prm.targetResolution = target;
    prm.paddingFactor = Vec3{pad[0], pad[1], pad[2]};
    RefVox* r = new RefVox();
    r->grid = Voxelizer::Voxelize(p, ix, prm);
    return r;
}

// dims[3] = nx, ny, nz; frame[4] = cellSize, origin.
void ref_vox_info(void* h, int* dims, float* frame) {
    const VoxelGrid& g = static_cast<RefVox*>(h)->grid;
    dims[0] = g.nx;
    dims[1] = g.ny;
    dims[2] = g.nz;
    frame[0] = g.cellSize;
    frame[1] = g.origin.x;
    frame[2] = g.origin.y;
    frame[3] = g.origin.z;
}

void ref_vox_copy(void* h, uint8_t* solid, float* sdf) {
    const VoxelGrid& g = static_cast<RefVox*>(h)->grid;
    if (!g.solid.empty()) std::memcpy(solid, g.solid.data(), g.solid.size());
    if (!g.signedDistance.empty()) std::memcpy(sdf, g.signedDistance.data(), g.signedDistance.size() * sizeof(float));
}

void ref_vox_free(void* h) { delete static_cast<RefVox*>(h); }

}  // extern "C"
