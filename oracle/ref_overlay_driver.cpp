// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A thin extern "C" driver around FlowCraft's own FieldOverlay::ComputeAerodynamicForces
// (FlowCraft/FieldOverlay.{h,cpp}), compiled from the reference sources where they lie
// by oracle/Makefile into oracle/_ref/libref_overlay.so.  It pins the restatement in
// lbm_oracle.cpp (AeroForces) to the real function.  No reference source is copied
// here: this file only includes the reference header and calls its public API
// (FieldOverlay.h:186-196).

#include <cstdint>
#include <vector>

#include "FieldOverlay.h"
#include "VoxelGrid.h"

extern "C" {

// field: 4 floats per cell (ReadbackField layout).  out[11] = forceX/Y/Z, minStress,
// maxStress, centerOfPressure (3), liftArrowBase (3).
void ref_aero_forces(int nx, int ny, int nz, float cell_size, const float* origin, const uint8_t* solid,
                     const float* field, float* out, int* has_nan) {
    VoxelGrid g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.cellSize = cell_size;
    g.origin = Vec3{origin[0], origin[1], origin[2]};
    g.solid.assign(solid, solid + g.CellCount());
    std::vector<float> f(field, field + g.CellCount() * 4);
    const FieldOverlay::AerodynamicForces a = FieldOverlay::ComputeAerodynamicForces(g, f);
    const float v[11] = {a.forceX, a.forceY, a.forceZ, a.minStress, a.maxStress,
                         a.centerOfPressure.x, a.centerOfPressure.y, a.centerOfPressure.z,
                         a.liftArrowBase.x, a.liftArrowBase.y, a.liftArrowBase.z};
    for (int i = 0; i < 11; ++i) out[i] = v[i];
    *has_nan = a.hasNaN ? 1 : 0;
}

}  // extern "C"
