// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A thin extern "C" driver around FlowCraft's own SmokeTrailSystem
// (FlowCraft/SmokeTrail.{h,cpp}, with FieldOverlay.cpp which it calls), compiled from the
// reference sources where they lie by oracle/Makefile into oracle/_ref/libref_smoke.so.
// It pins the restatement in smoke_oracle.cpp.  No reference source is copied here.  The
// class keeps its particles private and offers no accessor, so this one translation unit
// opens the class's access specifiers to read them (the layout is unchanged).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <random>
#include <string>
#include <vector>

#include "FieldOverlay.h"
#include "Mesh.h"
#include "VoxelGrid.h"
// everything SmokeTrail.h includes is already in, so only its own class is opened
#define private public
#include "SmokeTrail.h"
#undef private

namespace {
struct RefSmoke {
    VoxelGrid grid;
    SmokeTrailSystem sys;
    std::vector<float> field;  // the ReadbackField() vector the reference steps from
    int L = 10;
};
SmokeTrailSystem::PlaneConfinement Plane(int axis, int index) {
    return SmokeTrailSystem::PlaneConfinement{axis == 0 ? SliceAxis::X : axis == 1 ? SliceAxis::Y : SliceAxis::Z,
                                              index};
}
}  // namespace

extern "C" {

void* ref_smoke_create(int nx, int ny, int nz, float cs, const float* origin, const uint8_t* solid) {
    RefSmoke* r = new RefSmoke();
    r->grid.nx = nx;
    r->grid.ny = ny;
    r->grid.nz = nz;
    r->grid.cellSize = cs;
    r->grid.origin = Vec3{origin[0], origin[1], origin[2]};
    r->grid.solid.assign(solid, solid + r->grid.CellCount());
    return r;
}
void ref_smoke_destroy(void* h) { delete static_cast<RefSmoke*>(h); }
void ref_smoke_reset(void* h, int reverse, int count, int trail, int axis, int index) {
    RefSmoke* r = static_cast<RefSmoke*>(h);
    r->L = trail < 2 ? 2 : trail;
    const SmokeTrailSystem::PlaneConfinement c = Plane(axis, index);
    r->sys.Reset(r->grid, reverse != 0, count, trail, axis >= 0 ? &c : nullptr);
}
void ref_smoke_scatter(void* h, int axis, int index) {
    RefSmoke* r = static_cast<RefSmoke*>(h);
    const SmokeTrailSystem::PlaneConfinement c = Plane(axis, index);
    r->sys.ScatterRandom(r->grid, axis >= 0 ? &c : nullptr);
}
void ref_smoke_shift_plane(void* h, int axis, int index) {
    RefSmoke* r = static_cast<RefSmoke*>(h);
    r->sys.ShiftPlane(r->grid, Plane(axis, index));
}
// The field as FlowCraft holds it (a std::vector filled by ReadbackField), then the step
// on it: split so that a timing of the step excludes this copy.
void ref_smoke_load_field(void* h, const float* field) {
    RefSmoke* r = static_cast<RefSmoke*>(h);
    r->field.assign(field, field + r->grid.CellCount() * 4);
}
void ref_smoke_step_loaded(void* h, int reverse, float mmps, float dt, int suspend, int axis, int index) {
    RefSmoke* r = static_cast<RefSmoke*>(h);
    const SmokeTrailSystem::PlaneConfinement c = Plane(axis, index);
    r->sys.Step(r->grid, r->field, reverse != 0, mmps, dt, suspend != 0, axis >= 0 ? &c : nullptr);
}
void ref_smoke_step(void* h, const float* field, int reverse, float mmps, float dt, int suspend, int axis, int index) {
    ref_smoke_load_field(h, field);
    ref_smoke_step_loaded(h, reverse, mmps, dt, suspend, axis, index);
}
void ref_smoke_trails(void* h, float* out, int* lengths) {
    RefSmoke* r = static_cast<RefSmoke*>(h);
    const auto& ps = r->sys.particles_;
    for (size_t p = 0; p < ps.size(); ++p) {
        lengths[p] = static_cast<int>(ps[p].trail.size());
        for (size_t k = 0; k < ps[p].trail.size(); ++k) {
            out[(p * r->L + k) * 3] = ps[p].trail[k].x;
            out[(p * r->L + k) * 3 + 1] = ps[p].trail[k].y;
            out[(p * r->L + k) * 3 + 2] = ps[p].trail[k].z;
        }
    }
}

}  // extern "C"
