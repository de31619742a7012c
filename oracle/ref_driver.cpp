// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A thin extern "C" driver around FlowCraft's own LbmReference class
// (FlowCraft/LbmReference.{h,cpp}), compiled from the reference sources where
// they lie by oracle/Makefile into oracle/_ref/libref_lbm.so.  It lets tests
// pin the restatement in lbm_oracle.cpp to the real reference, and write the
// golden fixtures under tests/golden/.  No reference source is copied here:
// this file only includes the reference header and calls its public API
// (LbmReference.h:39-181).

#include <cstdint>
#include <cstring>
#include <vector>

#include "LbmReference.h"
#include "VoxelGrid.h"

namespace {
struct RefHandle {
    VoxelGrid grid;
    LbmReference solver;
    LbmReference::Params params;
    LatticeModel model = LatticeModel::D3Q19;
};
}  // namespace

extern "C" {

// q: 15 or 19.  sdf may be null (VoxelGrid::signedDistance left empty).
void* ref_create(int nx, int ny, int nz, int q, const uint8_t* solid, const float* sdf, float tau, float u,
                 int reverse, int fsy, int fsz, float cs, int seed_mid) {
    RefHandle* h = new RefHandle();
    h->grid.nx = nx;
    h->grid.ny = ny;
    h->grid.nz = nz;
    h->grid.solid.assign(solid, solid + h->grid.CellCount());
    if (sdf) h->grid.signedDistance.assign(sdf, sdf + h->grid.CellCount());
    h->params.tau = tau;
    h->params.inletVelocity = u;
    h->params.reverseFlow = reverse != 0;
    h->params.freeSlipY = fsy != 0;
    h->params.freeSlipZ = fsz != 0;
    h->params.smagorinskyCs = cs;
    h->model = q == 15 ? LatticeModel::D3Q15 : LatticeModel::D3Q19;
    h->solver.Init(h->grid, h->model, h->params, seed_mid != 0);
    return h;
}

void ref_destroy(void* p) { delete static_cast<RefHandle*>(p); }

void ref_set_params(void* p, float tau, float u, int reverse, int fsy, int fsz, float cs) {
    RefHandle* h = static_cast<RefHandle*>(p);
    h->params.tau = tau;
    h->params.inletVelocity = u;
    h->params.reverseFlow = reverse != 0;
    h->params.freeSlipY = fsy != 0;
    h->params.freeSlipZ = fsz != 0;
    h->params.smagorinskyCs = cs;
    h->solver.SetParams(h->params);
}

void ref_reset(void* p, int seed_mid) { static_cast<RefHandle*>(p)->solver.Reset(seed_mid != 0); }

void ref_reset_from_field(void* p, const float* field) {
    RefHandle* h = static_cast<RefHandle*>(p);
    std::vector<float> f(field, field + h->grid.CellCount() * 4);
    h->solver.ResetFromField(f);
}

void ref_step(void* p, int n) {
    RefHandle* h = static_cast<RefHandle*>(p);
    for (int i = 0; i < n; ++i) h->solver.Step();
}

void ref_readback_field(void* p, float* out) {
    RefHandle* h = static_cast<RefHandle*>(p);
    std::vector<float> f = h->solver.ReadbackField();
    std::memcpy(out, f.data(), f.size() * sizeof(float));
}

// out[11]: avgDensity, minDensity, maxDensity, avgSpeed, maxSpeed, forceX, forceY,
// forceZ, maxStress, totalMass, hasNaN
void ref_stats(void* p, double* out) {
    LbmReference::Stats s = static_cast<RefHandle*>(p)->solver.ComputeStats();
    out[0] = s.avgDensity;
    out[1] = s.minDensity;
    out[2] = s.maxDensity;
    out[3] = s.avgSpeed;
    out[4] = s.maxSpeed;
    out[5] = s.forceX;
    out[6] = s.forceY;
    out[7] = s.forceZ;
    out[8] = s.maxStress;
    out[9] = s.totalMass;
    out[10] = s.hasNaN ? 1.0 : 0.0;
}

int ref_has_diverged(void* p) { return static_cast<RefHandle*>(p)->solver.HasDiverged() ? 1 : 0; }

// Times `steps` reference steps after `warmup` untimed ones; returns seconds.
double ref_time_steps(void* p, int warmup, int steps);
}

#include <chrono>
extern "C" double ref_time_steps(void* p, int warmup, int steps) {
    ref_step(p, warmup);
    auto t0 = std::chrono::steady_clock::now();
    ref_step(p, steps);
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}
