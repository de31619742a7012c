// Drop-in replacement for FlowCraft's LbmSolver class (FlowCraft/LbmSolver.h:59-302),
// header-only over the C ABI of libfc_lbm.so (fc_lbm.h).
//
// Same public methods, argument meaning and error behaviour as the reference:
// Init() returns false and sets LastError(); every other method is a silent no-op
// on an invalid solver; ResetFromField() ignores a field of the wrong size; Step()/
// StepRange() only enqueue work (on the solver's HIP stream — the reference records
// into the caller's command buffer), readbacks block.  The SDL handle parameters
// are accepted and ignored, so reference call sites compile unchanged.
//
// LatticeModel and VoxelGrid come from the including application's own headers
// (FlowCraft/LatticeModel.h, FlowCraft/VoxelGrid.h); include/flowcraft/ carries
// minimal definitions with the same names and members for standalone builds.
#pragma once

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <string>
#include <vector>

#include "LatticeModel.h"
#include "VoxelGrid.h"
#include "fc_lbm.h"

struct SDL_GPUDevice;
struct SDL_GPUCommandBuffer;

class LbmSolver {
public:
    // LbmSolver::Params (FlowCraft/LbmSolver.h:61-100), same fields and defaults.
    struct Params {
        float tau = 0.8f;
        float inletVelocity = 0.05f;
        bool reverseFlow = false;
        bool freeSlipY = true;
        bool freeSlipZ = true;
        float smagorinskyCs = 0.12f;
    };

    // LbmSolver::Stats (FlowCraft/LbmSolver.h:172-184).  centerOfPressure/liftBaseZ
    // are FieldOverlay's, not the solver's: left zero here as in the reference's
    // ReadbackStats (FlowCraft/LbmSolver.cpp:739-788).
    struct Stats {
        float avgDensity = 0.0f, minDensity = 0.0f, maxDensity = 0.0f;
        float avgSpeed = 0.0f, maxSpeed = 0.0f;
        float forceX = 0.0f, forceY = 0.0f, forceZ = 0.0f;
        float maxStress = 0.0f;
        Vec3 centerOfPressure{0.0f, 0.0f, 0.0f};
        float liftBaseZ = 0.0f;
        bool hasNaN = false;
        bool valid = false;
        double totalMass = 0.0;  // extension (LbmReference::Stats::totalMass)
    };

    // Engine options with no reference counterpart (population storage and layout,
    // periodic y/z); Init() without them gives the FP32 ping-pong engine, the faster layout
    // (37.3 vs 32.1 GLUPS on a 512x256x256 lattice); FC_LBM_LAYOUT_INPLACE halves the
    // population memory, for lattices whose second copy would not fit.
    struct EngineOptions {
        int storage = FC_LBM_STORE_F32;
        int layout = FC_LBM_LAYOUT_PINGPONG;
        bool periodicY = false;
        bool periodicZ = false;
        void* stream = nullptr;  // hipStream_t; nullptr = the engine's own
    };

    LbmSolver() = default;
    LbmSolver(const LbmSolver&) = delete;
    LbmSolver& operator=(const LbmSolver&) = delete;
    ~LbmSolver() { Shutdown(); }

    bool Init(SDL_GPUDevice* /*device*/, const VoxelGrid& grid, LatticeModel model) {
        return Init(nullptr, grid, model, EngineOptions{});
    }

    bool Init(SDL_GPUDevice* /*device*/, const VoxelGrid& grid, LatticeModel model, const EngineOptions& opt) {
        Shutdown();
        if (grid.nx <= 0 || grid.ny <= 0 || grid.nz <= 0 || grid.solid.size() != grid.CellCount()) {
            lastError_ = "empty or inconsistent voxel grid";
            return false;
        }
        fc_lbm_desc d{};
        d.nx = grid.nx;
        d.ny = grid.ny;
        d.nz = grid.nz;
        d.model = model == LatticeModel::D3Q15 ? FC_LBM_D3Q15 : FC_LBM_D3Q19;
        d.storage = opt.storage;
        d.layout = opt.layout;
        d.periodic_y = opt.periodicY ? 1 : 0;
        d.periodic_z = opt.periodicZ ? 1 : 0;
        d.stream = opt.stream;
        fc_lbm* h = nullptr;
        if (fc_lbm_create(&d, &h) != FC_LBM_OK) {
            lastError_ = fc_lbm_last_error(h);
            if (h) fc_lbm_destroy(h);
            return false;
        }
        const float* sdf = grid.signedDistance.size() == grid.CellCount() ? grid.signedDistance.data() : nullptr;
        if (fc_lbm_set_geometry(h, grid.solid.data(), sdf) != FC_LBM_OK) {
            lastError_ = fc_lbm_last_error(h);
            fc_lbm_destroy(h);
            return false;
        }
        h_ = h;
        nx_ = grid.nx;
        ny_ = grid.ny;
        nz_ = grid.nz;
        model_ = model;
        solidCpu_ = grid.solid;
        lastError_.clear();
        return true;
    }

    void Shutdown() {
        if (h_) fc_lbm_destroy(h_);
        h_ = nullptr;
    }

    const std::string& LastError() const { return lastError_; }

    void Reset(const Params& params, bool seedAtMidSpeed = false) {
        params_ = params;
        if (!h_) return;
        PushParams();
        Check(fc_lbm_reset_uniform(h_, seedAtMidSpeed ? 1 : 0));
    }

    void ResetFromField(const std::vector<float>& field) {
        if (!h_ || field.size() != CellCount() * 4) return;
        Check(fc_lbm_reset_from_field(h_, field.data()));
    }

    void SetParams(const Params& params) {
        params_ = params;
        if (h_) PushParams();
    }

    void Step(SDL_GPUCommandBuffer* /*cmd*/, int numSubsteps) {
        if (h_) Check(fc_lbm_step(h_, numSubsteps));
    }

    void StepRange(SDL_GPUCommandBuffer* /*cmd*/, int zStart, int zCount) {
        if (h_) Check(fc_lbm_step_range(h_, zStart, zCount));
    }

    void AdvanceSubstep() {
        if (h_) Check(fc_lbm_advance(h_));
    }

    Stats ReadbackStats() {
        Stats s;
        if (!h_) return s;
        fc_lbm_stats r{};
        if (!Check(fc_lbm_readback_stats(h_, &r))) return s;
        s.avgDensity = static_cast<float>(r.avg_density);
        s.minDensity = static_cast<float>(r.min_density);
        s.maxDensity = static_cast<float>(r.max_density);
        s.avgSpeed = static_cast<float>(r.avg_speed);
        s.maxSpeed = static_cast<float>(r.max_speed);
        s.forceX = static_cast<float>(r.force_x);
        s.forceY = static_cast<float>(r.force_y);
        s.forceZ = static_cast<float>(r.force_z);
        s.maxStress = static_cast<float>(r.max_stress);
        s.hasNaN = r.has_nan != 0;
        s.totalMass = r.total_mass;
        s.valid = true;
        return s;
    }

    std::vector<float> ReadbackField() {
        std::vector<float> field;
        if (!h_) return field;
        field.resize(CellCount() * 4);
        Check(fc_lbm_readback_field(h_, field.data()));
        return field;
    }

    void ReadbackFieldRange(int zStart, int zCount, std::vector<float>& outField) {
        if (!h_ || outField.size() != CellCount() * 4) return;
        Check(fc_lbm_readback_field_range(h_, zStart, zCount, outField.data()));
    }

    // Extension: FieldOverlay::ComputeAerodynamicForces (FlowCraft/FieldOverlay.cpp:469-621)
    // of the current state, on the device; the reference computes it on the host from
    // ReadbackField().  Members as FieldOverlay::AerodynamicForces
    // (FlowCraft/FieldOverlay.h:186-195), world frame from the grid's cellSize/origin.
    struct AerodynamicForces {
        float forceX = 0.0f, forceY = 0.0f, forceZ = 0.0f;
        float minStress = 0.0f, maxStress = 0.0f;
        Vec3 centerOfPressure{0.0f, 0.0f, 0.0f};
        Vec3 liftArrowBase{0.0f, 0.0f, 0.0f};
        bool hasNaN = false;
    };

    AerodynamicForces ComputeAerodynamicForces(const VoxelGrid& grid) {
        AerodynamicForces a;
        if (!h_) return a;
        const float origin[3] = {grid.origin.x, grid.origin.y, grid.origin.z};
        fc_lbm_aero r{};
        if (!Check(fc_lbm_aero_forces(h_, grid.cellSize, origin, &r))) return a;
        a.forceX = r.force_x;
        a.forceY = r.force_y;
        a.forceZ = r.force_z;
        a.minStress = r.min_stress;
        a.maxStress = r.max_stress;
        a.centerOfPressure = Vec3{r.center_of_pressure[0], r.center_of_pressure[1], r.center_of_pressure[2]};
        a.liftArrowBase = Vec3{r.lift_arrow_base[0], r.lift_arrow_base[1], r.lift_arrow_base[2]};
        a.hasNaN = r.has_nan != 0;
        return a;
    }

    bool ProbeSliceHealth(int zLayer) {
        if (!h_ || nz_ <= 0) return false;
        int diverged = 0;
        if (!Check(fc_lbm_probe_health(h_, zLayer, &diverged))) return false;
        return diverged != 0;
    }

    const std::vector<uint8_t>& SolidMask() const { return solidCpu_; }
    bool IsValid() const { return h_ != nullptr; }
    LatticeModel Model() const { return model_; }
    int Nx() const { return nx_; }
    int Ny() const { return ny_; }
    int Nz() const { return nz_; }

    // Extension: FlowCraft's StepSolverPaced (FlowCraft/main.cpp:2372-2486) on this solver.
    // Each substep runs as z-chunks (StepRange) of adaptive depth, waiting for each chunk:
    // the depth halves when a chunk took longer than targetChunkMs and doubles when it took
    // under a quarter of it, within [4, 256] layers; AdvanceSubstep closes the substep, and
    // at most every probeIntervalMs a slice health probe runs (layers nz/2, nz/3, 2nz/3 in
    // turn).  Returns the substeps completed: fewer when `cancel` turns true between chunks
    // (FlowCraft's solver-epoch check), or i+1 after substep i's probe found a divergence
    // (state.diverged set).  `progress`, if given, receives (i + z/nz) / substeps.
    struct PacingState {
        int chunkZDepth = 32;
        int probePhase = 0;
        std::chrono::steady_clock::time_point lastProbe{};
        bool diverged = false;
    };

    int StepPaced(int substeps, PacingState& st, const std::atomic<bool>* cancel = nullptr,
                  double targetChunkMs = 3.0, double probeIntervalMs = 400.0, std::atomic<float>* progress = nullptr) {
        if (!h_) return 0;
        using Clock = std::chrono::steady_clock;
        const int nz = nz_;
        // A ping-pong substep writes the other copy, so chunks already run are dropped
        // by not advancing; in place they overwrite the state, so a started substep is
        // finished and a cancel takes effect at the next substep boundary.
        int nx_i, ny_i, nz_i, q_i, st_i, layout = FC_LBM_LAYOUT_INPLACE;
        fc_lbm_info(h_, &nx_i, &ny_i, &nz_i, &q_i, &st_i, &layout);
        const bool undoable = layout == FC_LBM_LAYOUT_PINGPONG;
        auto cancelled = [&]() { return cancel && cancel->load(std::memory_order_acquire); };
        if (progress) progress->store(0.0f, std::memory_order_relaxed);
        for (int i = 0; i < substeps; ++i) {
            for (int z0 = 0; z0 < nz;) {
                if ((z0 == 0 || undoable) && cancelled()) return i;
                const int zc = std::min(st.chunkZDepth, nz - z0);
                const Clock::time_point t0 = Clock::now();
                if (!Check(fc_lbm_step_range(h_, z0, zc)) || !Check(fc_lbm_sync(h_))) return i;
                const double ms = std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
                if (ms > targetChunkMs && st.chunkZDepth > 4)
                    st.chunkZDepth = std::max(4, st.chunkZDepth / 2);
                else if (ms < targetChunkMs * 0.25 && st.chunkZDepth < 256)
                    st.chunkZDepth = std::min(256, st.chunkZDepth * 2);
                z0 += zc;  // what this chunk covered, not the (possibly changed) depth
                if (progress)
                    progress->store((static_cast<float>(i) + static_cast<float>(z0) / static_cast<float>(nz)) /
                                        static_cast<float>(substeps),
                                    std::memory_order_relaxed);
            }
            if (undoable && cancelled()) return i;
            if (!Check(fc_lbm_advance(h_))) return i;
            const Clock::time_point now = Clock::now();
            if (std::chrono::duration<double, std::milli>(now - st.lastProbe).count() >= probeIntervalMs && nz > 0) {
                st.lastProbe = now;
                const int pz = st.probePhase == 0 ? nz / 2 : st.probePhase == 1 ? nz / 3 : (2 * nz) / 3;
                st.probePhase = (st.probePhase + 1) % 3;
                if (ProbeSliceHealth(pz)) {
                    st.diverged = true;
                    return i + 1;
                }
            }
        }
        return substeps;
    }

    // Extensions.
    fc_lbm* Handle() const { return h_; }
    void Sync() {
        if (h_) Check(fc_lbm_sync(h_));
    }

private:
    size_t CellCount() const { return static_cast<size_t>(nx_) * ny_ * nz_; }

    void PushParams() {
        fc_lbm_params p{};
        p.tau = params_.tau;
        p.inlet_velocity = params_.inletVelocity;
        p.reverse_flow = params_.reverseFlow ? 1 : 0;
        p.free_slip_y = params_.freeSlipY ? 1 : 0;
        p.free_slip_z = params_.freeSlipZ ? 1 : 0;
        p.smagorinsky_cs = params_.smagorinskyCs;
        Check(fc_lbm_set_params(h_, &p));
    }

    // Device errors have no reference counterpart (SDL_GPU calls there return
    // nothing); they are kept in LastError() for the caller to inspect.
    bool Check(int rc) {
        if (rc == FC_LBM_OK) return true;
        lastError_ = fc_lbm_last_error(h_);
        return false;
    }

    fc_lbm* h_ = nullptr;
    int nx_ = 0, ny_ = 0, nz_ = 0;
    LatticeModel model_ = LatticeModel::D3Q19;
    Params params_;
    std::vector<uint8_t> solidCpu_;
    std::string lastError_;
};
