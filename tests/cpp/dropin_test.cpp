// Drop-in check of include/flowcraft/LbmSolver.h: the call sequence of FlowCraft's
// own tests/gpu_vs_cpu_test.cpp (:27-80 and :149-172) written against the
// replacement class, compared with the CPU oracle (oracle/liblbm_oracle.so, test
// infrastructure) instead of LbmReference.  Exit 0 = pass.
#include <algorithm>
#include <cmath>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <cstdint>
#include <vector>

#include "LbmSolver.h"
#include "VoxelizeDevice.h"

extern "C" {
void* orc_create(int nx, int ny, int nz, int q, int real_bits, int storage, int py, int pz, int nthreads);
void orc_destroy(void*);
void orc_set_geometry(void*, const uint8_t*, const float*);
void orc_set_params(void*, float, float, int, int, int, float);
void orc_reset(void*, int);
void orc_step(void*, int);
void orc_readback_field_f32(void*, float*);
void orc_aero_forces(int nx, int ny, int nz, float cs, const float* origin, const uint8_t* solid, const float* field,
                     float* out, int* has_nan);
int orc_vox_frame(const float* pos, long long nv, long long ni, int target, const float* pad, int* dims, float* frame);
void orc_voxelize(const float* pos, long long nv, const uint32_t* idx, long long ni, int target, const float* pad,
                  uint8_t* solid, float* sdf);
}

// A closed latitude-longitude sphere, outward winding (FlowCraft's STL convention).
static void SphereMesh(float r, int nlon, int nlat, std::vector<Vec3>& pos, std::vector<uint32_t>& idx) {
    const float pi = 3.14159265f;
    pos.push_back(Vec3{0.0f, 0.0f, r});
    for (int i = 1; i < nlat; ++i)
        for (int j = 0; j < nlon; ++j) {
            const float t = pi * i / nlat, p = 2.0f * pi * j / nlon;
            pos.push_back(Vec3{r * std::sin(t) * std::cos(p), r * std::sin(t) * std::sin(p), r * std::cos(t)});
        }
    pos.push_back(Vec3{0.0f, 0.0f, -r});
    const uint32_t bot = static_cast<uint32_t>(pos.size() - 1);
    auto v = [&](int i, int j) { return static_cast<uint32_t>(1 + i * nlon + (j % nlon)); };
    for (int j = 0; j < nlon; ++j) {
        idx.insert(idx.end(), {0u, v(0, j), v(0, j + 1)});
        idx.insert(idx.end(), {bot, v(nlat - 2, j + 1), v(nlat - 2, j)});
    }
    for (int i = 0; i + 2 < nlat; ++i)
        for (int j = 0; j < nlon; ++j) {
            idx.insert(idx.end(), {v(i, j), v(i + 1, j), v(i + 1, j + 1)});
            idx.insert(idx.end(), {v(i, j), v(i + 1, j + 1), v(i, j + 1)});
        }
}

static VoxelGrid SphereGrid() {
    VoxelGrid grid;
    grid.nx = 48; grid.ny = 24; grid.nz = 24; grid.cellSize = 1.0f;
    grid.solid.assign(grid.CellCount(), 0);
    grid.signedDistance.assign(grid.CellCount(), 1e30f);
    const float cx = 20, cy = 12, cz = 12, r = 5;
    for (int z = 0; z < grid.nz; ++z)
        for (int y = 0; y < grid.ny; ++y)
            for (int x = 0; x < grid.nx; ++x) {
                float d = std::sqrt((x + .5f - cx) * (x + .5f - cx) + (y + .5f - cy) * (y + .5f - cy) +
                                    (z + .5f - cz) * (z + .5f - cz)) - r;
                size_t i = grid.Index(x, y, z);
                grid.signedDistance[i] = d;
                grid.solid[i] = d < 0 ? 1 : 0;
            }
    return grid;
}

static VoxelGrid WingGrid() {
    VoxelGrid grid;
    grid.nx = 48; grid.ny = 32; grid.nz = 24; grid.cellSize = 1.0f;
    grid.solid.assign(grid.CellCount(), 0);
    grid.signedDistance.assign(grid.CellCount(), 1e30f);
    const float cx = 20.0f, cy = 16.0f, cz = 12.0f;
    const float cosA = std::cos(0.1745f), sinA = std::sin(0.1745f);
    for (int z = 0; z < grid.nz; ++z)
        for (int y = 0; y < grid.ny; ++y)
            for (int x = 0; x < grid.nx; ++x) {
                float dx = (x + 0.5f) - cx, dy = (y + 0.5f) - cy, dz = (z + 0.5f) - cz;
                float rx = dx * cosA - dz * sinA, rz = dx * sinA + dz * cosA;
                float d = (std::sqrt((rx / 7.0f) * (rx / 7.0f) + (dy / 9.0f) * (dy / 9.0f) +
                                     (rz / 2.2f) * (rz / 2.2f)) - 1.0f) * 3.0f;
                size_t i = grid.Index(x, y, z);
                grid.signedDistance[i] = d;
                grid.solid[i] = d < 0.0f ? 1 : 0;
            }
    return grid;
}

int main() {
    int failures = 0;
    for (int layout : {FC_LBM_LAYOUT_PINGPONG, FC_LBM_LAYOUT_INPLACE}) {
        for (LatticeModel model : {LatticeModel::D3Q19, LatticeModel::D3Q15}) {
            VoxelGrid grid = SphereGrid();
            LbmSolver::Params gp; gp.tau = 0.6f; gp.inletVelocity = 0.05f; gp.smagorinskyCs = 0.12f;
            LbmSolver gpu;
            LbmSolver::EngineOptions opt; opt.layout = layout;
            if (!gpu.Init(nullptr, grid, model, opt)) { std::printf("Init failed: %s\n", gpu.LastError().c_str()); return 1; }
            gpu.Reset(gp, true);
            const int q = model == LatticeModel::D3Q19 ? 19 : 15;
            void* cpu = orc_create(grid.nx, grid.ny, grid.nz, q, 32, 0, 0, 0, 0);
            orc_set_geometry(cpu, grid.solid.data(), grid.signedDistance.data());
            orc_set_params(cpu, gp.tau, gp.inletVelocity, 0, 1, 1, gp.smagorinskyCs);
            orc_reset(cpu, 1);
            const int kSteps = 60;
            gpu.Step(nullptr, kSteps);
            orc_step(cpu, kSteps);
            std::vector<float> g = gpu.ReadbackField(), c(grid.CellCount() * 4);
            orc_readback_field_f32(cpu, c.data());
            orc_destroy(cpu);
            double maxRho = 0, maxU = 0, sumU = 0; size_t n = 0, bitdiff = 0; bool nan = false;
            for (size_t i = 0; i < grid.CellCount(); ++i) {
                for (int k = 0; k < 4; ++k) {
                    if (!std::isfinite(g[i * 4 + k])) nan = true;
                    if (g[i * 4 + k] != c[i * 4 + k]) ++bitdiff;
                }
                if (grid.solid[i]) continue;
                maxRho = std::max<double>(maxRho, std::fabs(g[i * 4] - c[i * 4]));
                for (int k = 1; k < 4; ++k) {
                    double e = std::fabs(g[i * 4 + k] - c[i * 4 + k]);
                    maxU = std::max(maxU, e); sumU += e; ++n;
                }
            }
            bool ok = !nan && maxRho <= 1e-5 && maxU <= 1e-5 * 0.06;
            std::printf("%s %s: max|drho|=%.3e max|du|=%.3e mean|du|=%.3e differing values=%zu -> %s\n",
                        q == 19 ? "D3Q19" : "D3Q15", layout == FC_LBM_LAYOUT_INPLACE ? "inplace" : "pingpong",
                        maxRho, maxU, sumU / double(n), bitdiff, ok ? "PASS" : "FAIL");
            if (!ok) ++failures;
        }
        {
            VoxelGrid grid = WingGrid();
            LbmSolver::Params lp; lp.tau = 0.6f; lp.inletVelocity = 0.05f; lp.smagorinskyCs = 0.12f;
            LbmSolver q19, q15;
            LbmSolver::EngineOptions opt; opt.layout = layout;
            q19.Init(nullptr, grid, LatticeModel::D3Q19, opt); q19.Reset(lp, true);
            q15.Init(nullptr, grid, LatticeModel::D3Q15, opt); q15.Reset(lp, true);
            // 400 steps: at the reference test's 80 the lift is still in the start-up
            // transient (LbmReference itself gives Fz < 0 there).
            q19.Step(nullptr, 400);
            q15.Step(nullptr, 400);
            LbmSolver::Stats s19 = q19.ReadbackStats(), s15 = q15.ReadbackStats();
            bool okWing = !s19.hasNaN && !s15.hasNaN && s19.forceX > 0.0f && s15.forceX > 0.0f &&
                          s19.forceZ > 0.0f && s15.forceZ > 0.0f && std::fabs(s19.avgSpeed - s15.avgSpeed) < 0.005f;
            std::printf("Wing D3Q19 (Fx=%.5f Fz=%.5f avgU=%.5f) vs D3Q15 (Fx=%.5f Fz=%.5f avgU=%.5f) -> %s\n",
                        s19.forceX, s19.forceZ, s19.avgSpeed, s15.forceX, s15.forceZ, s15.avgSpeed,
                        okWing ? "PASS" : "FAIL");
            if (!okWing) ++failures;
        }
    }
    {
        // ReadbackFieldRange on sub-ranges (FlowCraft/LbmSolver.cpp:584-651): z0 > 0, odd
        // counts, ranges clamped at either end, empty and out-of-lattice ranges.  The covered
        // layers equal ReadbackField() bit for bit; the rest of the caller's vector is untouched.
        VoxelGrid grid = SphereGrid();
        LbmSolver::Params p; p.tau = 0.6f; p.inletVelocity = 0.05f; p.smagorinskyCs = 0.12f;
        bool ok = true;
        const float sentinel = -7777.25f;
        const int ranges[][2] = {{0, 24}, {5, 7}, {23, 1}, {-4, 9}, {20, 100}, {24, 3}, {3, 0}, {6, -2}};
        for (int layout : {FC_LBM_LAYOUT_PINGPONG, FC_LBM_LAYOUT_INPLACE}) {
            LbmSolver s;
            LbmSolver::EngineOptions o; o.layout = layout;
            ok = ok && s.Init(nullptr, grid, LatticeModel::D3Q19, o);
            if (!ok) break;
            s.Reset(p, true);
            s.Step(nullptr, 30);
            const std::vector<float> whole = s.ReadbackField();
            const size_t layer = size_t(grid.nx) * grid.ny * 4;
            for (const auto& r : ranges) {
                std::vector<float> out(whole.size(), sentinel);
                s.ReadbackFieldRange(r[0], r[1], out);
                const int lo = std::max(r[0], 0), hi = std::min(r[0] + r[1], grid.nz);
                for (int z = 0; z < grid.nz; ++z) {
                    const float* a = out.data() + z * layer;
                    if (r[1] > 0 && z >= lo && z < hi) {
                        ok = ok && std::memcmp(a, whole.data() + z * layer, layer * sizeof(float)) == 0;
                    } else {
                        for (size_t k = 0; k < layer; ++k) ok = ok && a[k] == sentinel;
                    }
                }
            }
        }
        std::printf("ReadbackFieldRange sub-ranges (z0 > 0, odd counts, clamped, empty) -> %s\n", ok ? "PASS" : "FAIL");
        if (!ok) ++failures;
    }
    {
        // StepPaced (FlowCraft's StepSolverPaced loop) equals plain stepping, bit for bit,
        // with the chunk depth forced down to 4 layers and with it left to grow.
        VoxelGrid grid = SphereGrid();
        LbmSolver::Params p; p.tau = 0.6f; p.inletVelocity = 0.05f; p.smagorinskyCs = 0.12f;
        LbmSolver::EngineOptions opt; opt.layout = FC_LBM_LAYOUT_PINGPONG;
        bool ok = true;
        for (double target : {1e-6, 1e6}) {
            LbmSolver a, b;
            a.Init(nullptr, grid, LatticeModel::D3Q19, opt); a.Reset(p, true);
            b.Init(nullptr, grid, LatticeModel::D3Q19, opt); b.Reset(p, true);
            LbmSolver::PacingState st;
            const int done = a.StepPaced(12, st, nullptr, target);
            b.Step(nullptr, 12);
            std::vector<float> fa = a.ReadbackField(), fb = b.ReadbackField();
            ok = ok && done == 12 && !st.diverged && std::memcmp(fa.data(), fb.data(), fa.size() * 4) == 0 &&
                 st.chunkZDepth == (target < 1.0 ? 4 : 256);
        }
        // a cancel already raised stops before the first chunk in either layout: no substep
        // done, state untouched
        for (int layout : {FC_LBM_LAYOUT_PINGPONG, FC_LBM_LAYOUT_INPLACE}) {
            LbmSolver::EngineOptions o2; o2.layout = layout;
            LbmSolver a, b;
            a.Init(nullptr, grid, LatticeModel::D3Q19, o2); a.Reset(p, true); a.Step(nullptr, 3);
            b.Init(nullptr, grid, LatticeModel::D3Q19, o2); b.Reset(p, true); b.Step(nullptr, 3);
            std::atomic<bool> cancel{true};
            LbmSolver::PacingState st;
            const int done = a.StepPaced(5, st, &cancel);
            std::vector<float> fa = a.ReadbackField(), fb = b.ReadbackField();
            ok = ok && done == 0 && std::memcmp(fa.data(), fb.data(), fa.size() * 4) == 0;
        }
        std::printf("StepPaced vs Step (chunks of 4 and of 256 layers; raised cancel) -> %s\n", ok ? "PASS" : "FAIL");
        if (!ok) ++failures;
    }
    {
        // FlowCraft's STL -> voxelizer -> solver chain on the replacement: VoxelizeOnDevice
        // equals the oracle's restated Voxelizer::Voxelize bit for bit, the grid runs, and
        // the device aerodynamic summary equals the oracle's on the same populations' field.
        std::vector<Vec3> pos;
        std::vector<uint32_t> idx;
        SphereMesh(10.0f, 48, 24, pos, idx);
        VoxelGrid grid = VoxelizeOnDevice(pos, idx, 64);
        std::vector<float> flat;
        for (const Vec3& p : pos) flat.insert(flat.end(), {p.x, p.y, p.z});
        const float pad[3] = {0.3f, 1.0f, 0.5f};
        int dims[3];
        float frame[4];
        orc_vox_frame(flat.data(), (long long)pos.size(), (long long)idx.size(), 64, pad, dims, frame);
        std::vector<uint8_t> solid(size_t(dims[0]) * dims[1] * dims[2]);
        std::vector<float> sdf(solid.size());
        orc_voxelize(flat.data(), (long long)pos.size(), idx.data(), (long long)idx.size(), 64, pad, solid.data(),
                     sdf.data());
        bool okVox = grid.nx == dims[0] && grid.ny == dims[1] && grid.nz == dims[2] && grid.cellSize == frame[0] &&
                     grid.solid == solid &&
                     std::equal(sdf.begin(), sdf.end(), grid.signedDistance.begin(),
                                [](float a, float b) { return std::memcmp(&a, &b, 4) == 0; });
        LbmSolver s;
        LbmSolver::Params p; p.tau = 0.6f; p.inletVelocity = 0.05f; p.smagorinskyCs = 0.1f;
        bool okRun = s.Init(nullptr, grid, LatticeModel::D3Q19);
        if (okRun) {
            s.Reset(p, true);
            s.Step(nullptr, 40);
            const LbmSolver::AerodynamicForces a = s.ComputeAerodynamicForces(grid);
            std::vector<float> field = s.ReadbackField();
            const float org[3] = {grid.origin.x, grid.origin.y, grid.origin.z};
            float o[11];
            int nan = 0;
            orc_aero_forces(grid.nx, grid.ny, grid.nz, grid.cellSize, org, grid.solid.data(), field.data(), o, &nan);
            auto close = [](float x, float y) { return std::fabs(x - y) <= 1e-6f * std::max(std::fabs(y), 1e-20f); };
            okRun = !a.hasNaN && !nan && a.forceX > 0.0f && close(a.forceX, o[0]) && close(a.forceY, o[1]) &&
                    close(a.forceZ, o[2]) && a.maxStress == o[4] && close(a.centerOfPressure.x, o[5]) &&
                    close(a.centerOfPressure.y, o[6]) && close(a.centerOfPressure.z, o[7]) &&
                    a.liftArrowBase.z == o[10];
            std::printf("Mesh sphere: voxelized %dx%dx%d (%s), aero Fx=%.5f cop=(%.3f, %.3f, %.3f) -> %s\n", grid.nx,
                        grid.ny, grid.nz, okVox ? "bit-exact" : "DIFFERS", a.forceX, a.centerOfPressure.x,
                        a.centerOfPressure.y, a.centerOfPressure.z, okVox && okRun ? "PASS" : "FAIL");
        } else {
            std::printf("Mesh sphere: Init failed: %s -> FAIL\n", s.LastError().c_str());
        }
        if (!(okVox && okRun)) ++failures;
    }
    return failures ? 1 : 0;
}
