// Smoke trails on the device: FlowCraft's SmokeTrailSystem (FlowCraft/SmokeTrail.cpp)
// advecting its particles through the solver's velocity field where the field lives,
// instead of through a 4-floats-per-cell host readback every frame.
//
//   KSmokeAdvance  one thread per particle, SmokeTrail.cpp:277-345: a particle whose head
//                  has left the grid or entered a solid cell is marked for respawn (or
//                  cleared when respawns are suspended); the others move by the
//                  trilinearly interpolated velocity (SmokeTrail.cpp:52-99, solid cells
//                  read as zero, grid components swapped to world) times the clamped
//                  time step, appended to a ring buffer of the last trailLength points.
//   respawns       on the host, in particle order, with the reference's generator
//                  (std::mt19937 seeded 12345, the standard library's distributions):
//                  a respawn's draws depend only on its own particle, so drawing for the
//                  marked particles in index order reproduces the reference's stream.
//                  KSmokePlace writes the new one-point trails.
//
// Every float operation is the reference's, in its order (-ffp-contract=off).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/flowcraft/fc_lbm.h"

struct fc_smoke {
    int nx = 0, ny = 0, nz = 0;
    float cs = 1.0f, o[3] = {0, 0, 0};
    int P = 0, L = 2;
    std::vector<uint8_t> solid;  // host copy: the respawn's cell picks
    uint8_t* dsolid = nullptr;
    float* dtrail = nullptr;  // P * L points (x, y, z), ring buffer per particle
    int* dhead = nullptr;     // oldest point of each ring
    int* dcount = nullptr;    // points held (0..L)
    uint8_t* dflag = nullptr; // after an advance: 0 moved or idle, 1 respawn, 2 cleared
    int* dcells = nullptr;    // stepping from a solver: 8 corner cells per particle
    float* dgath = nullptr;   // and the solver's field at them (4 floats each)
    int gath_cap = 0;         // particles the two buffers hold
    float* dplace = nullptr;  // respawn positions (4 floats: index bits, x, y, z)
    int place_cap = 0;
    hipStream_t stream = nullptr;
    std::mt19937 rng{12345};
    std::string err;
};

namespace {

int SmokeFail(fc_smoke* s, int code, const std::string& m) {
    if (s) s->err = m;
    return code;
}

#define SMOKE_HIP(s, x)                                                                               \
    do {                                                                                              \
        hipError_t e_ = (x);                                                                          \
        if (e_ != hipSuccess) return SmokeFail(s, FC_LBM_EDEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct SmokeGrid {
    int nx, ny, nz;
    float cs, ox, oy, oz;
};

__device__ __forceinline__ bool Solid(const SmokeGrid& g, const uint8_t* solid, int x, int y, int z) {
    return solid[(static_cast<size_t>(z) * g.ny + y) * g.nx + x] != 0;
}

// The 8 cell centres around a world position (clamped at the grid edge) and the
// trilinear weights, grid x from world Y and grid y from world X.
__device__ __forceinline__ void CornerCells(const SmokeGrid& g, float px, float py, float pz, int (&cell)[8], float& tx,
                                            float& ty, float& tz) {
    const float lx = px - g.ox, ly = py - g.oy, lz = pz - g.oz;
    const float gxf = ly / g.cs - 0.5f, gyf = lx / g.cs - 0.5f, gzf = lz / g.cs - 0.5f;
    const int gx0 = static_cast<int>(floorf(gxf)), gy0 = static_cast<int>(floorf(gyf)),
              gz0 = static_cast<int>(floorf(gzf));
    tx = gxf - static_cast<float>(gx0);
    ty = gyf - static_cast<float>(gy0);
    tz = gzf - static_cast<float>(gz0);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int x = min(max(gx0 + (k & 1), 0), g.nx - 1);
        const int y = min(max(gy0 + ((k >> 1) & 1), 0), g.ny - 1);
        const int z = min(max(gz0 + (k >> 2), 0), g.nz - 1);
        cell[k] = (z * g.ny + y) * g.nx + x;
    }
}

// Velocity at a world position: trilinear over the 8 surrounding cell centres, solid
// cells zero, grid (ux, uy, uz) returned as world (uy, ux, uz).  fetch(k, cell) gives the
// 4-float field entry of corner k (the whole field, or values gathered per particle).
template <class Fetch>
__device__ __forceinline__ void Sample(const SmokeGrid& g, const uint8_t* solid, Fetch fetch, float px, float py,
                                       float pz, float& vx, float& vy, float& vz) {
    int cell[8];
    float tx, ty, tz;
    CornerCells(g, px, py, pz, cell, tx, ty, tz);
    float c[8][3];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (solid[cell[k]]) {
            c[k][0] = c[k][1] = c[k][2] = 0.0f;
        } else {
            const float* f = fetch(k, cell[k]);
            c[k][0] = f[2];
            c[k][1] = f[1];
            c[k][2] = f[3];
        }
    }
    float r[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float c00 = c[0][d] + (c[1][d] - c[0][d]) * tx;
        const float c10 = c[2][d] + (c[3][d] - c[2][d]) * tx;
        const float c01 = c[4][d] + (c[5][d] - c[4][d]) * tx;
        const float c11 = c[6][d] + (c[7][d] - c[6][d]) * tx;
        const float c0 = c00 + (c10 - c00) * ty;
        const float c1 = c01 + (c11 - c01) * ty;
        r[d] = c0 + (c1 - c0) * tz;
    }
    vx = r[0];
    vy = r[1];
    vz = r[2];
}

// The newest trail point of particle p (its head).
__device__ __forceinline__ const float* Head(const float* trail, const int* head, int p, int n, int L) {
    const int last = (head[p] + n - 1) % L;
    return trail + (static_cast<size_t>(p) * L + last) * 3;
}

// Corner cells of every particle's head for a gathered field (8 per particle; a particle
// with no points, or a head outside the grid, lists valid cells whose values go unused).
__global__ void KSmokeCorners(const SmokeGrid g, const float* __restrict__ trail, const int* __restrict__ head,
                              const int* __restrict__ count, int P, int L, int* __restrict__ cells) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    int c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int n = count[p];
    if (n > 0) {
        const float* t = Head(trail, head, p, n, L);
        float tx, ty, tz;
        CornerCells(g, t[0], t[1], t[2], c, tx, ty, tz);
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) cells[static_cast<size_t>(p) * 8 + k] = c[k];
}

// GATHER: `field` holds 8 gathered entries per particle (KSmokeCorners' cells) instead
// of the whole field.
template <bool GATHER>
__global__ void KSmokeAdvance(const SmokeGrid g, const uint8_t* __restrict__ solid, const float* __restrict__ field,
                              float* __restrict__ trail, int* __restrict__ head, int* __restrict__ count,
                              uint8_t* __restrict__ flag, int P, int L, float scale, int suspend, int confine_axis) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int n = count[p];
    if (n == 0) {
        flag[p] = suspend ? 0 : 1;
        return;
    }
    const int h = head[p];
    float* t = trail + static_cast<size_t>(p) * L * 3;
    const int last = (h + n - 1) % L;
    const float cx = t[last * 3], cy = t[last * 3 + 1], cz = t[last * 3 + 2];
    // the cell under the head (SmokeTrail.cpp:36-50): grid-x from world Y, grid-y from X
    const int ix = static_cast<int>(floorf((cy - g.oy) / g.cs));
    const int iy = static_cast<int>(floorf((cx - g.ox) / g.cs));
    const int iz = static_cast<int>(floorf((cz - g.oz) / g.cs));
    const bool in = ix >= 0 && iy >= 0 && iz >= 0 && ix < g.nx && iy < g.ny && iz < g.nz;
    if (!in || Solid(g, solid, ix, iy, iz)) {
        if (suspend) {
            count[p] = 0;
            flag[p] = 2;
        } else {
            flag[p] = 1;
        }
        return;
    }
    float vx, vy, vz;
    if constexpr (GATHER)
        Sample(g, solid, [&](int k, int) { return field + (static_cast<size_t>(p) * 8 + k) * 4; }, cx, cy, cz, vx, vy,
               vz);
    else
        Sample(g, solid, [&](int, int c) { return field + static_cast<size_t>(c) * 4; }, cx, cy, cz, vx, vy, vz);
    if (confine_axis == 0) vy = 0.0f;  // grid-x plane: world Y fixed
    if (confine_axis == 1) vx = 0.0f;  // grid-y plane: world X fixed
    if (confine_axis == 2) vz = 0.0f;
    const float nx_ = cx + vx * scale, ny_ = cy + vy * scale, nz_ = cz + vz * scale;
    if (n < L) {
        const int slot = (h + n) % L;
        t[slot * 3] = nx_;
        t[slot * 3 + 1] = ny_;
        t[slot * 3 + 2] = nz_;
        count[p] = n + 1;
    } else {
        t[h * 3] = nx_;  // overwrite the oldest
        t[h * 3 + 1] = ny_;
        t[h * 3 + 2] = nz_;
        head[p] = (h + 1) % L;
    }
    flag[p] = 0;
}

// Respawned particles: a fresh trail of one point.  rec[k] = (particle, x, y, z).
__global__ void KSmokePlace(const float* __restrict__ rec, int n, float* __restrict__ trail, int* __restrict__ head,
                            int* __restrict__ count, int L) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int p = __float_as_int(rec[k * 4]);
    float* t = trail + static_cast<size_t>(p) * L * 3;
    t[0] = rec[k * 4 + 1];
    t[1] = rec[k * 4 + 2];
    t[2] = rec[k * 4 + 3];
    head[p] = 0;
    count[p] = 1;
}

// --- host side: the reference's random cell picks (SmokeTrail.cpp:1-35, 148-200)

bool HostSolid(const fc_smoke* s, int x, int y, int z) {
    return x >= 0 && y >= 0 && z >= 0 && x < s->nx && y < s->ny && z < s->nz &&
           s->solid[(static_cast<size_t>(z) * s->ny + y) * s->nx + x] != 0;
}

void PickCell(fc_smoke* s, int fx, int fy, int fz, int& ox, int& oy, int& oz) {
    std::uniform_int_distribution<int> xd(0, std::max(s->nx - 1, 0));
    std::uniform_int_distribution<int> yd(0, std::max(s->ny - 1, 0));
    std::uniform_int_distribution<int> zd(0, std::max(s->nz - 1, 0));
    ox = fx >= 0 ? fx : s->nx / 2;
    oy = fy >= 0 ? fy : s->ny / 2;
    oz = fz >= 0 ? fz : s->nz / 2;
    for (int attempt = 0; attempt < 8; ++attempt) {
        const int tx = fx >= 0 ? fx : xd(s->rng);
        const int ty = fy >= 0 ? fy : yd(s->rng);
        const int tz = fz >= 0 ? fz : zd(s->rng);
        if (!HostSolid(s, tx, ty, tz)) {
            ox = tx;
            oy = ty;
            oz = tz;
            return;
        }
    }
}

// A cell centre (VoxelGrid::CellCenter) jittered within the cell along the free axes.
void Jittered(fc_smoke* s, int x, int y, int z, bool xf, bool yf, bool zf, float* out) {
    std::uniform_real_distribution<float> jitter(-0.5f, 0.5f);
    const float jx = xf ? 0.0f : jitter(s->rng) * s->cs;
    const float jy = yf ? 0.0f : jitter(s->rng) * s->cs;
    const float jz = zf ? 0.0f : jitter(s->rng) * s->cs;
    out[0] = (s->o[0] + (y + 0.5f) * s->cs) + jy;
    out[1] = (s->o[1] + (x + 0.5f) * s->cs) + jx;
    out[2] = (s->o[2] + (z + 0.5f) * s->cs) + jz;
}

int Fixed(int axis, int want, int index, int n) { return axis == want ? std::clamp(index, 0, n - 1) : -1; }

void Respawn(fc_smoke* s, int reverse, int axis, int index, float* out) {
    const int fx = axis == 0 ? std::clamp(index, 0, s->nx - 1) : (reverse ? s->nx - 1 : 0);
    const int fy = Fixed(axis, 1, index, s->ny), fz = Fixed(axis, 2, index, s->nz);
    int x, y, z;
    PickCell(s, fx, fy, fz, x, y, z);
    Jittered(s, x, y, z, true, fy < 0, fz < 0, out);
}

int Place(fc_smoke* s, const std::vector<float>& rec) {
    const int n = static_cast<int>(rec.size() / 4);
    if (n == 0) return FC_LBM_OK;
    if (n > s->place_cap) {
        if (s->dplace) SMOKE_HIP(s, hipFree(s->dplace));
        s->dplace = nullptr;
        SMOKE_HIP(s, hipMalloc(&s->dplace, rec.size() * sizeof(float)));
        s->place_cap = n;
    }
    SMOKE_HIP(s, hipMemcpyAsync(s->dplace, rec.data(), rec.size() * sizeof(float), hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(KSmokePlace, dim3((n + 255) / 256), dim3(256), 0, s->stream, s->dplace, n, s->dtrail, s->dhead,
                       s->dcount, s->L);
    SMOKE_HIP(s, hipGetLastError());
    SMOKE_HIP(s, hipStreamSynchronize(s->stream));
    return FC_LBM_OK;
}

void PushRec(std::vector<float>& rec, int p, const float* pos) {
    float pb;
    static_assert(sizeof(pb) == sizeof(p), "index bits");
    std::memcpy(&pb, &p, sizeof(pb));
    rec.insert(rec.end(), {pb, pos[0], pos[1], pos[2]});
}

int CheckAxis(fc_smoke* s, int axis) {
    if (axis < -1 || axis > 2) return SmokeFail(s, FC_LBM_EINVAL, "confine axis must be -1 (none), 0, 1 or 2");
    return FC_LBM_OK;
}

// After an advance (complete on the device): respawn the marked particles, in particle
// order, with the reference's generator.
int FinishStep(fc_smoke* s, int reverse_flow, int suspend_respawn, int confine_axis, int confine_index) {
    if (suspend_respawn) return FC_LBM_OK;
    std::vector<uint8_t> flag(s->P);
    SMOKE_HIP(s, hipMemcpyAsync(flag.data(), s->dflag, s->P, hipMemcpyDeviceToHost, s->stream));
    SMOKE_HIP(s, hipStreamSynchronize(s->stream));
    std::vector<float> rec;
    for (int p = 0; p < s->P; ++p) {
        if (flag[p] != 1) continue;
        float pos[3];
        Respawn(s, reverse_flow, confine_axis, confine_index, pos);
        PushRec(rec, p, pos);
    }
    return Place(s, rec);
}

}  // namespace

extern "C" {

int fc_smoke_create(int nx, int ny, int nz, float cell_size, const float origin[3], const uint8_t* solid,
                    fc_smoke** out) {
    if (!out) return FC_LBM_EINVAL;
    *out = nullptr;
    if (nx <= 0 || ny <= 0 || nz <= 0 || !origin || !solid || !(cell_size > 0.0f)) return FC_LBM_EINVAL;
    if (static_cast<long long>(nx) * ny * nz >= (1LL << 31)) return FC_LBM_EINVAL;  // cell indices are int
    fc_smoke* s = new fc_smoke();
    *out = s;  // returned even on failure so fc_smoke_last_error can explain; destroy it either way
    s->nx = nx;
    s->ny = ny;
    s->nz = nz;
    s->cs = cell_size;
    for (int k = 0; k < 3; ++k) s->o[k] = origin[k];
    const size_t N = static_cast<size_t>(nx) * ny * nz;
    s->solid.assign(solid, solid + N);
    SMOKE_HIP(s, hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    SMOKE_HIP(s, hipMalloc(&s->dsolid, N));
    SMOKE_HIP(s, hipMemcpy(s->dsolid, solid, N, hipMemcpyHostToDevice));
    return FC_LBM_OK;
}

int fc_smoke_destroy(fc_smoke* s) {
    if (!s) return FC_LBM_EINVAL;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (void* p : {static_cast<void*>(s->dsolid), static_cast<void*>(s->dtrail), static_cast<void*>(s->dhead),
                    static_cast<void*>(s->dcount), static_cast<void*>(s->dflag), static_cast<void*>(s->dcells),
                    static_cast<void*>(s->dgath), static_cast<void*>(s->dplace)})
        if (p) (void)hipFree(p);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return FC_LBM_OK;
}

const char* fc_smoke_last_error(const fc_smoke* s) { return s ? s->err.c_str() : "null fc_smoke handle"; }

int fc_smoke_scatter(fc_smoke* s, int confine_axis, int confine_index) {
    if (!s) return FC_LBM_EINVAL;
    int rc = CheckAxis(s, confine_axis);
    if (rc) return rc;
    const int fx = Fixed(confine_axis, 0, confine_index, s->nx), fy = Fixed(confine_axis, 1, confine_index, s->ny),
              fz = Fixed(confine_axis, 2, confine_index, s->nz);
    std::vector<float> rec;
    rec.reserve(static_cast<size_t>(s->P) * 4);
    for (int p = 0; p < s->P; ++p) {
        int x, y, z;
        PickCell(s, fx, fy, fz, x, y, z);
        float pos[3];
        Jittered(s, x, y, z, fx >= 0, fy >= 0, fz >= 0, pos);
        PushRec(rec, p, pos);
    }
    return Place(s, rec);
}

int fc_smoke_reset(fc_smoke* s, int reverse_flow, int particle_count, int trail_length, int confine_axis,
                   int confine_index) {
    if (!s) return FC_LBM_EINVAL;
    (void)reverse_flow;  // the initial scatter does not use the inlet face (SmokeTrail.cpp:261-273)
    int rc = CheckAxis(s, confine_axis);
    if (rc) return rc;
    const int P = std::max(particle_count, 0), L = std::max(trail_length, 2);
    for (void* p : {static_cast<void*>(s->dtrail), static_cast<void*>(s->dhead), static_cast<void*>(s->dcount),
                    static_cast<void*>(s->dflag)})
        if (p) SMOKE_HIP(s, hipFree(p));
    s->dtrail = nullptr;
    s->dhead = s->dcount = nullptr;
    s->dflag = nullptr;
    s->P = P;
    s->L = L;
    if (P > 0) {
        SMOKE_HIP(s, hipMalloc(&s->dtrail, static_cast<size_t>(P) * L * 3 * sizeof(float)));
        SMOKE_HIP(s, hipMalloc(&s->dhead, P * sizeof(int)));
        SMOKE_HIP(s, hipMalloc(&s->dcount, P * sizeof(int)));
        SMOKE_HIP(s, hipMalloc(&s->dflag, P));
        SMOKE_HIP(s, hipMemsetAsync(s->dcount, 0, P * sizeof(int), s->stream));
    }
    return fc_smoke_scatter(s, confine_axis, confine_index);
}

int fc_smoke_step_field(fc_smoke* s, const float* field_device, int reverse_flow, float mm_per_sec_per_lattice_unit,
                        float dt_seconds, int suspend_respawn, int confine_axis, int confine_index) {
    if (!s) return FC_LBM_EINVAL;
    int rc = CheckAxis(s, confine_axis);
    if (rc) return rc;
    if (!field_device || dt_seconds <= 0.0f) return FC_LBM_OK;  // SmokeTrail.cpp:280: nothing to do
    if (s->P == 0) return FC_LBM_OK;
    const float dt = std::min(dt_seconds, 0.1f);
    const float scale = mm_per_sec_per_lattice_unit * dt;
    const SmokeGrid g{s->nx, s->ny, s->nz, s->cs, s->o[0], s->o[1], s->o[2]};
    hipLaunchKernelGGL(KSmokeAdvance<false>, dim3((s->P + 255) / 256), dim3(256), 0, s->stream, g, s->dsolid,
                       field_device, s->dtrail, s->dhead, s->dcount, s->dflag, s->P, s->L, scale,
                       suspend_respawn ? 1 : 0, confine_axis);
    SMOKE_HIP(s, hipGetLastError());
    // blocking, so the caller may reuse the field at once
    SMOKE_HIP(s, hipStreamSynchronize(s->stream));
    return FinishStep(s, reverse_flow, suspend_respawn, confine_axis, confine_index);
}

int fc_smoke_step(fc_smoke* s, fc_lbm* h, int reverse_flow, float mm_per_sec_per_lattice_unit, float dt_seconds,
                  int suspend_respawn, int confine_axis, int confine_index) {
    if (!s || !h) return FC_LBM_EINVAL;
    int rc = CheckAxis(s, confine_axis);
    if (rc) return rc;
    int nx, ny, nz, q, st, lay;
    if (fc_lbm_info(h, &nx, &ny, &nz, &q, &st, &lay) != FC_LBM_OK || nx != s->nx || ny != s->ny || nz != s->nz)
        return SmokeFail(s, FC_LBM_EINVAL, "solver lattice does not match the smoke grid");
    if (dt_seconds <= 0.0f || s->P == 0) return FC_LBM_OK;
    // Only the 8 cells around each head are read: their field values are gathered from the
    // solver's populations (fc_lbm_field_at, the same values as the whole field there),
    // all on the solver's stream, behind the steps already queued on it.
    if (s->gath_cap < s->P) {
        if (s->dcells) SMOKE_HIP(s, hipFree(s->dcells));
        if (s->dgath) SMOKE_HIP(s, hipFree(s->dgath));
        s->dcells = nullptr;
        s->dgath = nullptr;
        s->gath_cap = 0;
        SMOKE_HIP(s, hipMalloc(&s->dcells, static_cast<size_t>(s->P) * 8 * sizeof(int)));
        SMOKE_HIP(s, hipMalloc(&s->dgath, static_cast<size_t>(s->P) * 8 * 4 * sizeof(float)));
        s->gath_cap = s->P;
    }
    const hipStream_t hs = static_cast<hipStream_t>(fc_lbm_stream(h));
    const SmokeGrid g{s->nx, s->ny, s->nz, s->cs, s->o[0], s->o[1], s->o[2]};
    const dim3 grid((s->P + 255) / 256), block(256);
    // the trails were last written on the smoke stream
    SMOKE_HIP(s, hipStreamSynchronize(s->stream));
    hipLaunchKernelGGL(KSmokeCorners, grid, block, 0, hs, g, s->dtrail, s->dhead, s->dcount, s->P, s->L, s->dcells);
    SMOKE_HIP(s, hipGetLastError());
    rc = fc_lbm_field_at(h, s->dcells, static_cast<long long>(s->P) * 8, s->dgath);
    if (rc) return SmokeFail(s, rc, fc_lbm_last_error(h));
    const float dt = std::min(dt_seconds, 0.1f);
    hipLaunchKernelGGL(KSmokeAdvance<true>, grid, block, 0, hs, g, s->dsolid, s->dgath, s->dtrail, s->dhead, s->dcount,
                       s->dflag, s->P, s->L, mm_per_sec_per_lattice_unit * dt, suspend_respawn ? 1 : 0, confine_axis);
    SMOKE_HIP(s, hipGetLastError());
    SMOKE_HIP(s, hipStreamSynchronize(hs));
    return FinishStep(s, reverse_flow, suspend_respawn, confine_axis, confine_index);
}

int fc_smoke_shift_plane(fc_smoke* s, int axis, int index) {
    if (!s || axis < 0 || axis > 2) return FC_LBM_EINVAL;
    if (s->P == 0) return FC_LBM_OK;
    // the plane's world coordinate, VoxelGrid::CellCenter of the clamped index
    float coord;
    int comp;
    if (axis == 0) {
        coord = s->o[1] + (std::clamp(index, 0, s->nx - 1) + 0.5f) * s->cs;
        comp = 1;
    } else if (axis == 1) {
        coord = s->o[0] + (std::clamp(index, 0, s->ny - 1) + 0.5f) * s->cs;
        comp = 0;
    } else {
        coord = s->o[2] + (std::clamp(index, 0, s->nz - 1) + 0.5f) * s->cs;
        comp = 2;
    }
    std::vector<float> t(static_cast<size_t>(s->P) * s->L * 3);
    SMOKE_HIP(s, hipMemcpyAsync(t.data(), s->dtrail, t.size() * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    SMOKE_HIP(s, hipStreamSynchronize(s->stream));
    for (size_t i = comp; i < t.size(); i += 3) t[i] = coord;  // every stored point; unused slots are never read
    SMOKE_HIP(s, hipMemcpyAsync(s->dtrail, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice, s->stream));
    SMOKE_HIP(s, hipStreamSynchronize(s->stream));
    return FC_LBM_OK;
}

int fc_smoke_trails(fc_smoke* s, float* out, int* lengths) {
    if (!s || (!out && s->P > 0) || (!lengths && s->P > 0)) return FC_LBM_EINVAL;
    if (s->P == 0) return FC_LBM_OK;
    const size_t n = static_cast<size_t>(s->P) * s->L * 3;
    std::vector<float> t(n);
    std::vector<int> head(s->P), count(s->P);
    SMOKE_HIP(s, hipMemcpyAsync(t.data(), s->dtrail, n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    SMOKE_HIP(s, hipMemcpyAsync(head.data(), s->dhead, s->P * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    SMOKE_HIP(s, hipMemcpyAsync(count.data(), s->dcount, s->P * sizeof(int), hipMemcpyDeviceToHost, s->stream));
    SMOKE_HIP(s, hipStreamSynchronize(s->stream));
    for (int p = 0; p < s->P; ++p) {
        lengths[p] = count[p];
        for (int k = 0; k < count[p]; ++k) {
            const size_t src = (static_cast<size_t>(p) * s->L + (head[p] + k) % s->L) * 3;
            const size_t dst = (static_cast<size_t>(p) * s->L + k) * 3;
            out[dst] = t[src];
            out[dst + 1] = t[src + 1];
            out[dst + 2] = t[src + 2];
        }
    }
    return FC_LBM_OK;
}

int fc_smoke_info(const fc_smoke* s, int* particles, int* trail_length) {
    if (!s) return FC_LBM_EINVAL;
    if (particles) *particles = s->P;
    if (trail_length) *trail_length = s->L;
    return FC_LBM_OK;
}

}  // extern "C"
