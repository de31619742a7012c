// Mesh voxelizer on the device: FlowCraft's Voxelizer::Voxelize
// (FlowCraft/Voxelizer.cpp:128-432) — triangle mesh in world space to the solver's
// solid mask and signed-distance field — with the same results bit for bit.
//
//   frame        host: padded mesh bounds -> cell size, origin, nx/ny/nz (grid-x along
//                world Y, grid-y along world X; Voxelizer.cpp:137-166).
//   setup        host, O(triangles): unit normal, degenerate skip, each triangle's cell
//                box (the reference's conservative-radius AABB), and a brick index: for
//                every 8x8x8-cell brick, the triangles whose box reaches it, in triangle
//                order.
//   KVoxRaster   one workgroup per non-empty brick, one thread per cell.  Each cell walks
//                its brick's triangles in index order, exactly the order in which the
//                reference's per-triangle loops reach it, so the running nearest distance
//                and its order-dependent edge tie-break (Voxelizer.cpp:244-290) evolve
//                identically.  Triangles are staged through LDS in chunks.
//   flood fill   the exterior is the 6-connected component of cell (0,0,0) among the
//                non-shell cells (Voxelizer.cpp:295-412).  Union-find on the device:
//                KUfRuns links each x-run of non-shell cells to its first cell, KUfHook
//                unites runs that touch across y and z, once per overlap (lock-free:
//                roots are hooked under smaller roots with a CAS, so parent[i] <= i
//                always and every walk terminates), KUfCompress stores each cell's root.
//                Component membership does not depend on the order.
//   KVoxFinal    cells outside the seed's component become solid; the signed distance is
//                reconciled with the final classification (Voxelizer.cpp:414-432).
//
// Every float operation is the reference's, in its order (the library is built with
// -ffp-contract=off; division and square root are IEEE).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/flowcraft/fc_lbm.h"

namespace {

thread_local std::string g_vox_err;

int VoxFail(int code, const std::string& msg) {
    g_vox_err = msg;
    return code;
}

#define VOX_HIP(x)                                                                                  \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) return VoxFail(FC_LBM_EDEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int kBrick = 8;  // brick edge in cells; 512 cells per workgroup
constexpr int kChunk = 128;  // triangles staged in LDS at a time

struct V3 {
    float x, y, z;
};
__host__ __device__ __forceinline__ V3 Add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ __forceinline__ V3 Sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ __forceinline__ V3 Mul(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__host__ __device__ __forceinline__ float Dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ __forceinline__ V3 Cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ __forceinline__ float Len(V3 v) { return sqrtf(v.x * v.x + v.y * v.y + v.z * v.z); }

// Nearest point of triangle (a, b, c) to p by Voronoi region (Voxelizer.cpp:85-123).
__device__ __forceinline__ V3 Nearest(V3 p, V3 a, V3 b, V3 c) {
    const V3 ab = Sub(b, a), ac = Sub(c, a), ap = Sub(p, a);
    const float d1 = Dot(ab, ap), d2 = Dot(ac, ap);
    if (d1 <= 0.0f && d2 <= 0.0f) return a;
    const V3 bp = Sub(p, b);
    const float d3 = Dot(ab, bp), d4 = Dot(ac, bp);
    if (d3 >= 0.0f && d4 <= d3) return b;
    const float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) return Add(a, Mul(ab, d1 / (d1 - d3)));
    const V3 cp = Sub(p, c);
    const float d5 = Dot(ab, cp), d6 = Dot(ac, cp);
    if (d6 >= 0.0f && d5 <= d6) return c;
    const float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) return Add(a, Mul(ac, d2 / (d2 - d6)));
    const float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f)
        return Add(b, Mul(Sub(c, b), (d4 - d3) / ((d4 - d3) + (d5 - d6))));
    const float inv = 1.0f / (va + vb + vc);
    return Add(Add(a, Mul(ab, vb * inv)), Mul(ac, vc * inv));
}

// One triangle as the raster kernel reads it: vertices, unit normal, cell box.
struct alignas(16) TriRec {
    float v[9];
    float n[3];
    int box[6];  // x0, x1, y0, y1, z0, z1 (inclusive, clipped to the grid)
    int pad[2];
};
static_assert(sizeof(TriRec) == 80, "TriRec layout");

struct VoxArgs {
    int nx, ny, nz;
    float cs, ox, oy, oz;
    float radius, tol;
    int bx, by;  // bricks along x and y
};

__global__ __launch_bounds__(512) void KVoxRaster(const VoxArgs a, const TriRec* __restrict__ tris,
                                                  const uint32_t* __restrict__ brick_start,
                                                  const uint32_t* __restrict__ brick_tris,
                                                  const uint32_t* __restrict__ active, uint8_t* __restrict__ solid,
                                                  float* __restrict__ sdf) {
    __shared__ TriRec st[kChunk];
    const uint32_t b = active[blockIdx.x];
    const int bxi = static_cast<int>(b % a.bx), byi = static_cast<int>((b / a.bx) % a.by),
              bzi = static_cast<int>(b / (static_cast<uint32_t>(a.bx) * a.by));
    const int x = bxi * kBrick + (threadIdx.x & 7), y = byi * kBrick + ((threadIdx.x >> 3) & 7),
              z = bzi * kBrick + (threadIdx.x >> 6);
    const bool in = x < a.nx && y < a.ny && z < a.nz;
    // cell centre as VoxelGrid::CellCenter
    const V3 ctr{a.ox + (static_cast<float>(y) + 0.5f) * a.cs, a.oy + (static_cast<float>(x) + 0.5f) * a.cs,
                 a.oz + (static_cast<float>(z) + 0.5f) * a.cs};
    float sd = __builtin_inff(), proj = -1.0f;
    bool shell = false;
    const uint32_t t0 = brick_start[b], t1 = brick_start[b + 1];
    for (uint32_t base = t0; base < t1; base += kChunk) {
        const int cnt = static_cast<int>(min(static_cast<uint32_t>(kChunk), t1 - base));
        __syncthreads();
        // stage cnt records (20 words each) with all 512 threads
        const uint32_t* src = reinterpret_cast<const uint32_t*>(tris);
        uint32_t* dst = reinterpret_cast<uint32_t*>(st);
        for (int w = threadIdx.x; w < cnt * 20; w += 512) {
            const int k = w / 20, f = w - k * 20;
            dst[w] = src[static_cast<size_t>(brick_tris[base + k]) * 20 + f];
        }
        __syncthreads();
        if (!in) continue;
        for (int k = 0; k < cnt; ++k) {
            const TriRec& r = st[k];
            if (x < r.box[0] || x > r.box[1] || y < r.box[2] || y > r.box[3] || z < r.box[4] || z > r.box[5])
                continue;
            const V3 va{r.v[0], r.v[1], r.v[2]}, vb{r.v[3], r.v[4], r.v[5]}, vc{r.v[6], r.v[7], r.v[8]};
            const V3 n{r.n[0], r.n[1], r.n[2]};
            const V3 diff = Sub(ctr, Nearest(ctr, va, vb, vc));
            const float dist = Len(diff);
            const float nd = Dot(diff, n);
            const float and_ = fabsf(nd);
            const float best = fabsf(sd);
            if (dist + a.tol < best) {
                sd = nd < 0.0f ? -dist : dist;
                proj = and_;
            } else if (dist <= best + a.tol) {
                if (and_ > proj) {
                    const float m = fminf(dist, best);
                    sd = nd < 0.0f ? -m : m;
                    proj = and_;
                } else if (dist < best) {
                    sd = sd < 0.0f ? -dist : dist;
                }
            }
            if (dist <= a.radius) shell = true;
        }
    }
    if (in) {
        const size_t i = (static_cast<size_t>(z) * a.ny + y) * a.nx + x;
        solid[i] = shell ? 1 : 0;
        sdf[i] = sd;
    }
}

__global__ void KVoxInit(uint8_t* __restrict__ solid, float* __restrict__ sdf, uint32_t* __restrict__ parent,
                         long long n) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    solid[i] = 0;
    sdf[i] = __builtin_inff();
    parent[i] = static_cast<uint32_t>(i);
}

// Root of a's set; halves the path on the way (every value written is an ancestor,
// and parent[x] <= x holds throughout, so the walk strictly descends and ends).
__device__ __forceinline__ uint32_t UfFind(uint32_t* parent, uint32_t a) {
    uint32_t p = __hip_atomic_load(&parent[a], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (p != a) {
        const uint32_t gp = __hip_atomic_load(&parent[p], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (gp != p) atomicMin(&parent[a], gp);
        a = p;
        p = gp;
    }
    return a;
}

// Lock-free union: hook the larger root under the smaller with a CAS; a failed CAS means
// that root was hooked meanwhile, and the retry starts from strictly smaller roots.
__device__ __forceinline__ void UfUnion(uint32_t* parent, uint32_t a, uint32_t b) {
    for (;;) {
        a = UfFind(parent, a);
        b = UfFind(parent, b);
        if (a == b) return;
        if (a < b) {
            const uint32_t t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(&parent[a], a, b) == a) return;
    }
}

// x-runs first, without atomics: one thread per grid row points every non-shell cell at
// the first cell of its run of consecutive non-shell cells (a root, and <= the cell).
__global__ void KUfRuns(const uint8_t* __restrict__ solid, uint32_t* __restrict__ parent, int nx, long long rows) {
    const long long row = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (row >= rows) return;
    const long long base = row * nx;
    uint32_t start = 0;
    bool prev = false;
    for (int x = 0; x < nx; ++x) {
        const bool fluid = solid[base + x] == 0;
        if (fluid && !prev) start = static_cast<uint32_t>(base + x);
        if (fluid) parent[base + x] = start;
        prev = fluid;
    }
}

// Unions across y and z.  A cell whose left neighbour and whose partner's left neighbour
// are both non-shell skips the union: that pair was united one cell to the left, and
// each cell is in its left neighbour's run.  So only the first cell of each overlap of
// two runs unites them.
__global__ void KUfHook(const uint8_t* __restrict__ solid, uint32_t* __restrict__ parent, int nx, int ny, int nz) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long layer = static_cast<long long>(nx) * ny;
    if (i >= layer * nz || solid[i]) return;
    const int x = static_cast<int>(i % nx), y = static_cast<int>((i / nx) % ny), z = static_cast<int>(i / layer);
    const bool left = x > 0 && solid[i - 1] == 0;
    const uint32_t u = static_cast<uint32_t>(i);
    if (y + 1 < ny && !solid[i + nx] && !(left && solid[i + nx - 1] == 0))
        UfUnion(parent, u, static_cast<uint32_t>(i + nx));
    if (z + 1 < nz && !solid[i + layer] && !(left && solid[i + layer - 1] == 0))
        UfUnion(parent, u, static_cast<uint32_t>(i + layer));
}

__global__ void KUfCompress(const uint8_t* __restrict__ solid, uint32_t* __restrict__ parent, long long n) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || solid[i]) return;
    const uint32_t r = UfFind(parent, static_cast<uint32_t>(i));
    __hip_atomic_store(&parent[i], r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Interior = every non-shell cell outside the component of cell 0 (all of them when
// cell 0 is itself shell); then the signed distance takes the final classification.
__global__ void KVoxFinal(uint8_t* __restrict__ solid, float* __restrict__ sdf, const uint32_t* __restrict__ parent,
                          long long n, float radius) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool seed = solid[0] == 0;
    const uint32_t seed_root = parent[0];
    bool s = solid[i] != 0;
    if (!s && (!seed || parent[i] != seed_root)) s = true;
    solid[i] = s ? 1 : 0;
    const float d = fabsf(sdf[i]);
    if (!s)
        sdf[i] = d;
    else if (d > radius)
        sdf[i] = -d;
}

struct Frame {
    int nx = 0, ny = 0, nz = 0;
    float cs = 1.0f;
    V3 o{0, 0, 0};
};

// Grid frame of Voxelizer.cpp:137-166, host float arithmetic in the same order.
Frame MakeFrame(const float* pos, long long nv, const fc_vox_params& p) {
    V3 lo{std::numeric_limits<float>::max(), std::numeric_limits<float>::max(), std::numeric_limits<float>::max()};
    V3 hi{std::numeric_limits<float>::lowest(), std::numeric_limits<float>::lowest(),
          std::numeric_limits<float>::lowest()};
    for (long long i = 0; i < nv; ++i) {
        lo = {std::min(lo.x, pos[3 * i]), std::min(lo.y, pos[3 * i + 1]), std::min(lo.z, pos[3 * i + 2])};
        hi = {std::max(hi.x, pos[3 * i]), std::max(hi.y, pos[3 * i + 1]), std::max(hi.z, pos[3 * i + 2])};
    }
    const V3 size = Sub(hi, lo);
    const V3 pad{size.x * p.padding[0], size.y * p.padding[1], size.z * p.padding[2]};
    const V3 dmin = Sub(lo, pad), dmax = Add(hi, pad), ds = Sub(dmax, dmin);
    const float longest = std::max({ds.x, ds.y, ds.z});
    Frame f;
    f.cs = std::max(longest / static_cast<float>(std::max(p.target_resolution, 4)), 1e-6f);
    f.o = dmin;
    f.nx = std::max(static_cast<int>(std::ceil(ds.y / f.cs)), 1);
    f.ny = std::max(static_cast<int>(std::ceil(ds.x / f.cs)), 1);
    f.nz = std::max(static_cast<int>(std::ceil(ds.z / f.cs)), 1);
    return f;
}

int CheckMesh(const float* pos, long long nv, const uint32_t* idx, long long ni, const fc_vox_params* p) {
    if (!p) return VoxFail(FC_LBM_EINVAL, "null voxelizer parameters");
    if (nv < 0 || ni < 0 || (nv > 0 && !pos) || (ni > 0 && !idx))
        return VoxFail(FC_LBM_EINVAL, "null mesh arrays with a nonzero count");
    // the triangles read are indices [0, 3 * floor(ni / 3))
    for (long long k = 0; k < ni - ni % 3; ++k)
        if (idx[k] >= static_cast<unsigned long long>(nv)) return VoxFail(FC_LBM_EINVAL, "vertex index out of range");
    return FC_LBM_OK;
}

template <class T>
struct DevBuf {
    T* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

extern "C" {

const char* fc_voxelize_last_error(void) { return g_vox_err.c_str(); }

int fc_vox_params_default(fc_vox_params* p) {
    if (!p) return FC_LBM_EINVAL;
    p->target_resolution = 640;
    p->padding[0] = 0.3f;
    p->padding[1] = 1.0f;
    p->padding[2] = 0.5f;
    return FC_LBM_OK;
}

int fc_voxelize_frame(const float* positions, long long nverts, const uint32_t* indices, long long nindices,
                      const fc_vox_params* params, fc_vox_grid* out) {
    if (!out) return VoxFail(FC_LBM_EINVAL, "null output");
    *out = fc_vox_grid{0, 0, 0, 1.0f, {0.0f, 0.0f, 0.0f}};
    int rc = CheckMesh(positions, nverts, indices, nindices, params);
    if (rc) return rc;
    if (nverts == 0 || nindices < 3) return FC_LBM_OK;  // the reference's empty grid
    const Frame f = MakeFrame(positions, nverts, *params);
    *out = fc_vox_grid{f.nx, f.ny, f.nz, f.cs, {f.o.x, f.o.y, f.o.z}};
    return FC_LBM_OK;
}

// The whole voxelizer; results into host arrays (copied back) or, device_out, straight
// into the caller's device buffers of `capacity` cells each (checked against the frame on
// the host, before any device work).
static int Voxelize(const float* positions, long long nverts, const uint32_t* indices, long long nindices,
                    const fc_vox_params* params, uint8_t* solid, float* sdf, bool device_out,
                    long long capacity = 0) {
    int rc = CheckMesh(positions, nverts, indices, nindices, params);
    if (rc) return rc;
    if (nverts == 0 || nindices < 3) return FC_LBM_OK;
    if (!solid || !sdf) return VoxFail(FC_LBM_EINVAL, "null output arrays");
    const Frame f = MakeFrame(positions, nverts, *params);
    const long long N = static_cast<long long>(f.nx) * f.ny * f.nz;
    if (N >= (1LL << 32) - 1) return VoxFail(FC_LBM_EINVAL, "grid too large (cells must fit 32-bit indices)");
    if (device_out && capacity < N)
        return VoxFail(FC_LBM_EINVAL, "output buffers hold " + std::to_string(capacity) + " cells, the frame has " +
                                          std::to_string(N));
    const float radius = f.cs * 0.8660254f, tol = f.cs * 1e-3f;

    // triangle setup and brick index (host, triangle order)
    const int bx = (f.nx + kBrick - 1) / kBrick, by = (f.ny + kBrick - 1) / kBrick, bz = (f.nz + kBrick - 1) / kBrick;
    const long long nbricks = static_cast<long long>(bx) * by * bz;
    std::vector<TriRec> recs;
    recs.reserve(static_cast<size_t>(nindices / 3));
    for (long long t = 0; t + 2 < nindices; t += 3) {
        auto V = [&](uint32_t k) { return V3{positions[3 * k], positions[3 * k + 1], positions[3 * k + 2]}; };
        const V3 a = V(indices[t]), b = V(indices[t + 1]), c = V(indices[t + 2]);
        const V3 nraw = Cross(Sub(b, a), Sub(c, a));
        const float nlen = Len(nraw);
        if (nlen <= 1e-12f) continue;
        const V3 n = Mul(nraw, 1.0f / nlen);
        const V3 tmin{std::min({a.x, b.x, c.x}), std::min({a.y, b.y, c.y}), std::min({a.z, b.z, c.z})};
        const V3 tmax{std::max({a.x, b.x, c.x}), std::max({a.y, b.y, c.y}), std::max({a.z, b.z, c.z})};
        TriRec r{};
        r.box[0] = std::max(static_cast<int>(std::floor((tmin.y - radius - f.o.y) / f.cs)), 0);
        r.box[2] = std::max(static_cast<int>(std::floor((tmin.x - radius - f.o.x) / f.cs)), 0);
        r.box[4] = std::max(static_cast<int>(std::floor((tmin.z - radius - f.o.z) / f.cs)), 0);
        r.box[1] = std::min(static_cast<int>(std::ceil((tmax.y + radius - f.o.y) / f.cs)), f.nx - 1);
        r.box[3] = std::min(static_cast<int>(std::ceil((tmax.x + radius - f.o.x) / f.cs)), f.ny - 1);
        r.box[5] = std::min(static_cast<int>(std::ceil((tmax.z + radius - f.o.z) / f.cs)), f.nz - 1);
        if (r.box[0] > r.box[1] || r.box[2] > r.box[3] || r.box[4] > r.box[5]) continue;
        const float v[9] = {a.x, a.y, a.z, b.x, b.y, b.z, c.x, c.y, c.z};
        std::memcpy(r.v, v, sizeof(v));
        r.n[0] = n.x;
        r.n[1] = n.y;
        r.n[2] = n.z;
        recs.push_back(r);
    }
    std::vector<uint32_t> start(static_cast<size_t>(nbricks) + 1, 0);
    auto for_bricks = [&](const TriRec& r, auto&& fn) {
        for (int k = r.box[4] / kBrick; k <= r.box[5] / kBrick; ++k)
            for (int j = r.box[2] / kBrick; j <= r.box[3] / kBrick; ++j)
                for (int i = r.box[0] / kBrick; i <= r.box[1] / kBrick; ++i)
                    fn((static_cast<size_t>(k) * by + j) * bx + i);
    };
    unsigned long long pairs = 0;
    for (const TriRec& r : recs)
        for_bricks(r, [&](size_t b) {
            ++start[b + 1];
            ++pairs;
        });
    if (pairs >= (1ULL << 32)) return VoxFail(FC_LBM_EINVAL, "too many triangle-brick pairs");
    for (size_t b = 0; b < static_cast<size_t>(nbricks); ++b) start[b + 1] += start[b];
    std::vector<uint32_t> list(std::max<unsigned long long>(pairs, 1));
    {
        std::vector<uint32_t> fill(start.begin(), start.end() - 1);
        for (size_t t = 0; t < recs.size(); ++t) for_bricks(recs[t], [&](size_t b) { list[fill[b]++] = t; });
    }
    std::vector<uint32_t> active;
    for (size_t b = 0; b < static_cast<size_t>(nbricks); ++b)
        if (start[b + 1] > start[b]) active.push_back(static_cast<uint32_t>(b));

    DevBuf<uint8_t> own_solid;
    DevBuf<float> own_sdf;
    DevBuf<uint32_t> dparent, dstart, dlist, dactive;
    DevBuf<TriRec> dtris;
    if (!device_out) {
        VOX_HIP(hipMalloc(&own_solid.p, N));
        VOX_HIP(hipMalloc(&own_sdf.p, N * sizeof(float)));
    }
    uint8_t* const dsolid = device_out ? solid : own_solid.p;
    float* const dsdf = device_out ? sdf : own_sdf.p;
    VOX_HIP(hipMalloc(&dparent.p, N * sizeof(uint32_t)));
    VOX_HIP(hipMalloc(&dtris.p, std::max<size_t>(recs.size(), 1) * sizeof(TriRec)));
    VOX_HIP(hipMalloc(&dstart.p, start.size() * sizeof(uint32_t)));
    VOX_HIP(hipMalloc(&dlist.p, list.size() * sizeof(uint32_t)));
    VOX_HIP(hipMalloc(&dactive.p, std::max<size_t>(active.size(), 1) * sizeof(uint32_t)));
    if (!recs.empty()) VOX_HIP(hipMemcpy(dtris.p, recs.data(), recs.size() * sizeof(TriRec), hipMemcpyHostToDevice));
    VOX_HIP(hipMemcpy(dstart.p, start.data(), start.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    VOX_HIP(hipMemcpy(dlist.p, list.data(), list.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    if (!active.empty())
        VOX_HIP(hipMemcpy(dactive.p, active.data(), active.size() * sizeof(uint32_t), hipMemcpyHostToDevice));

    const unsigned cells_blocks = static_cast<unsigned>((N + 255) / 256);
    hipLaunchKernelGGL(KVoxInit, dim3(cells_blocks), dim3(256), 0, 0, dsolid, dsdf, dparent.p, N);
    const VoxArgs va{f.nx, f.ny, f.nz, f.cs, f.o.x, f.o.y, f.o.z, radius, tol, bx, by};
    if (!active.empty())
        hipLaunchKernelGGL(KVoxRaster, dim3(static_cast<unsigned>(active.size())), dim3(512), 0, 0, va, dtris.p,
                           dstart.p, dlist.p, dactive.p, dsolid, dsdf);
    const long long rows = static_cast<long long>(f.ny) * f.nz;
    hipLaunchKernelGGL(KUfRuns, dim3(static_cast<unsigned>((rows + 255) / 256)), dim3(256), 0, 0, dsolid, dparent.p,
                       f.nx, rows);
    hipLaunchKernelGGL(KUfHook, dim3(cells_blocks), dim3(256), 0, 0, dsolid, dparent.p, f.nx, f.ny, f.nz);
    hipLaunchKernelGGL(KUfCompress, dim3(cells_blocks), dim3(256), 0, 0, dsolid, dparent.p, N);
    hipLaunchKernelGGL(KVoxFinal, dim3(cells_blocks), dim3(256), 0, 0, dsolid, dsdf, dparent.p, N, radius);
    VOX_HIP(hipGetLastError());
    if (device_out) {
        VOX_HIP(hipStreamSynchronize(0));
    } else {
        VOX_HIP(hipMemcpy(solid, dsolid, N, hipMemcpyDeviceToHost));
        VOX_HIP(hipMemcpy(sdf, dsdf, N * sizeof(float), hipMemcpyDeviceToHost));
    }
    return FC_LBM_OK;
}

int fc_voxelize(const float* positions, long long nverts, const uint32_t* indices, long long nindices,
                const fc_vox_params* params, uint8_t* solid, float* sdf) {
    return Voxelize(positions, nverts, indices, nindices, params, solid, sdf, false);
}

int fc_voxelize_device(const float* positions, long long nverts, const uint32_t* indices, long long nindices,
                       const fc_vox_params* params, uint8_t* solid_device, float* sdf_device, long long capacity) {
    return Voxelize(positions, nverts, indices, nindices, params, solid_device, sdf_device, true, capacity);
}

}  // extern "C"
