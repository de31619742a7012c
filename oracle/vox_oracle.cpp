// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// Clean-room CPU restatement of FlowCraft's mesh voxelizer, Voxelizer::Voxelize
// (FlowCraft/Voxelizer.cpp:128-432), checking the engine's device voxelizer
// (flowcraft_amd/csrc/voxelize.hip).  Single-threaded; every float operation in the
// reference's order (built with -ffp-contract=off), so the solid mask and the signed
// distance come out bit for bit:
//   grid frame from the padded mesh bounds ........ Voxelizer.cpp:137-166
//   conservative rasterization + nearest-triangle
//   signed distance with the edge tie-break ....... Voxelizer.cpp:167-293
//   closest point on a triangle (Ericson 5.1.5) ... Voxelizer.cpp:85-123
//   exterior flood fill from cell (0,0,0) ......... Voxelizer.cpp:295-412 (a plain
//                                                   BFS here: reachability is the same)
//   signed-distance reconciliation ................ Voxelizer.cpp:414-432
// Vector helpers as FlowCraft/Mesh.h:14-45 (Dot, Cross, Length, component-wise ops).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

namespace {

struct V3 {
    float x, y, z;
};
V3 Add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 Sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 Mul(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
float Dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 Cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
float Len(V3 v) { return std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }

// Nearest point of triangle (a, b, c) to p by Voronoi region: vertex, edge or face.
V3 Nearest(V3 p, V3 a, V3 b, V3 c) {
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

struct Frame {
    int nx = 0, ny = 0, nz = 0;
    float cs = 1.0f;
    V3 o{0, 0, 0};
};

Frame MakeFrame(const float* pos, long long nv, int target, const float* pad) {
    V3 lo{std::numeric_limits<float>::max(), std::numeric_limits<float>::max(), std::numeric_limits<float>::max()};
    V3 hi{std::numeric_limits<float>::lowest(), std::numeric_limits<float>::lowest(),
          std::numeric_limits<float>::lowest()};
    for (long long i = 0; i < nv; ++i) {
        lo = {std::min(lo.x, pos[3 * i]), std::min(lo.y, pos[3 * i + 1]), std::min(lo.z, pos[3 * i + 2])};
        hi = {std::max(hi.x, pos[3 * i]), std::max(hi.y, pos[3 * i + 1]), std::max(hi.z, pos[3 * i + 2])};
    }
    const V3 size = Sub(hi, lo);
    const V3 p{size.x * pad[0], size.y * pad[1], size.z * pad[2]};
    const V3 dmin = Sub(lo, p), dmax = Add(hi, p), ds = Sub(dmax, dmin);
    const float longest = std::max({ds.x, ds.y, ds.z});
    const int res = std::max(target, 4);
    Frame f;
    f.cs = std::max(longest / static_cast<float>(res), 1e-6f);
    f.o = dmin;
    // grid-x along world Y, grid-y along world X (VoxelGrid::CellCenter)
    f.nx = std::max(static_cast<int>(std::ceil(ds.y / f.cs)), 1);
    f.ny = std::max(static_cast<int>(std::ceil(ds.x / f.cs)), 1);
    f.nz = std::max(static_cast<int>(std::ceil(ds.z / f.cs)), 1);
    return f;
}

}  // namespace

extern "C" {

// Grid frame of a mesh: dims[3] = nx, ny, nz; frame[4] = cellSize, origin.  0 for an
// empty mesh (no vertices or fewer than 3 indices: the reference's empty grid).
int orc_vox_frame(const float* pos, long long nv, long long ni, int target, const float* pad, int* dims,
                  float* frame) {
    dims[0] = dims[1] = dims[2] = 0;
    frame[0] = 1.0f;
    frame[1] = frame[2] = frame[3] = 0.0f;
    if (nv <= 0 || ni < 3) return 0;
    const Frame f = MakeFrame(pos, nv, target, pad);
    dims[0] = f.nx;
    dims[1] = f.ny;
    dims[2] = f.nz;
    frame[0] = f.cs;
    frame[1] = f.o.x;
    frame[2] = f.o.y;
    frame[3] = f.o.z;
    return 1;
}

// Voxelizer::Voxelize into solid (N bytes) and sdf (N floats), N from orc_vox_frame.
void orc_voxelize(const float* pos, long long nv, const uint32_t* idx, long long ni, int target, const float* pad,
                  uint8_t* solid, float* sdf) {
    if (nv <= 0 || ni < 3) return;
    const Frame f = MakeFrame(pos, nv, target, pad);
    const size_t N = static_cast<size_t>(f.nx) * f.ny * f.nz;
    std::fill(solid, solid + N, uint8_t{0});
    std::fill(sdf, sdf + N, std::numeric_limits<float>::infinity());
    std::vector<float> proj(N, -1.0f);  // |normal component| of the triangle owning sdf[i]
    const float radius = f.cs * 0.8660254f, tol = f.cs * 1e-3f;
    auto vert = [&](uint32_t k) { return V3{pos[3 * k], pos[3 * k + 1], pos[3 * k + 2]}; };
    for (long long t = 0; t + 2 < ni; t += 3) {
        const V3 a = vert(idx[t]), b = vert(idx[t + 1]), c = vert(idx[t + 2]);
        const V3 nraw = Cross(Sub(b, a), Sub(c, a));
        const float nlen = Len(nraw);
        if (nlen <= 1e-12f) continue;
        const V3 n = Mul(nraw, 1.0f / nlen);
        const V3 tmin{std::min({a.x, b.x, c.x}), std::min({a.y, b.y, c.y}), std::min({a.z, b.z, c.z})};
        const V3 tmax{std::max({a.x, b.x, c.x}), std::max({a.y, b.y, c.y}), std::max({a.z, b.z, c.z})};
        const int x0 = std::max(static_cast<int>(std::floor((tmin.y - radius - f.o.y) / f.cs)), 0);
        const int y0 = std::max(static_cast<int>(std::floor((tmin.x - radius - f.o.x) / f.cs)), 0);
        const int z0 = std::max(static_cast<int>(std::floor((tmin.z - radius - f.o.z) / f.cs)), 0);
        const int x1 = std::min(static_cast<int>(std::ceil((tmax.y + radius - f.o.y) / f.cs)), f.nx - 1);
        const int y1 = std::min(static_cast<int>(std::ceil((tmax.x + radius - f.o.x) / f.cs)), f.ny - 1);
        const int z1 = std::min(static_cast<int>(std::ceil((tmax.z + radius - f.o.z) / f.cs)), f.nz - 1);
        for (int z = z0; z <= z1; ++z)
            for (int y = y0; y <= y1; ++y)
                for (int x = x0; x <= x1; ++x) {
                    const size_t i = (static_cast<size_t>(z) * f.ny + y) * f.nx + x;
                    const V3 ctr = Add(f.o, V3{(y + 0.5f) * f.cs, (x + 0.5f) * f.cs, (z + 0.5f) * f.cs});
                    const V3 diff = Sub(ctr, Nearest(ctr, a, b, c));
                    const float dist = Len(diff);
                    const float nd = Dot(diff, n);
                    const float and_ = std::fabs(nd);
                    const float best = std::fabs(sdf[i]);
                    if (dist + tol < best) {
                        sdf[i] = nd < 0.0f ? -dist : dist;
                        proj[i] = and_;
                    } else if (dist <= best + tol) {
                        // near-tie (a shared edge or vertex): the face seen more head-on decides the sign
                        if (and_ > proj[i]) {
                            const float m = std::min(dist, best);
                            sdf[i] = nd < 0.0f ? -m : m;
                            proj[i] = and_;
                        } else if (dist < best) {
                            sdf[i] = sdf[i] < 0.0f ? -dist : dist;
                        }
                    }
                    if (dist <= radius) solid[i] = 1;
                }
    }
    // exterior: 6-connected non-shell cells reachable from cell (0,0,0)
    std::vector<uint8_t> reached(N, 0);
    std::vector<size_t> stack;
    if (!solid[0]) {
        reached[0] = 1;
        stack.push_back(0);
    }
    const size_t layer = static_cast<size_t>(f.nx) * f.ny;
    while (!stack.empty()) {
        const size_t i = stack.back();
        stack.pop_back();
        const int z = static_cast<int>(i / layer), y = static_cast<int>((i % layer) / f.nx),
                  x = static_cast<int>(i % f.nx);
        const int nb[6][3] = {{x - 1, y, z}, {x + 1, y, z}, {x, y - 1, z}, {x, y + 1, z}, {x, y, z - 1}, {x, y, z + 1}};
        for (const auto& q : nb) {
            if (q[0] < 0 || q[1] < 0 || q[2] < 0 || q[0] >= f.nx || q[1] >= f.ny || q[2] >= f.nz) continue;
            const size_t j = (static_cast<size_t>(q[2]) * f.ny + q[1]) * f.nx + q[0];
            if (solid[j] || reached[j]) continue;
            reached[j] = 1;
            stack.push_back(j);
        }
    }
    for (size_t i = 0; i < N; ++i) {
        if (!solid[i] && !reached[i]) solid[i] = 1;  // enclosed by the shell: interior
        const float d = std::fabs(sdf[i]);
        if (!solid[i])
            sdf[i] = d;
        else if (d > radius)
            sdf[i] = -d;
    }
}

}  // extern "C"
