// Lattice velocity sets, shared by host and device code.
//
// Same directions, order, weights and opposite pairing as FlowCraft's
// LatticeModel.h:20-64 (D3Q19 and D3Q15), so population index i means the same
// thing on both sides of the boundary.  Opposite directions are the consecutive
// pairs (1,2), (3,4), ...; the in-place layout relies on that pairing.
// Mirror tables are found by search, as LbmReference::Init does
// (LbmReference.cpp:172-183).
#pragma once

#include <hip/hip_runtime.h>

namespace fc {

template <int Q>
struct Lat;

template <>
struct Lat<19> {
    static constexpr int kQ = 19;
    static constexpr int ex[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    static constexpr int ey[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
    static constexpr int ez[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
    static constexpr float w[19] = {1.0f / 3.0f,  1.0f / 18.0f, 1.0f / 18.0f, 1.0f / 18.0f, 1.0f / 18.0f,
                                    1.0f / 18.0f, 1.0f / 18.0f, 1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f,
                                    1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f,
                                    1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f, 1.0f / 36.0f};
    static constexpr double wd[19] = {1.0 / 3.0,  1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
                                      1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                                      1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                                      1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
};

template <>
struct Lat<15> {
    static constexpr int kQ = 15;
    static constexpr int ex[15] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
    static constexpr int ey[15] = {0, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, -1, 1, 1, -1};
    static constexpr int ez[15] = {0, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, 1, -1, 1, -1};
    static constexpr float w[15] = {2.0f / 9.0f,  1.0f / 9.0f,  1.0f / 9.0f,  1.0f / 9.0f,  1.0f / 9.0f,
                                    1.0f / 9.0f,  1.0f / 9.0f,  1.0f / 72.0f, 1.0f / 72.0f, 1.0f / 72.0f,
                                    1.0f / 72.0f, 1.0f / 72.0f, 1.0f / 72.0f, 1.0f / 72.0f, 1.0f / 72.0f};
    static constexpr double wd[15] = {2.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,  1.0 / 9.0,
                                      1.0 / 9.0,  1.0 / 9.0,  1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0,
                                      1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0};
};

// Opposite direction: the consecutive pairing of both tables.
__host__ __device__ constexpr int Opp(int i) { return i == 0 ? 0 : ((i & 1) ? i + 1 : i - 1); }

template <int Q>
__host__ __device__ constexpr int MirrorY(int i) {
    int r = 0;
    for (int j = 0; j < Q; ++j)
        if (Lat<Q>::ex[j] == Lat<Q>::ex[i] && Lat<Q>::ey[j] == -Lat<Q>::ey[i] && Lat<Q>::ez[j] == Lat<Q>::ez[i]) r = j;
    return r;
}

template <int Q>
__host__ __device__ constexpr int MirrorZ(int i) {
    int r = 0;
    for (int j = 0; j < Q; ++j)
        if (Lat<Q>::ex[j] == Lat<Q>::ex[i] && Lat<Q>::ey[j] == Lat<Q>::ey[i] && Lat<Q>::ez[j] == -Lat<Q>::ez[i]) r = j;
    return r;
}

// Link kinds of the boundary table (per fluid cell and direction i, about the
// direction-i source cell x - e_i).
enum LinkKind : unsigned {
    kPull = 0,      // source is an in-domain fluid cell: ordinary streaming
    kPlainBB = 2,   // solid source without usable SDF data: halfway bounce-back
    kBouzidiLo = 3, // solid source, q < 1/2 and x + e_i is an in-domain fluid cell
    kBouzidiHi = 4, // solid source, q >= 1/2
};

// Per-cell flag bits (one byte per cell).
enum CellFlag : unsigned char {
    kSolid = 1,     // solid cell
    kNearSolid = 2, // fluid cell with at least one solid source: has a boundary record
    kSurface = 4,   // solid cell with at least one fluid neighbour: on the force list
};

}  // namespace fc
