// Minimal stand-alone definition of FlowCraft's VoxelGrid (FlowCraft/VoxelGrid.h:13-61)
// with the members the solver reads, for builds of LbmSolver.h outside FlowCraft;
// inside FlowCraft the application's own header is found first on the include path.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#ifndef FLOWCRAFT_HAS_VEC3
struct Vec3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
    constexpr Vec3() = default;
    constexpr Vec3(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}
};
#endif

struct VoxelGrid {
    int nx = 0, ny = 0, nz = 0;
    float cellSize = 1.0f;
    Vec3 origin{0.0f, 0.0f, 0.0f};
    std::vector<uint8_t> solid;         // nx*ny*nz, 1 = solid
    std::vector<float> signedDistance;  // same indexing, or empty (plain bounce-back)

    size_t CellCount() const { return static_cast<size_t>(nx) * ny * nz; }
    size_t Index(int x, int y, int z) const { return (static_cast<size_t>(z) * ny + y) * nx + x; }
};
