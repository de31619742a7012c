// hbm_probe: measures HBM copy bandwidth on the current device, and serves as the
// known-size reference kernel for calibrating rocprofv3 FETCH_SIZE / WRITE_SIZE.
//
//   hbm_probe [GiB per buffer (default 2)] [repeats (default 20)]
//
// Kernels (each reads one buffer and writes another of the same size):
//   copy16    16 bytes per lane (float4)
//   copy4      4 bytes per lane
//   copy2      2 bytes per lane
//   pull19_4   the D3Q19 pull pattern without arithmetic: 19 SoA planes of 4-byte
//              values, each lane reads plane d at x - e_d and writes plane d at x
//   pull19_2   same with 2-byte values
//   pull19a_4  pull19_4 without the x shift (every load 256-byte aligned)
//   pull19x4_4 pull19_4 with four consecutive cells per lane (16-byte loads and stores)
//   pull19nt_4 pull19_4 with non-temporal stores
//   blkB_4     pull19_4 on a blocked layout (runs of B cells keep their 19 populations
//              together); blk64nt_4 with non-temporal stores
//   blkflag_4  blk64nt_4 plus a one-byte-per-cell flag stream; blksep_4 all loads before
//              the stores; blkflagsep_4 both (the step kernel's access pattern)
//   blk16_2    blkflagsep_4 with 2-byte values; blk16c2_2 with two cells of consecutive
//              rows per thread (the 16-bit step kernel's decomposition)
//   ip_blk64_4 in place on the blocked layout: each plane read and written at the cell
//   ip_esoteric_4  in place with the esoteric-pull shifts (half the planes read, half
//              written, one cell to the right)
// Prints one line per kernel: bytes moved per launch and GB/s (HIP events).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                                          \
    do {                                                                                  \
        hipError_t e = (x);                                                               \
        if (e != hipSuccess) {                                                            \
            std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e));           \
            std::exit(1);                                                                 \
        }                                                                                 \
    } while (0)

template <class T>
__global__ void KCopy(const T* __restrict__ a, T* __restrict__ b, long long n) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__constant__ int kEx[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
__constant__ int kEy[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
__constant__ int kEz[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};

// Interior cells of an nx*ny*nz block with a one-cell shell in y and z (so every source
// index stays inside the plane); x sources wrap inside the row.
template <class T>
__global__ void KPull19(const T* __restrict__ a, T* __restrict__ b, int nx, int ny, int nz, long long P) {
    const long long lin = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (lin >= static_cast<long long>(nx) * ny * nz) return;
    const int x = static_cast<int>(lin % nx);
    const int y = static_cast<int>((lin / nx) % ny);
    const int z = static_cast<int>(lin / (static_cast<long long>(nx) * ny));
    const int sy = nx, sz = nx * (ny + 2);
    const int self = (z + 1) * sz + (y + 1) * sy + x;
    float acc = 0.f;
#pragma unroll
    for (int d = 0; d < 19; ++d) {
        int xs = x - kEx[d];
        xs = xs < 0 ? xs + nx : (xs >= nx ? xs - nx : xs);
        const int src = (z + 1 - kEz[d]) * sz + (y + 1 - kEy[d]) * sy + xs;
        acc += static_cast<float>(a[d * P + src]);
        b[d * P + self] = static_cast<T>(acc);
    }
}

// As KPull19 with the x shift dropped (kAligned) and/or four cells per lane.
template <bool kAligned>
__global__ void KPull19Var(const float* __restrict__ a, float* __restrict__ b, int nx, int ny, int nz, long long P) {
    const long long lin = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (lin >= static_cast<long long>(nx) * ny * nz) return;
    const int x = static_cast<int>(lin % nx);
    const int y = static_cast<int>((lin / nx) % ny);
    const int z = static_cast<int>(lin / (static_cast<long long>(nx) * ny));
    const int sy = nx, sz = nx * (ny + 2);
    const int self = (z + 1) * sz + (y + 1) * sy + x;
    float acc = 0.f;
#pragma unroll
    for (int d = 0; d < 19; ++d) {
        int xs = kAligned ? x : x - kEx[d];
        xs = xs < 0 ? xs + nx : (xs >= nx ? xs - nx : xs);
        const int src = (z + 1 - kEz[d]) * sz + (y + 1 - kEy[d]) * sy + xs;
        acc += a[d * P + src];
        b[d * P + self] = acc;
    }
}

// Four consecutive cells per lane; the shifted loads are 4-byte aligned 16-byte loads
// (rows are read one element early or late, staying inside the padded plane).
__global__ void KPull19x4(const float* __restrict__ a, float* __restrict__ b, int nx, int ny, int nz, long long P) {
    const int gx = nx / 4;
    const long long lin = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (lin >= static_cast<long long>(gx) * ny * nz) return;
    const int x = static_cast<int>(lin % gx) * 4;
    const int y = static_cast<int>((lin / gx) % ny);
    const int z = static_cast<int>(lin / (static_cast<long long>(gx) * ny));
    const int sy = nx, sz = nx * (ny + 2);
    const int self = (z + 1) * sz + (y + 1) * sy + x;
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int d = 0; d < 19; ++d) {
        const int src = (z + 1 - kEz[d]) * sz + (y + 1 - kEy[d]) * sy + x - kEx[d];
        float4 v;
        __builtin_memcpy(&v, a + d * P + src, sizeof(v));
        acc.x += v.x;
        acc.y += v.y;
        acc.z += v.z;
        acc.w += v.w;
        *reinterpret_cast<float4*>(b + d * P + self) = acc;
    }
}

// The D3Q19 pull pattern on a blocked layout: runs of BX consecutive x cells keep their
// 19 populations together, element (d, x, y, z) at
//   ((((z+1)*(ny+2) + y+1) * (nx/BX) + x/BX) * 19 + d) * BX + x%BX,
// so one wave's 19 loads of its own run fall in one 19*BX*4-byte region instead of 19
// planes.  NT: non-temporal stores.
template <int BX, bool NT>
__global__ void KPull19Blk(const float* __restrict__ a, float* __restrict__ b, int nx, int ny, int nz) {
    const long long lin = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (lin >= static_cast<long long>(nx) * ny * nz) return;
    const int x = static_cast<int>(lin % nx);
    const int y = static_cast<int>((lin / nx) % ny);
    const int z = static_cast<int>(lin / (static_cast<long long>(nx) * ny));
    const int nb = nx / BX;
    auto idx = [&](int d, int xx, int yy, int zz) -> long long {
        return ((static_cast<long long>(zz + 1) * (ny + 2) + (yy + 1)) * nb + xx / BX) * (19 * BX) + d * BX + xx % BX;
    };
    float acc = 0.f;
#pragma unroll
    for (int d = 0; d < 19; ++d) {
        int xs = x - kEx[d];
        xs = xs < 0 ? xs + nx : (xs >= nx ? xs - nx : xs);
        acc += a[idx(d, xs, y - kEy[d], z - kEz[d])];
        if (NT)
            __builtin_nontemporal_store(acc, &b[idx(d, x, y, z)]);
        else
            b[idx(d, x, y, z)] = acc;
    }
}

// In-place variants on the blocked layout (one buffer, non-temporal stores): plane d of
// the run is read and written at the cell itself (SHIFT = false), or — the esoteric-pull
// pattern — the odd planes are read at the cell and written one cell to the right and the
// even planes read one cell to the right and written at the cell (SHIFT = true).  Every
// slot is read and then written by one lane; values are meaningless (timing only).
template <bool SHIFT>
__global__ void KInPlaceBlk(float* __restrict__ a, int nx, int ny, int nz) {
    const long long lin = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (lin >= static_cast<long long>(nx) * ny * nz) return;
    const int x = static_cast<int>(lin % nx);
    const int y = static_cast<int>((lin / nx) % ny);
    const int z = static_cast<int>(lin / (static_cast<long long>(nx) * ny));
    const int nb = nx / 64;
    auto idx = [&](int d, int xx) -> long long {
        return ((static_cast<long long>(z + 1) * (ny + 2) + (y + 1)) * nb + xx / 64) * (19 * 64) + d * 64 + xx % 64;
    };
    const int xr = x + 1 < nx ? x + 1 : 0;
    float v[19];
#pragma unroll
    for (int d = 0; d < 19; ++d) v[d] = a[idx(d, SHIFT && d != 0 && (d & 1) == 0 ? xr : x)];
    float acc = 0.f;
#pragma unroll
    for (int d = 0; d < 19; ++d) {
        acc += v[d];
        __builtin_nontemporal_store(acc, &a[idx(d, SHIFT && (d & 1) ? xr : x)]);
    }
}

// KPull19Blk<64, true> variants closer to the step kernel: FLAG also reads one byte per
// cell from a separate array (the step's cell-type flags); SEP issues all 19 loads before
// the 19 stores (the step's order: pull, collide, store).
template <bool FLAG, bool SEP>
__global__ void KPull19BlkX(const float* __restrict__ a, float* __restrict__ b, const unsigned char* __restrict__ fl,
                            int nx, int ny, int nz) {
    const long long lin = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (lin >= static_cast<long long>(nx) * ny * nz) return;
    const int x = static_cast<int>(lin % nx);
    const int y = static_cast<int>((lin / nx) % ny);
    const int z = static_cast<int>(lin / (static_cast<long long>(nx) * ny));
    const int nb = nx / 64;
    auto idx = [&](int d, int xx, int yy, int zz) -> long long {
        return ((static_cast<long long>(zz + 1) * (ny + 2) + (yy + 1)) * nb + xx / 64) * (19 * 64) + d * 64 + xx % 64;
    };
    float v[19];
#pragma unroll
    for (int d = 0; d < 19; ++d) {
        int xs = x - kEx[d];
        xs = xs < 0 ? xs + nx : (xs >= nx ? xs - nx : xs);
        v[d] = a[idx(d, xs, y - kEy[d], z - kEz[d])];
        if (!SEP) __builtin_nontemporal_store(v[d] * 0.5f, &b[idx(d, x, y, z)]);
    }
    float f = FLAG ? static_cast<float>(fl[lin]) : 0.0f;
    if (SEP) {
        float acc = f;
#pragma unroll
        for (int d = 0; d < 19; ++d) acc += v[d];
#pragma unroll
        for (int d = 0; d < 19; ++d) __builtin_nontemporal_store(v[d] - acc * 0.05f, &b[idx(d, x, y, z)]);
    } else if (FLAG && f != 0.0f) {
        b[idx(0, x, y, z)] = f;
    }
}

// The step kernel's pattern with 2-byte values on the blocked layout: each thread pulls
// CPT cells of consecutive rows (all loads first), reads their flag bytes, then stores
// with non-temporal stores.
template <int CPT>
__global__ void KPull19Blk16(const unsigned short* __restrict__ a, unsigned short* __restrict__ b,
                             const unsigned char* __restrict__ fl, int nx, int ny, int nz) {
    const long long lin = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int nyc = ny / CPT;
    if (lin >= static_cast<long long>(nx) * nyc * nz) return;
    const int x = static_cast<int>(lin % nx);
    const int y0 = static_cast<int>((lin / nx) % nyc) * CPT;
    const int z = static_cast<int>(lin / (static_cast<long long>(nx) * nyc));
    const int nb = nx / 64;
    auto idx = [&](int d, int xx, int yy, int zz) -> long long {
        return ((static_cast<long long>(zz + 1) * (ny + 2) + (yy + 1)) * nb + xx / 64) * (19 * 64) + d * 64 + xx % 64;
    };
    unsigned short v[CPT][19];
#pragma unroll
    for (int c = 0; c < CPT; ++c)
#pragma unroll
        for (int d = 0; d < 19; ++d) {
            int xs = x - kEx[d];
            xs = xs < 0 ? xs + nx : (xs >= nx ? xs - nx : xs);
            v[c][d] = a[idx(d, xs, y0 + c - kEy[d], z - kEz[d])];
        }
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
        const unsigned f = fl[(static_cast<long long>(z) * ny + y0 + c) * nx + x];
        unsigned acc = f;
#pragma unroll
        for (int d = 0; d < 19; ++d) acc += v[c][d];
#pragma unroll
        for (int d = 0; d < 19; ++d)
            __builtin_nontemporal_store(static_cast<unsigned short>(v[c][d] ^ (acc & 1u)), &b[idx(d, x, y0 + c, z)]);
    }
}

// KPull19<float> with non-temporal stores.
__global__ void KPull19Nt(const float* __restrict__ a, float* __restrict__ b, int nx, int ny, int nz, long long P) {
    const long long lin = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (lin >= static_cast<long long>(nx) * ny * nz) return;
    const int x = static_cast<int>(lin % nx);
    const int y = static_cast<int>((lin / nx) % ny);
    const int z = static_cast<int>(lin / (static_cast<long long>(nx) * ny));
    const int sy = nx, sz = nx * (ny + 2);
    const int self = (z + 1) * sz + (y + 1) * sy + x;
    float acc = 0.f;
#pragma unroll
    for (int d = 0; d < 19; ++d) {
        int xs = x - kEx[d];
        xs = xs < 0 ? xs + nx : (xs >= nx ? xs - nx : xs);
        const int src = (z + 1 - kEz[d]) * sz + (y + 1 - kEy[d]) * sy + xs;
        acc += a[d * P + src];
        __builtin_nontemporal_store(acc, &b[d * P + self]);
    }
}

template <class F>
float TimeMs(F launch, int reps) {
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    launch();
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(e0));
    for (int r = 0; r < reps; ++r) launch();
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    CHECK(hipEventDestroy(e0));
    CHECK(hipEventDestroy(e1));
    return ms / reps;
}

int main(int argc, char** argv) {
    const double gib = argc > 1 ? std::atof(argv[1]) : 2.0;
    const int reps = argc > 2 ? std::atoi(argv[2]) : 20;
    const size_t bytes = static_cast<size_t>(gib * (1ull << 30)) / 256 * 256;
    void *a, *b;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMemset(a, 0, bytes));
    CHECK(hipMemset(b, 0, bytes));
    auto report = [&](const char* name, double moved, float ms) {
        std::printf("%-9s bytes/launch %.0f  avg %.4f ms  %.1f GB/s\n", name, moved, ms, moved / (ms * 1e-3) / 1e9);
    };
    {
        const long long n = bytes / 16;
        float ms = TimeMs([&] { hipLaunchKernelGGL(KCopy<float4>, dim3((n + 255) / 256), dim3(256), 0, 0,
                                                   (const float4*)a, (float4*)b, n); }, reps);
        report("copy16", 2.0 * bytes, ms);
    }
    {
        const long long n = bytes / 4;
        float ms = TimeMs([&] { hipLaunchKernelGGL(KCopy<float>, dim3((n + 255) / 256), dim3(256), 0, 0,
                                                   (const float*)a, (float*)b, n); }, reps);
        report("copy4", 2.0 * bytes, ms);
    }
    {
        const long long n = bytes / 2;
        float ms = TimeMs([&] { hipLaunchKernelGGL(KCopy<unsigned short>, dim3((n + 255) / 256), dim3(256), 0, 0,
                                                   (const unsigned short*)a, (unsigned short*)b, n); }, reps);
        report("copy2", 2.0 * bytes, ms);
    }
    // pull pattern on a 512 x 256 x nz block sized to the buffer
    for (int s : {4, 2}) {
        const int nx = 512, ny = 256;
        const long long plane_pad = static_cast<long long>(nx) * (ny + 2);
        const int nz = static_cast<int>(bytes / s / 19 / plane_pad) - 2;
        const long long P = plane_pad * (nz + 2);
        const long long cells = static_cast<long long>(nx) * ny * nz;
        float ms;
        if (s == 4)
            ms = TimeMs([&] { hipLaunchKernelGGL(KPull19<float>, dim3((cells + 255) / 256), dim3(256), 0, 0,
                                                 (const float*)a, (float*)b, nx, ny, nz, P); }, reps);
        else
            ms = TimeMs([&] { hipLaunchKernelGGL(KPull19<_Float16>, dim3((cells + 255) / 256), dim3(256), 0, 0,
                                                 (const _Float16*)a, (_Float16*)b, nx, ny, nz, P); }, reps);
        report(s == 4 ? "pull19_4" : "pull19_2", 2.0 * 19 * s * cells, ms);
        if (s == 2) {
            unsigned char* fl;
            CHECK(hipMalloc(&fl, cells));
            CHECK(hipMemset(fl, 0, cells));
            ms = TimeMs([&] { hipLaunchKernelGGL(KPull19Blk16<1>, dim3((cells + 255) / 256), dim3(256), 0, 0,
                                                 (const unsigned short*)a, (unsigned short*)b, fl, nx, ny, nz); }, reps);
            report("blk16_2", 2.0 * 19 * s * cells, ms);
            ms = TimeMs([&] { hipLaunchKernelGGL(KPull19Blk16<2>, dim3((cells / 2 + 255) / 256), dim3(256), 0, 0,
                                                 (const unsigned short*)a, (unsigned short*)b, fl, nx, ny, nz); }, reps);
            report("blk16c2_2", 2.0 * 19 * s * cells, ms);
            CHECK(hipFree(fl));
        }
        if (s == 4) {
            ms = TimeMs([&] { hipLaunchKernelGGL(KPull19Var<true>, dim3((cells + 255) / 256), dim3(256), 0, 0,
                                                 (const float*)a, (float*)b, nx, ny, nz, P); }, reps);
            report("pull19a_4", 2.0 * 19 * s * cells, ms);
            ms = TimeMs([&] { hipLaunchKernelGGL(KPull19x4, dim3((cells / 4 + 255) / 256), dim3(256), 0, 0,
                                                 (const float*)a, (float*)b, nx, ny, nz, P); }, reps);
            report("pull19x4_4", 2.0 * 19 * s * cells, ms);
            ms = TimeMs([&] { hipLaunchKernelGGL(KPull19Nt, dim3((cells + 255) / 256), dim3(256), 0, 0,
                                                 (const float*)a, (float*)b, nx, ny, nz, P); }, reps);
            report("pull19nt_4", 2.0 * 19 * s * cells, ms);
            // blocked layouts: the same element count (nx*(ny+2)*(nz+2)*19) as the planes
            ms = TimeMs([&] { hipLaunchKernelGGL((KPull19Blk<64, false>), dim3((cells + 255) / 256), dim3(256), 0, 0,
                                                 (const float*)a, (float*)b, nx, ny, nz); }, reps);
            report("blk64_4", 2.0 * 19 * s * cells, ms);
            ms = TimeMs([&] { hipLaunchKernelGGL((KPull19Blk<64, true>), dim3((cells + 255) / 256), dim3(256), 0, 0,
                                                 (const float*)a, (float*)b, nx, ny, nz); }, reps);
            report("blk64nt_4", 2.0 * 19 * s * cells, ms);
            ms = TimeMs([&] { hipLaunchKernelGGL((KPull19Blk<256, false>), dim3((cells + 255) / 256), dim3(256), 0, 0,
                                                 (const float*)a, (float*)b, nx, ny, nz); }, reps);
            report("blk256_4", 2.0 * 19 * s * cells, ms);
            ms = TimeMs([&] { hipLaunchKernelGGL((KPull19Blk<512, false>), dim3((cells + 255) / 256), dim3(256), 0, 0,
                                                 (const float*)a, (float*)b, nx, ny, nz); }, reps);
            report("blk512_4", 2.0 * 19 * s * cells, ms);
            {
                unsigned char* fl;
                CHECK(hipMalloc(&fl, cells));
                CHECK(hipMemset(fl, 0, cells));
                ms = TimeMs([&] { hipLaunchKernelGGL((KPull19BlkX<true, false>), dim3((cells + 255) / 256), dim3(256), 0, 0,
                                                     (const float*)a, (float*)b, fl, nx, ny, nz); }, reps);
                report("blkflag_4", 2.0 * 19 * s * cells, ms);
                ms = TimeMs([&] { hipLaunchKernelGGL((KPull19BlkX<false, true>), dim3((cells + 255) / 256), dim3(256), 0, 0,
                                                     (const float*)a, (float*)b, fl, nx, ny, nz); }, reps);
                report("blksep_4", 2.0 * 19 * s * cells, ms);
                ms = TimeMs([&] { hipLaunchKernelGGL((KPull19BlkX<true, true>), dim3((cells + 255) / 256), dim3(256), 0, 0,
                                                     (const float*)a, (float*)b, fl, nx, ny, nz); }, reps);
                report("blkflagsep_4", 2.0 * 19 * s * cells, ms);
                CHECK(hipFree(fl));
            }
            ms = TimeMs([&] { hipLaunchKernelGGL((KInPlaceBlk<false>), dim3((cells + 255) / 256), dim3(256), 0, 0,
                                                 (float*)a, nx, ny, nz); }, reps);
            report("ip_blk64_4", 2.0 * 19 * s * cells, ms);
            ms = TimeMs([&] { hipLaunchKernelGGL((KInPlaceBlk<true>), dim3((cells + 255) / 256), dim3(256), 0, 0,
                                                 (float*)a, nx, ny, nz); }, reps);
            report("ip_esoteric_4", 2.0 * 19 * s * cells, ms);
        }
    }
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    return 0;
}
