// hbm_copy: HBM copy bandwidth of the current device, and the known-size reference kernel
// that scales rocprofv3's FETCH_SIZE / WRITE_SIZE counters to bytes (the step kernel's
// measured traffic is reported in those scaled units).  A standalone program, not part of
// the engine's libraries.
//
//   hbm_copy [GiB per buffer (default 2)] [repeats (default 20)]
//
// Each kernel reads one buffer and writes another of the same size, one element per lane:
//   copy16    16 bytes per lane (four floats)
//   copy16nt  the same with non-temporal stores
//   copy4      4 bytes per lane
// Prints one line per kernel: bytes moved per launch, average launch time and GB/s (HIP
// events around `repeats` launches after one warm-up launch).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                                \
    do {                                                                        \
        hipError_t e_ = (x);                                                    \
        if (e_ != hipSuccess) {                                                 \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));        \
            std::exit(1);                                                       \
        }                                                                       \
    } while (0)

typedef float f32x4 __attribute__((ext_vector_type(4)));  // 16 bytes per lane

template <class T, bool NT>
__global__ void KCopy(const T* __restrict__ a, T* __restrict__ b, long long n) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if constexpr (NT)
        __builtin_nontemporal_store(a[i], &b[i]);
    else
        b[i] = a[i];
}

template <class T, bool NT>
static float TimeCopy(const void* a, void* b, long long n, int reps) {
    const dim3 grid(static_cast<unsigned>((n + 255) / 256)), block(256);
    hipLaunchKernelGGL((KCopy<T, NT>), grid, block, 0, 0, static_cast<const T*>(a), static_cast<T*>(b), n);
    CHECK(hipGetLastError());
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    CHECK(hipEventRecord(e0, 0));
    for (int r = 0; r < reps; ++r)
        hipLaunchKernelGGL((KCopy<T, NT>), grid, block, 0, 0, static_cast<const T*>(a), static_cast<T*>(b), n);
    CHECK(hipEventRecord(e1, 0));
    CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    CHECK(hipEventDestroy(e0));
    CHECK(hipEventDestroy(e1));
    return ms / reps;
}

int main(int argc, char** argv) {
    const double gib = argc > 1 ? std::atof(argv[1]) : 2.0;
    const int reps = argc > 2 ? std::atoi(argv[2]) : 20;
    if (!(gib > 0.0) || gib > 64.0 || reps < 1) {
        std::fprintf(stderr, "usage: hbm_copy [GiB per buffer, (0, 64]] [repeats >= 1]\n");
        return 2;
    }
    const size_t bytes = static_cast<size_t>(gib * (1ull << 30)) / 256 * 256;
    void *a = nullptr, *b = nullptr;
    CHECK(hipMalloc(&a, bytes));
    CHECK(hipMalloc(&b, bytes));
    CHECK(hipMemset(a, 0, bytes));
    CHECK(hipMemset(b, 0, bytes));
    auto report = [&](const char* name, float ms) {
        const double moved = 2.0 * static_cast<double>(bytes);
        std::printf("%-9s bytes/launch %.0f  avg %.4f ms  %.1f GB/s\n", name, moved, ms, moved / (ms * 1e-3) / 1e9);
    };
    report("copy16", TimeCopy<f32x4, false>(a, b, static_cast<long long>(bytes / 16), reps));
    report("copy16nt", TimeCopy<f32x4, true>(a, b, static_cast<long long>(bytes / 16), reps));
    report("copy4", TimeCopy<float, false>(a, b, static_cast<long long>(bytes / 4), reps));
    CHECK(hipFree(a));
    CHECK(hipFree(b));
    return 0;
}
