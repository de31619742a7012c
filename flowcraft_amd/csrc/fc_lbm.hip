// Host side of libfc_lbm.so: the C ABI declared in include/flowcraft/fc_lbm.h.
//
// One handle = one lattice resident in HBM on the current HIP device.  All
// device work is enqueued on the handle's stream; only the readback/stats/probe
// calls synchronise.  See lbm_kernels.hpp for the data layout.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/flowcraft/fc_lbm.h"
#include "lbm_aux.hpp"
#include "lbm_kernels.hpp"
#include "step_launch.hpp"

using namespace fc;

// Staging of the allocation-free readback (fc_lbm_readback_field_range_noalloc): the
// field of a z range passes through this array in chunks of whole layers.  Each GPU holds
// its own copy for the life of the process, shared by every handle on that GPU.  Host code
// never names it directly: its device address is resolved per device at handle creation
// (ScratchAddress) and kept in the handle.
constexpr size_t kScratchBytes = size_t(32) << 20;
static __device__ __align__(256) float scratch[kScratchBytes / sizeof(float)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (which must be `device`), resolved
// once per device.
hipError_t ScratchAddress(int device, void** out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}
}  // namespace

struct fc_lbm {
    int nx = 0, ny = 0, nz = 0, Q = 19, storage = FC_LBM_STORE_F32, layout = FC_LBM_LAYOUT_PINGPONG;
    int pery = 0, perz = 0;
    long long N = 0, P = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    size_t tbytes = 4, rbytes = 4;  // storage element / compute element sizes
    void* pops[2] = {nullptr, nullptr};
    uint8_t* flags = nullptr;
    uint8_t* dsolid = nullptr;
    int* bidx = nullptr;
    void* recs = nullptr;
    long long* nearlist = nullptr;  // padded indices of the near-solid fluid cells (64-bit)
    int nrec = 0;
    long long* surf = nullptr;      // padded indices of the surface solid cells
    int nsurf = 0;
    void* side_mom = nullptr;
    void* sforce = nullptr;
    double* dforces = nullptr;
    double* stat_part = nullptr;
    double* stat_out = nullptr;
    double* aero_part = nullptr;  // per-block partials of the aerodynamic-force passes
    double* aero_out = nullptr;   // [0, 13) surface sums, [16, 21) geometric sums
    bool aero_geom_valid = false;  // geometric sums cached for aero_frame
    float aero_frame[4] = {0, 0, 0, 0};
    double aero_geom[kAeroGVals] = {0, 0, 0, 0, -1};
    int* dflag = nullptr;   // sticky divergence flag
    int* dprobe = nullptr;
    float* staging = nullptr;
    size_t staging_bytes = 0;
    float* scratch = nullptr;  // this device's copy of the file-scope `scratch` array
    // field snapshots: device copy + pinned host copy, moved on their own stream
    float* snap_dev = nullptr;
    float* snap_host = nullptr;
    hipStream_t copy_stream = nullptr;
    hipEvent_t snap_taken = nullptr;   // the device snapshot is complete (solver stream)
    hipEvent_t snap_copied = nullptr;  // the host copy is complete (copy stream)
    bool snap_pending = false;
    fc_lbm_params params{};
    long long t = 0;
    bool has_geometry = false;
    bool has_state = false;
    int cells_per_thread = 1;  // cells each step-kernel thread updates (1 or 2)
    bool resident = false;     // populations fit in the last-level cache: cached stores
    std::vector<uint8_t> solid_host;
    std::string err;
};

namespace {

const char* kNoHandle = "null fc_lbm handle";

// Lattices whose population buffers together fit in the 256 MB last-level cache (the
// Infinity Cache) keep their stores cached, since the next step pulls them back; larger
// ones stream the stores to HBM.  Measured on MI355X (ms per step, streamed -> cached):
// the periodic 128x64x64 sphere FP32 ping-pong (80 MB) 21.7 -> 19.2 us, 128x64x128 (167
// MB) 35.6 -> 29.3 us, 128^3 FP32 in place (164 MB) 57.1 -> 52.6 us; past the cache size
// cached stores lose (192x96x96 FP32 ping-pong, 280 MB: 58.8 -> 60.0 us; 256x128x128
// FP16 ping-pong, 328 MB: 67.3 -> 71.1 us).  FC_LBM_RESIDENT=0/1 overrides the choice.
constexpr long long kResidentBytes = 256LL << 20;
// Largest lattice, in cells: every one-work-item-per-cell launch stays below 2^32
// work-items (HIP's limit for one grid dimension), with room for a partial block.
constexpr long long kMaxCells = (1LL << 32) - (1LL << 20);
// Step launches are split by z range when one would exceed this many work-items.
constexpr long long kMaxStepItems = 1LL << 31;
constexpr size_t kPopMargin = 256;

// Elements of one population copy: the front margin, nz+2 padded layers of ny+2 rows of
// population blocks, then one block and kPopMargin elements (the in-place anchors and the
// step's speculative bulk pull of a cell at the far corner reach up to one block beyond
// the last row).
size_t PopElements(const fc_lbm* h) {
    const long long rs = PopRowStride(h->nx, h->Q);
    return static_cast<size_t>(PopFront(rs, h->Q) + rs * (h->ny + 2) * (h->nz + 2) + h->Q * kBlk + kPopMargin);
}

void DevFree(void* p) {
    if (p) (void)hipFree(p);
}

// A temporary device buffer freed on every return path, after the stream that may still
// read it has drained.
struct DevTmp {
    hipStream_t st;
    void* p = nullptr;
    ~DevTmp() {
        if (p) {
            (void)hipStreamSynchronize(st);
            (void)hipFree(p);
        }
    }
};

#define FC_HIP(h, call)                                                                         \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            (h)->err = std::string(#call) + ": " + hipGetErrorString(e_);                       \
            return FC_LBM_EDEVICE;                                                              \
        }                                                                                       \
    } while (0)

int Fail(fc_lbm* h, int code, const std::string& msg) {
    h->err = msg;
    return code;
}

template <int Q, class S, class F>
void DispatchL(const fc_lbm* h, F& f) {
    if (h->layout == FC_LBM_LAYOUT_INPLACE)
        f.template operator()<Q, S, InPlace>();
    else
        f.template operator()<Q, S, PingPong>();
}

// Storage dispatch of every kernel but the step: the stochastically rounded formats hold
// the same elements as their round-to-nearest twins and seed/restore/read back alike.
template <int Q, class F>
void DispatchS(const fc_lbm* h, F& f) {
    switch (h->storage) {
        case FC_LBM_STORE_F16:
        case FC_LBM_STORE_F16_SR: DispatchL<Q, StoreF16<Q>>(h, f); break;
        case FC_LBM_STORE_BF16:
        case FC_LBM_STORE_BF16_SR: DispatchL<Q, StoreBF16<Q>>(h, f); break;
        case FC_LBM_STORE_F64: DispatchL<Q, StoreF64<Q>>(h, f); break;
        default: DispatchL<Q, StoreF32<Q>>(h, f); break;
    }
}

template <class F>
void Dispatch(const fc_lbm* h, F&& f) {
    if (h->Q == 15)
        DispatchS<15>(h, f);
    else
        DispatchS<19>(h, f);
}

// The step kernel's dispatch, where the rounding of the stores differs.
template <int Q, class F>
void DispatchStepS(const fc_lbm* h, F& f) {
    switch (h->storage) {
        case FC_LBM_STORE_F16_SR: DispatchL<Q, StoreF16SR<Q>>(h, f); break;
        case FC_LBM_STORE_BF16_SR: DispatchL<Q, StoreBF16SR<Q>>(h, f); break;
        default: DispatchS<Q>(h, f); break;
    }
}

template <class F>
void DispatchStep(const fc_lbm* h, F&& f) {
    if (h->Q == 15)
        DispatchStepS<15>(h, f);
    else
        DispatchStepS<19>(h, f);
}

unsigned Blocks(long long n, int bs) { return static_cast<unsigned>((n + bs - 1) / bs); }

void FreeGeometry(fc_lbm* h) {
    DevFree(h->flags);
    DevFree(h->dsolid);
    DevFree(h->bidx);
    DevFree(h->recs);
    DevFree(h->nearlist);
    DevFree(h->surf);
    DevFree(h->sforce);
    h->flags = h->dsolid = nullptr;
    h->bidx = nullptr;
    h->nearlist = h->surf = nullptr;
    h->recs = h->sforce = nullptr;
    h->nrec = h->nsurf = 0;
    h->has_geometry = false;
    h->has_state = false;
}

// Deterministic compaction of the cells whose flag has `mask` (padded indices, in
// global cell order).  Allocates *out (at least one element).
int Compact(fc_lbm* h, uint8_t mask, long long** out, int* count) {
    const unsigned chunks = Blocks(h->N, kChunk);
    DevTmp counts_buf{h->stream};
    FC_HIP(h, hipMalloc(&counts_buf.p, sizeof(int) * chunks));
    int* dcounts = static_cast<int*>(counts_buf.p);
    hipLaunchKernelGGL(KCountFlags, dim3(chunks), dim3(256), 0, h->stream, h->flags, mask, h->nx, h->ny, h->nz,
                       dcounts);
    std::vector<int> counts(chunks);
    std::vector<long long> offs(chunks);
    FC_HIP(h, hipMemcpyAsync(counts.data(), dcounts, sizeof(int) * chunks, hipMemcpyDeviceToHost, h->stream));
    FC_HIP(h, hipStreamSynchronize(h->stream));
    long long total = 0;
    for (unsigned i = 0; i < chunks; ++i) {
        offs[i] = total;
        total += counts[i];
    }
    if (total >= (1LL << 31)) return Fail(h, FC_LBM_EINVAL, "more than 2^31 - 1 boundary cells");
    DevTmp offs_buf{h->stream};
    FC_HIP(h, hipMalloc(&offs_buf.p, sizeof(long long) * chunks));
    FC_HIP(h, hipMemcpyAsync(offs_buf.p, offs.data(), sizeof(long long) * chunks, hipMemcpyHostToDevice, h->stream));
    FC_HIP(h, hipMalloc(out, sizeof(long long) * std::max<long long>(total, 1)));
    hipLaunchKernelGGL(KScatterFlags, dim3(chunks), dim3(256), 0, h->stream, h->flags, mask, h->nx, h->ny, h->nz,
                       static_cast<const long long*>(offs_buf.p), *out);
    FC_HIP(h, hipStreamSynchronize(h->stream));
    *count = static_cast<int>(total);
    return FC_LBM_OK;
}

int EnsureStaging(fc_lbm* h, size_t bytes) {
    if (bytes <= h->staging_bytes) return FC_LBM_OK;
    DevFree(h->staging);
    h->staging = nullptr;
    h->staging_bytes = 0;
    FC_HIP(h, hipMalloc(&h->staging, bytes));
    h->staging_bytes = bytes;
    return FC_LBM_OK;
}

// Host copy of the solid mask, made from the device copy when the geometry came from
// device buffers.
int EnsureSolidHost(fc_lbm* h) {
    if (static_cast<long long>(h->solid_host.size()) == h->N) return FC_LBM_OK;
    h->solid_host.resize(h->N);
    FC_HIP(h, hipMemcpyAsync(h->solid_host.data(), h->dsolid, h->N, hipMemcpyDeviceToHost, h->stream));
    FC_HIP(h, hipStreamSynchronize(h->stream));
    return FC_LBM_OK;
}

int CheckReady(fc_lbm* h) {
    if (!h->has_geometry) return Fail(h, FC_LBM_ESTATE, "no geometry: call fc_lbm_set_geometry first");
    if (!h->has_state) return Fail(h, FC_LBM_ESTATE, "no state: call a reset first");
    return FC_LBM_OK;
}

// In-place layout: side values must describe the state after a seed/restore.
void RefreshSide(fc_lbm* h) {
    if (h->layout != FC_LBM_LAYOUT_INPLACE) return;
    const long long work = static_cast<long long>(h->nrec) + 2LL * h->ny * h->nz;
    Dispatch(h, [&]<int Q, class S, class Lay>() {
        using R = typename S::R;
        hipLaunchKernelGGL((KRefreshSide<Q, S, Lay>), dim3(Blocks(work, 256)), dim3(256), 0, h->stream, h->pops[0],
                           h->pops[1], h->P, h->t, h->nx, h->ny, h->nz, h->pery, h->perz, h->nearlist, h->nrec,
                           static_cast<BRec<R>*>(h->recs), static_cast<R*>(h->side_mom));
    });
}

void LaunchSurfaceForce(fc_lbm* h) {
    if (h->nsurf == 0) return;
    Dispatch(h, [&]<int Q, class S, class Lay>() {
        using R = typename S::R;
        hipLaunchKernelGGL((KSurfaceForce<Q, S, Lay>), dim3(Blocks(h->nsurf, 256)), dim3(256), 0, h->stream,
                           h->pops[0], h->pops[1], h->P, h->t, h->nx, h->ny, h->nz, h->pery, h->perz, h->flags,
                           h->bidx, static_cast<const BRec<R>*>(h->recs), h->surf, h->nsurf,
                           static_cast<R*>(h->sforce));
    });
}

void LaunchReduceForces(fc_lbm* h, double* out) {
    if (h->rbytes == 8)
        hipLaunchKernelGGL(KReduceForces<double>, dim3(1), dim3(1024), 0, h->stream,
                           static_cast<const double*>(h->sforce), h->nsurf, out);
    else
        hipLaunchKernelGGL(KReduceForces<float>, dim3(1), dim3(1024), 0, h->stream,
                           static_cast<const float*>(h->sforce), h->nsurf, out);
}

// The field of layers [z0, z0+zc) into dout (4 floats per cell, every cell written).
// surface_ready: the surface forces of the current state are already in sforce.
int FieldChunk(fc_lbm* h, int z0, int zc, float* dout, bool surface_ready = false) {
    const long long cnt = static_cast<long long>(h->nx) * h->ny * zc;
    if (!surface_ready) LaunchSurfaceForce(h);
    Dispatch(h, [&]<int Q, class S, class Lay>() {
        hipLaunchKernelGGL((KField<Q, S, Lay>), dim3(Blocks(cnt, 256)), dim3(256), 0, h->stream, h->pops[0],
                           h->pops[1], h->P, h->t, h->nx, h->ny, h->nz, h->pery, h->perz, h->flags, z0, zc, dout);
    });
    if (h->nsurf > 0) {
        if (h->rbytes == 8)
            hipLaunchKernelGGL(KScatterSurface<double>, dim3(Blocks(h->nsurf, 256)), dim3(256), 0, h->stream, h->surf,
                               h->nsurf, static_cast<const double*>(h->sforce), h->nx, h->ny, z0, zc, dout);
        else
            hipLaunchKernelGGL(KScatterSurface<float>, dim3(Blocks(h->nsurf, 256)), dim3(256), 0, h->stream, h->surf,
                               h->nsurf, static_cast<const float*>(h->sforce), h->nx, h->ny, z0, zc, dout);
    }
    FC_HIP(h, hipGetLastError());
    return FC_LBM_OK;
}

int Seed(fc_lbm* h, const float* dfield, double rho0, double u0) {
    Dispatch(h, [&]<int Q, class S, class Lay>() {
        hipLaunchKernelGGL((KSeed<Q, S, Lay>), dim3(Blocks(h->N, 256)), dim3(256), 0, h->stream, h->pops[0],
                           h->pops[1], h->P, h->t, h->nx, h->ny, h->nz, h->pery, h->perz, h->dsolid, dfield, rho0, u0);
    });
    FC_HIP(h, hipMemsetAsync(h->dflag, 0, sizeof(int), h->stream));
    RefreshSide(h);
    FC_HIP(h, hipGetLastError());
    h->has_state = true;
    return FC_LBM_OK;
}

}  // namespace

extern "C" {

void fc_lbm_params_default(fc_lbm_params* p) {
    if (!p) return;
    p->tau = 0.8f;
    p->inlet_velocity = 0.05f;
    p->reverse_flow = 0;
    p->free_slip_y = 1;
    p->free_slip_z = 1;
    p->smagorinsky_cs = 0.12f;
}

int fc_lbm_create(const fc_lbm_desc* desc, fc_lbm** out) {
    if (!out) return FC_LBM_EINVAL;
    *out = nullptr;
    if (!desc) return FC_LBM_EINVAL;
    fc_lbm* h = new fc_lbm();
    *out = h;  // returned even on failure so fc_lbm_last_error can explain; destroy it either way
    if (desc->nx < 4 || desc->ny < 1 || desc->nz < 1)
        return Fail(h, FC_LBM_EINVAL, "lattice must be at least 4 x 1 x 1");
    if (desc->model != FC_LBM_D3Q15 && desc->model != FC_LBM_D3Q19) return Fail(h, FC_LBM_EINVAL, "unknown model");
    if (desc->storage < FC_LBM_STORE_F32 || desc->storage > FC_LBM_STORE_BF16_SR)
        return Fail(h, FC_LBM_EINVAL, "unknown storage");
    if (desc->layout != FC_LBM_LAYOUT_PINGPONG && desc->layout != FC_LBM_LAYOUT_INPLACE)
        return Fail(h, FC_LBM_EINVAL, "unknown layout");
    h->nx = desc->nx;
    h->ny = desc->ny;
    h->nz = desc->nz;
    h->Q = desc->model == FC_LBM_D3Q15 ? 15 : 19;
    h->storage = desc->storage;
    h->layout = desc->layout;
    h->pery = desc->periodic_y ? 1 : 0;
    h->perz = desc->periodic_z ? 1 : 0;
    h->N = static_cast<long long>(h->nx) * h->ny * h->nz;
    h->P = static_cast<long long>(h->nx) * (h->ny + 2) * (h->nz + 2);
    // Kernels over every cell launch one work-item per cell in a one-dimensional grid, whose
    // size is a 32-bit count of work-items; cell and padded-cell indices are 64-bit.
    if (h->N > kMaxCells)
        return Fail(h, FC_LBM_EINVAL, "lattice too large (at most " + std::to_string(kMaxCells) + " cells)");
    h->tbytes = h->storage == FC_LBM_STORE_F64 ? 8 : (h->storage == FC_LBM_STORE_F32 ? 4 : 2);
    // The step kernel addresses three padded z layers of all planes with 32-bit byte
    // offsets from one base.
    if ((3LL * (h->ny + 2) + 2) * PopRowStride(h->nx, h->Q) * h->tbytes >= (1LL << 31))
        return Fail(h, FC_LBM_EINVAL, "lattice cross-section too large (3*Q*nx*(ny+2) population bytes must be < 2 GiB)");
    h->rbytes = h->storage == FC_LBM_STORE_F64 ? 8 : 4;
    fc_lbm_params_default(&h->params);
    // One cell per thread: with every pull of a thread issued before the flag branch,
    // two cells per thread (twice the bytes in flight per wave, at a lower occupancy)
    // lose 3-8% in every format and layout (512x256x256: FP16 ping-pong 69.7 vs 66.7
    // GLUPS, FP32 36.7 vs 35.6).  FC_LBM_CELLS_PER_THREAD=2 selects the two-cell variant.
    h->cells_per_thread = 1;
    if (const char* e = std::getenv("FC_LBM_CELLS_PER_THREAD")) h->cells_per_thread = std::atoi(e) == 2 ? 2 : 1;
    FC_HIP(h, hipGetDevice(&h->device));
    {
        void* p = nullptr;
        FC_HIP(h, ScratchAddress(h->device, &p));
        h->scratch = static_cast<float*>(p);
    }
    if (desc->stream) {
        h->stream = static_cast<hipStream_t>(desc->stream);
    } else {
        FC_HIP(h, hipStreamCreate(&h->stream));
        h->own_stream = true;
    }
    const size_t popbytes = PopElements(h) * h->tbytes;
    const int copies = h->layout == FC_LBM_LAYOUT_INPLACE ? 1 : 2;
    for (int i = 0; i < copies; ++i) {
        if (hipMalloc(&h->pops[i], popbytes) != hipSuccess) {
            (void)hipGetLastError();
            return Fail(h, FC_LBM_ENOMEM, "population allocation failed (" + std::to_string(popbytes) + " bytes)");
        }
        FC_HIP(h, hipMemsetAsync(h->pops[i], 0, popbytes, h->stream));
    }
    if (copies == 1) h->pops[1] = h->pops[0];
    h->resident = static_cast<long long>(popbytes) * copies <= kResidentBytes;
    if (const char* e = std::getenv("FC_LBM_RESIDENT")) h->resident = std::atoi(e) != 0;
    FC_HIP(h, hipMalloc(&h->side_mom, 16ull * h->ny * h->nz * h->rbytes));
    FC_HIP(h, hipMemsetAsync(h->side_mom, 0, 16ull * h->ny * h->nz * h->rbytes, h->stream));
    FC_HIP(h, hipMalloc(&h->dforces, 4 * sizeof(double)));
    FC_HIP(h, hipMalloc(&h->stat_part, kStatBlocks * 8 * sizeof(double)));
    FC_HIP(h, hipMalloc(&h->stat_out, 8 * sizeof(double)));
    FC_HIP(h, hipMalloc(&h->aero_part, kAeroPartDoubles * sizeof(double)));
    FC_HIP(h, hipMalloc(&h->aero_out, 24 * sizeof(double)));
    FC_HIP(h, hipMalloc(&h->dflag, sizeof(int)));
    FC_HIP(h, hipMalloc(&h->dprobe, sizeof(int)));
    FC_HIP(h, hipMemsetAsync(h->dflag, 0, sizeof(int), h->stream));
    FC_HIP(h, hipStreamSynchronize(h->stream));
    return FC_LBM_OK;
}

int fc_lbm_destroy(fc_lbm* h) {
    if (!h) return FC_LBM_EINVAL;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    FreeGeometry(h);
    if (h->pops[1] != h->pops[0]) DevFree(h->pops[1]);
    DevFree(h->pops[0]);
    DevFree(h->side_mom);
    DevFree(h->dforces);
    DevFree(h->stat_part);
    DevFree(h->stat_out);
    DevFree(h->aero_part);
    DevFree(h->aero_out);
    DevFree(h->dflag);
    DevFree(h->dprobe);
    DevFree(h->staging);
    if (h->copy_stream) (void)hipStreamSynchronize(h->copy_stream);
    DevFree(h->snap_dev);
    if (h->snap_host) (void)hipHostFree(h->snap_host);
    if (h->snap_taken) (void)hipEventDestroy(h->snap_taken);
    if (h->snap_copied) (void)hipEventDestroy(h->snap_copied);
    if (h->copy_stream) (void)hipStreamDestroy(h->copy_stream);
    if (h->own_stream && h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return FC_LBM_OK;
}

const char* fc_lbm_last_error(const fc_lbm* h) { return h ? h->err.c_str() : kNoHandle; }

// Mask and SDF from host arrays or (device) from device buffers.
static int SetGeometry(fc_lbm* h, const uint8_t* solid, const float* sdf, bool device) {
    if (!h) return FC_LBM_EINVAL;
    if (!h->pops[0]) return Fail(h, FC_LBM_ESTATE, "handle was not created successfully");
    if (!solid) return Fail(h, FC_LBM_EINVAL, "solid mask is required");
    FC_HIP(h, hipStreamSynchronize(h->stream));
    FreeGeometry(h);
    const hipMemcpyKind in_kind = device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    FC_HIP(h, hipMalloc(&h->dsolid, h->N));
    FC_HIP(h, hipMemcpyAsync(h->dsolid, solid, h->N, in_kind, h->stream));
    // A host copy of the mask serves the aerodynamic summary's lift-arrow column search; from
    // host input it is kept now, from device input it is made on first use (EnsureSolidHost).
    h->solid_host.clear();
    if (!device) h->solid_host.assign(solid, solid + h->N);
    // the device copy of the SDF lives only while the link table is built
    DevTmp sdfbuf{h->stream};
    float* dsdf = nullptr;
    if (sdf) {
        FC_HIP(h, hipMalloc(&sdfbuf.p, h->N * sizeof(float)));
        dsdf = static_cast<float*>(sdfbuf.p);
        FC_HIP(h, hipMemcpyAsync(dsdf, sdf, h->N * sizeof(float), in_kind, h->stream));
    }
    FC_HIP(h, hipMalloc(&h->flags, h->P));
    FC_HIP(h, hipMemsetAsync(h->flags, 0, h->P, h->stream));
    if (h->Q == 15)
        hipLaunchKernelGGL(KClassify<15>, dim3(Blocks(h->N, 256)), dim3(256), 0, h->stream, h->dsolid, h->flags, h->nx,
                           h->ny, h->nz, h->pery, h->perz);
    else
        hipLaunchKernelGGL(KClassify<19>, dim3(Blocks(h->N, 256)), dim3(256), 0, h->stream, h->dsolid, h->flags, h->nx,
                           h->ny, h->nz, h->pery, h->perz);
    FC_HIP(h, hipGetLastError());
    int rc = Compact(h, kNearSolid, &h->nearlist, &h->nrec);
    if (rc) return rc;
    rc = Compact(h, kSurface, &h->surf, &h->nsurf);
    if (rc) return rc;
    FC_HIP(h, hipMalloc(&h->bidx, h->P * sizeof(int)));
    FC_HIP(h, hipMemsetAsync(h->bidx, 0xFF, h->P * sizeof(int), h->stream));
    const size_t recbytes = h->rbytes == 8 ? sizeof(BRec<double>) : sizeof(BRec<float>);
    FC_HIP(h, hipMalloc(&h->recs, recbytes * std::max(h->nrec, 1)));
    if (h->nrec > 0) {
        Dispatch(h, [&]<int Q, class S, class Lay>() {
            using R = typename S::R;
            hipLaunchKernelGGL((KBuildRecords<Q, R>), dim3(Blocks(h->nrec, 256)), dim3(256), 0, h->stream, h->dsolid,
                               dsdf, h->nearlist, h->nrec, h->bidx, static_cast<BRec<R>*>(h->recs), h->nx, h->ny,
                               h->nz, h->pery, h->perz);
        });
    }
    FC_HIP(h, hipMalloc(&h->sforce, 4 * h->rbytes * std::max(h->nsurf, 1)));
    FC_HIP(h, hipGetLastError());
    FC_HIP(h, hipStreamSynchronize(h->stream));
    h->has_geometry = true;
    h->aero_geom_valid = false;
    h->has_state = false;
    return FC_LBM_OK;
}

int fc_lbm_set_geometry(fc_lbm* h, const uint8_t* solid, const float* sdf) {
    return SetGeometry(h, solid, sdf, false);
}

int fc_lbm_set_geometry_device(fc_lbm* h, const uint8_t* solid_device, const float* sdf_device) {
    return SetGeometry(h, solid_device, sdf_device, true);
}

int fc_lbm_solid_mask(fc_lbm* h, uint8_t* out) {
    if (!h || !out) return FC_LBM_EINVAL;
    if (!h->has_geometry) return Fail(h, FC_LBM_ESTATE, "no geometry: call fc_lbm_set_geometry first");
    FC_HIP(h, hipMemcpyAsync(out, h->dsolid, h->N, hipMemcpyDeviceToHost, h->stream));
    FC_HIP(h, hipStreamSynchronize(h->stream));
    return FC_LBM_OK;
}

int fc_lbm_set_params(fc_lbm* h, const fc_lbm_params* p) {
    if (!h || !p) return FC_LBM_EINVAL;
    h->params = *p;
    return FC_LBM_OK;
}

int fc_lbm_get_params(const fc_lbm* h, fc_lbm_params* p) {
    if (!h || !p) return FC_LBM_EINVAL;
    *p = h->params;
    return FC_LBM_OK;
}

int fc_lbm_reset_uniform(fc_lbm* h, int seed_mid) {
    if (!h) return FC_LBM_EINVAL;
    if (!h->has_geometry) return Fail(h, FC_LBM_ESTATE, "no geometry: call fc_lbm_set_geometry first");
    h->t = 0;
    double rho0 = 1.0, u0 = 0.0;
    if (seed_mid) {
        if (h->rbytes == 4) {
            // LbmReference::Reset (LbmReference.cpp:192-202), in float.
            float u = h->params.inlet_velocity;
            float mid = 0.5f * u;
            float ux = h->params.reverse_flow ? -mid : mid;
            float dev = 0.5f * (u * u - mid * mid);
            rho0 = 1.0f + 3.0f * dev;
            u0 = ux;
        } else {
            double u = h->params.inlet_velocity;
            double mid = 0.5 * u;
            u0 = h->params.reverse_flow ? -mid : mid;
            rho0 = 1.0 + 3.0 * (0.5 * (u * u - mid * mid));
        }
    }
    return Seed(h, nullptr, rho0, u0);
}

int fc_lbm_reset_from_field_device(fc_lbm* h, const float* field4_device) {
    if (!h || !field4_device) return FC_LBM_EINVAL;
    if (!h->has_geometry) return Fail(h, FC_LBM_ESTATE, "no geometry: call fc_lbm_set_geometry first");
    h->t = 0;
    return Seed(h, field4_device, 1.0, 0.0);
}

int fc_lbm_reset_from_field(fc_lbm* h, const float* field4) {
    if (!h || !field4) return FC_LBM_EINVAL;
    if (!h->has_geometry) return Fail(h, FC_LBM_ESTATE, "no geometry: call fc_lbm_set_geometry first");
    int rc = EnsureStaging(h, static_cast<size_t>(h->N) * 16);
    if (rc) return rc;
    FC_HIP(h, hipMemcpyAsync(h->staging, field4, static_cast<size_t>(h->N) * 16, hipMemcpyHostToDevice, h->stream));
    rc = fc_lbm_reset_from_field_device(h, h->staging);
    if (rc) return rc;
    FC_HIP(h, hipStreamSynchronize(h->stream));
    return FC_LBM_OK;
}

// Byte offsets of the step kernel's bulk pull and stores from the workgroup base (the
// padded z layer below the cells' own, lowered by StepBias elements; blocked populations,
// see lbm_kernels.hpp): for direction d of a cell at padded (x, y, 1) relative to that
// layer, with the x shift applied inside the cell's block (the kernel moves the two edge
// lanes of a block across the boundary).  The bias makes every offset non-negative.
static void StepOffsets(const fc_lbm* h, int par, StepArgs& a) {
    const long long rs = PopRowStride(h->nx, h->Q);
    const long long tb = h->tbytes;
    const long long bias = StepBias(rs);
    a.layer_bytes = (h->ny + 2) * rs * tb;
    a.base_bytes = (PopFront(rs, h->Q) - bias) * tb;
    auto elem = [&](int layer, int plane, int dy, int dx) {
        return bias + layer * (h->ny + 2) * rs + dy * rs + plane * kBlk + dx;
    };
    auto lat = [&](int d, int& ex, int& ey, int& ez) {
        if (h->Q == 19) {
            ex = Lat<19>::ex[d]; ey = Lat<19>::ey[d]; ez = Lat<19>::ez[d];
        } else {
            ex = Lat<15>::ex[d]; ey = Lat<15>::ey[d]; ez = Lat<15>::ez[d];
        }
    };
    for (int d = 0; d < 19; ++d) {
        a.ld_off[d] = a.st_off[d] = 0;
        if (d >= h->Q) continue;
        long long ld, st, prev;
        if (h->layout != FC_LBM_LAYOUT_INPLACE) {
            int ex, ey, ez;
            lat(d, ex, ey, ez);
            ld = elem(1 - ez, d, -ey, -ex);  // pull from x - e_d
            st = prev = elem(1, d, 0, 0);
        } else {
            const int pa = d == 0 ? 0 : ((d & 1) ? d : d - 1);
            int ex, ey, ez;
            lat(pa, ex, ey, ez);
            int pl;
            bool anc;
            LoadSlot(d, par, pl, anc);
            ld = anc ? elem(1 + ez, pl, ey, ex) : elem(1, pl, 0, 0);
            StoreSlot(d, par, pl, anc);
            st = anc ? elem(1 + ez, pl, ey, ex) : elem(1, pl, 0, 0);
            StoreSlot(d, par ^ 1, pl, anc);
            prev = anc ? elem(1 + ez, pl, ey, ex) : elem(1, pl, 0, 0);
        }
        a.ld_off[d] = static_cast<uint32_t>(ld * tb);
        a.st_off[d] = static_cast<uint32_t>(st * tb);
        a.prev_off[d] = static_cast<uint32_t>(prev * tb);
    }
}

// Inlet populations for the current parameters, with the step's own arithmetic.
extern "C++" template <int Q, class S>
static void InletPops(const fc_lbm* h, StepArgs& a) {
    using R = typename S::R;
    R fo[Q];
    const R u = static_cast<R>(h->params.reverse_flow ? -h->params.inlet_velocity : h->params.inlet_velocity);
    if constexpr (S::kOffset)
        EqOffset<Q, S>(0.0f, u, 0.0f, 0.0f, fo);
    else
        FeqAll<Q, R>(R(1), u, R(0), R(0), fo);
    for (int d = 0; d < 19; ++d) {
        a.inlet_f[d] = d < Q ? static_cast<float>(fo[d]) : 0.0f;
        a.inlet_d[d] = d < Q ? static_cast<double>(fo[d]) : 0.0;
    }
}

int fc_lbm_step_range(fc_lbm* h, int z0, int zcount) {
    if (!h) return FC_LBM_EINVAL;
    int rc = CheckReady(h);
    if (rc) return rc;
    if (z0 < 0) {
        zcount += z0;
        z0 = 0;
    }
    if (z0 + zcount > h->nz) zcount = h->nz - z0;
    if (zcount <= 0) return FC_LBM_OK;
    StepArgs a{};
    const int par = static_cast<int>(h->t & 1);
    a.pop_src = h->pops[par];
    a.pop_dst = h->layout == FC_LBM_LAYOUT_INPLACE ? h->pops[0] : h->pops[par ^ 1];
    a.flags = h->flags;
    a.bidx = h->bidx;
    a.recs = h->recs;
    a.side_mom = h->side_mom;
    a.diverged = h->dflag;
    a.P = h->P;
    StepOffsets(h, par, a);
    a.nx = h->nx;
    a.ny = h->ny;
    a.nz = h->nz;
    a.z0 = z0;
    a.zc = zcount;
    a.parity = par;
    a.step = static_cast<uint32_t>(h->t);
    a.reverse = h->params.reverse_flow ? 1 : 0;
    a.fsy = h->params.free_slip_y ? 1 : 0;
    a.fsz = h->params.free_slip_z ? 1 : 0;
    a.pery = h->pery;
    a.perz = h->perz;
    a.tau = h->params.tau;
    a.u_in = h->params.inlet_velocity;
    a.cs = h->params.smagorinsky_cs;
    a.tau_d = h->params.tau;
    a.u_in_d = h->params.inlet_velocity;
    a.cs_d = h->params.smagorinsky_cs;
    const int cpt = h->cells_per_thread;
    const int tx = TileX(static_cast<int>(h->tbytes)), ty = TileY(static_cast<int>(h->tbytes));
    const long long layer_tiles = static_cast<long long>((h->nx + tx - 1) / tx) * ((h->ny + ty * cpt - 1) / (ty * cpt));
    // Lattices past 2^31 work-items per launch step in z ranges (one step's ranges are
    // independent: each cell's update reads only the previous state).
    const int zmax = static_cast<int>(std::max<long long>(1, kMaxStepItems / (StepThreads(static_cast<int>(h->tbytes)) * layer_tiles)));
    hipError_t e = hipSuccess;
    DispatchStep(h, [&]<int Q, class S, class Lay>() {
        InletPops<Q, S>(h, a);
        for (int z = z0; z < z0 + zcount && e == hipSuccess; z += zmax) {
            a.z0 = z;
            a.zc = std::min(zmax, z0 + zcount - z);
            e = LaunchStep<Q, S, Lay>(a, static_cast<unsigned>(layer_tiles * a.zc), cpt, (h->pery | h->perz) != 0,
                                      h->resident, h->stream);
        }
    });
    FC_HIP(h, e);
    return FC_LBM_OK;
}

int fc_lbm_advance(fc_lbm* h) {
    if (!h) return FC_LBM_EINVAL;
    ++h->t;
    return FC_LBM_OK;
}

int fc_lbm_step(fc_lbm* h, int n) {
    if (!h) return FC_LBM_EINVAL;
    for (int i = 0; i < n; ++i) {
        int rc = fc_lbm_step_range(h, 0, h->nz);
        if (rc) return rc;
        ++h->t;
    }
    return FC_LBM_OK;
}

int fc_lbm_field_device(fc_lbm* h, float* out_device) {
    if (!h || !out_device) return FC_LBM_EINVAL;
    int rc = CheckReady(h);
    if (rc) return rc;
    return FieldChunk(h, 0, h->nz, out_device);
}

int fc_lbm_field_at(fc_lbm* h, const int* cells_device, long long n, float* out_device) {
    if (!h || n < 0 || (n > 0 && (!cells_device || !out_device))) return FC_LBM_EINVAL;
    int rc = CheckReady(h);
    if (rc) return rc;
    if (n == 0) return FC_LBM_OK;
    Dispatch(h, [&]<int Q, class S, class Lay>() {
        hipLaunchKernelGGL((KFieldAt<Q, S, Lay>), dim3(Blocks(n, 256)), dim3(256), 0, h->stream, h->pops[0],
                           h->pops[1], h->t, h->nx, h->ny, h->nz, h->pery, h->perz, h->flags, cells_device, n,
                           out_device);
    });
    FC_HIP(h, hipGetLastError());
    return FC_LBM_OK;
}

int fc_lbm_snapshot_begin(fc_lbm* h) {
    if (!h) return FC_LBM_EINVAL;
    int rc = CheckReady(h);
    if (rc) return rc;
    const size_t bytes = static_cast<size_t>(h->N) * 4 * sizeof(float);
    // first use: each resource on its own, so a call after a failed set-up completes it
    if (!h->snap_dev) FC_HIP(h, hipMalloc(&h->snap_dev, bytes));
    if (!h->snap_host) FC_HIP(h, hipHostMalloc(reinterpret_cast<void**>(&h->snap_host), bytes, hipHostMallocDefault));
    if (!h->copy_stream) FC_HIP(h, hipStreamCreateWithFlags(&h->copy_stream, hipStreamNonBlocking));
    if (!h->snap_taken) FC_HIP(h, hipEventCreateWithFlags(&h->snap_taken, hipEventDisableTiming));
    if (!h->snap_copied) {
        FC_HIP(h, hipEventCreateWithFlags(&h->snap_copied, hipEventDisableTiming));
        FC_HIP(h, hipEventRecord(h->snap_copied, h->copy_stream));
    }
    // the previous snapshot's copy reads snap_dev (and fills snap_host, which the caller
    // may still be reading): the new field waits for it on the device, the host waits here
    FC_HIP(h, hipEventSynchronize(h->snap_copied));
    FC_HIP(h, hipStreamWaitEvent(h->stream, h->snap_copied, 0));
    rc = FieldChunk(h, 0, h->nz, h->snap_dev);
    if (rc) return rc;
    FC_HIP(h, hipEventRecord(h->snap_taken, h->stream));
    FC_HIP(h, hipStreamWaitEvent(h->copy_stream, h->snap_taken, 0));
    FC_HIP(h, hipMemcpyAsync(h->snap_host, h->snap_dev, bytes, hipMemcpyDeviceToHost, h->copy_stream));
    FC_HIP(h, hipEventRecord(h->snap_copied, h->copy_stream));
    h->snap_pending = true;
    return FC_LBM_OK;
}

int fc_lbm_snapshot_poll(fc_lbm* h, int* ready) {
    if (!h || !ready) return FC_LBM_EINVAL;
    if (!h->snap_pending) return Fail(h, FC_LBM_ESTATE, "no snapshot was started");
    const hipError_t e = hipEventQuery(h->snap_copied);
    if (e == hipErrorNotReady) {
        *ready = 0;
        return FC_LBM_OK;
    }
    FC_HIP(h, e);
    *ready = 1;
    return FC_LBM_OK;
}

int fc_lbm_snapshot_wait(fc_lbm* h, const float** field) {
    if (!h || !field) return FC_LBM_EINVAL;
    *field = nullptr;
    if (!h->snap_pending) return Fail(h, FC_LBM_ESTATE, "no snapshot was started");
    FC_HIP(h, hipEventSynchronize(h->snap_copied));
    *field = h->snap_host;
    return FC_LBM_OK;
}

int fc_lbm_readback_field_range(fc_lbm* h, int z0, int zcount, float* out) {
    if (!h || !out) return FC_LBM_EINVAL;
    int rc = CheckReady(h);
    if (rc) return rc;
    if (z0 < 0) {
        zcount += z0;
        z0 = 0;
    }
    if (z0 + zcount > h->nz) zcount = h->nz - z0;
    if (zcount <= 0) return FC_LBM_OK;
    const size_t layer = static_cast<size_t>(h->nx) * h->ny * 4;
    rc = EnsureStaging(h, layer * zcount * sizeof(float));
    if (rc) return rc;
    rc = FieldChunk(h, z0, zcount, h->staging);
    if (rc) return rc;
    FC_HIP(h, hipMemcpyAsync(out + layer * z0, h->staging, layer * zcount * sizeof(float), hipMemcpyDeviceToHost,
                             h->stream));
    FC_HIP(h, hipStreamSynchronize(h->stream));
    return FC_LBM_OK;
}

int fc_lbm_readback_field(fc_lbm* h, float* out) {
    if (!h) return FC_LBM_EINVAL;
    return fc_lbm_readback_field_range(h, 0, h->nz, out);
}

int fc_lbm_readback_field_range_noalloc(fc_lbm* h, int z0, int zcount, float* out) {
    if (!h || !out) return FC_LBM_EINVAL;
    int rc = CheckReady(h);
    if (rc) return rc;
    if (!h->scratch) return Fail(h, FC_LBM_ESTATE, "handle was not created successfully");
    const size_t layer = static_cast<size_t>(h->nx) * h->ny * 4;
    if (layer * sizeof(float) > kScratchBytes)
        return Fail(h, FC_LBM_EINVAL,
                    "readback_field_range_noalloc: one z layer (nx*ny*16 = " + std::to_string(layer * sizeof(float)) +
                        " bytes) exceeds the " + std::to_string(kScratchBytes >> 20) + " MiB scratch");
    if (z0 < 0) {
        zcount += z0;
        z0 = 0;
    }
    if (z0 + zcount > h->nz) zcount = h->nz - z0;
    if (zcount <= 0) return FC_LBM_OK;
    const int per = static_cast<int>(kScratchBytes / (layer * sizeof(float)));
    LaunchSurfaceForce(h);
    for (int z = z0; z < z0 + zcount; z += per) {
        const int zc = std::min(per, z0 + zcount - z);
        rc = FieldChunk(h, z, zc, h->scratch, true);
        if (rc) return rc;
        FC_HIP(h, hipMemcpyAsync(out + layer * z, h->scratch, layer * zc * sizeof(float), hipMemcpyDeviceToHost,
                                 h->stream));
    }
    FC_HIP(h, hipStreamSynchronize(h->stream));
    return FC_LBM_OK;
}

int fc_lbm_forces_device(fc_lbm* h, double* out_device) {
    if (!h || !out_device) return FC_LBM_EINVAL;
    int rc = CheckReady(h);
    if (rc) return rc;
    LaunchSurfaceForce(h);
    LaunchReduceForces(h, out_device);
    FC_HIP(h, hipGetLastError());
    return FC_LBM_OK;
}

int fc_lbm_forces(fc_lbm* h, double out[4]) {
    if (!h || !out) return FC_LBM_EINVAL;
    int rc = fc_lbm_forces_device(h, h->dforces);
    if (rc) return rc;
    FC_HIP(h, hipMemcpyAsync(out, h->dforces, 4 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    FC_HIP(h, hipStreamSynchronize(h->stream));
    return FC_LBM_OK;
}

int fc_lbm_aero_forces(fc_lbm* h, float cell_size, const float origin[3], fc_lbm_aero* out) {
    if (!h || !out || !origin) return FC_LBM_EINVAL;
    *out = fc_lbm_aero{};
    int rc = CheckReady(h);
    if (rc) return rc;
    rc = EnsureSolidHost(h);
    if (rc) return rc;
    const float ox = origin[0], oy = origin[1], oz = origin[2];
    LaunchSurfaceForce(h);
    if (h->rbytes == 8)
        hipLaunchKernelGGL(KAeroSurface<double>, dim3(kAeroBlocks), dim3(256), 0, h->stream, h->surf, h->nsurf,
                           static_cast<const double*>(h->sforce), h->nx, h->ny, cell_size, ox, oy, oz, h->aero_part);
    else
        hipLaunchKernelGGL(KAeroSurface<float>, dim3(kAeroBlocks), dim3(256), 0, h->stream, h->surf, h->nsurf,
                           static_cast<const float*>(h->sforce), h->nx, h->ny, cell_size, ox, oy, oz, h->aero_part);
    hipLaunchKernelGGL(KAeroFold<kAeroSVals>, dim3(1), dim3(256), 0, h->stream, h->aero_part, kAeroBlocks, true,
                       false, h->aero_out);
    // The geometric sums depend only on the geometry and the frame: computed on the first
    // call for a frame and kept on the host.
    const float frame[4] = {cell_size, ox, oy, oz};
    const bool geom = !h->aero_geom_valid || std::memcmp(frame, h->aero_frame, sizeof(frame)) != 0;
    if (geom) {
        double* gpart = h->aero_part + kAeroBlocks * kAeroSVals;
        hipLaunchKernelGGL(KAeroGeom, dim3(kGeomBlocks), dim3(256), 0, h->stream, h->flags, h->nx, h->ny, h->nz,
                           cell_size, ox, oy, oz, gpart);
        hipLaunchKernelGGL(KAeroFold<kAeroGVals>, dim3(1), dim3(256), 0, h->stream, gpart, kGeomBlocks, false, true,
                           h->aero_out + 16);
    }
    FC_HIP(h, hipGetLastError());
    double a[18];
    FC_HIP(h, hipMemcpyAsync(a, h->aero_out, kAeroSVals * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    if (geom)
        FC_HIP(h, hipMemcpyAsync(h->aero_geom, h->aero_out + 16, kAeroGVals * sizeof(double), hipMemcpyDeviceToHost,
                                 h->stream));
    FC_HIP(h, hipStreamSynchronize(h->stream));
    if (geom) {
        std::memcpy(h->aero_frame, frame, sizeof(frame));
        h->aero_geom_valid = true;
    }
    for (int i = 0; i < kAeroGVals; ++i) a[13 + i] = h->aero_geom[i];
    // Host epilogue, as FieldOverlay.cpp:553-618: stress range anchored at 0, centre of
    // pressure by |Fz|, else by |F|, else the solid centroid, and the lift-arrow base on
    // the top solid cell of the 5x5 columns around it.
    out->force_x = static_cast<float>(a[0]);
    out->force_y = static_cast<float>(a[1]);
    out->force_z = static_cast<float>(a[2]);
    out->has_nan = a[12] > 0.0 ? 1 : 0;
    out->min_stress = 0.0f;
    out->max_stress = a[11] > 0.0 ? static_cast<float>(a[11]) : 1e-6f;
    float cp[3] = {0.0f, 0.0f, 0.0f};
    const double solid_cells = a[13];
    if (a[3] > 1e-9) {
        for (int i = 0; i < 3; ++i) cp[i] = static_cast<float>(a[4 + i] / a[3]);
    } else if (a[7] > 1e-9) {
        for (int i = 0; i < 3; ++i) cp[i] = static_cast<float>(a[8 + i] / a[7]);
    } else if (solid_cells > 0.0) {
        const double inv = 1.0 / solid_cells;
        for (int i = 0; i < 3; ++i) cp[i] = static_cast<float>(a[14 + i] * inv);
    }
    const int gx0 = std::clamp(static_cast<int>(std::floor((cp[1] - oy) / cell_size)), 0, std::max(h->nx - 1, 0));
    const int gy0 = std::clamp(static_cast<int>(std::floor((cp[0] - ox) / cell_size)), 0, std::max(h->ny - 1, 0));
    int col_top = -1;
    for (int dy = -2; dy <= 2; ++dy) {
        const int gy = std::clamp(gy0 + dy, 0, std::max(h->ny - 1, 0));
        for (int dx = -2; dx <= 2; ++dx) {
            const int gx = std::clamp(gx0 + dx, 0, std::max(h->nx - 1, 0));
            for (int z = h->nz - 1; z >= 0; --z) {
                if (h->solid_host[(static_cast<size_t>(z) * h->ny + gy) * h->nx + gx]) {
                    col_top = std::max(col_top, z);
                    break;
                }
            }
        }
    }
    const int top = col_top >= 0 ? col_top : std::max(static_cast<int>(a[17]), 0);
    for (int i = 0; i < 3; ++i) out->center_of_pressure[i] = cp[i];
    out->lift_arrow_base[0] = cp[0];
    out->lift_arrow_base[1] = cp[1];
    out->lift_arrow_base[2] = oz + (static_cast<float>(top) + 1.0f) * cell_size;
    return FC_LBM_OK;
}

int fc_lbm_readback_stats(fc_lbm* h, fc_lbm_stats* out) {
    if (!h || !out) return FC_LBM_EINVAL;
    int rc = CheckReady(h);
    if (rc) return rc;
    LaunchSurfaceForce(h);
    LaunchReduceForces(h, h->dforces);
    Dispatch(h, [&]<int Q, class S, class Lay>() {
        hipLaunchKernelGGL((KStatsPartial<Q, S, Lay>), dim3(kStatBlocks), dim3(256), 0, h->stream, h->pops[0],
                           h->pops[1], h->P, h->t, h->nx, h->ny, h->nz, h->pery, h->perz, h->flags, h->stat_part);
    });
    hipLaunchKernelGGL(KStatsFinal, dim3(1), dim3(64), 0, h->stream, h->stat_part, h->stat_out);
    FC_HIP(h, hipGetLastError());
    double f[4], s[8];
    FC_HIP(h, hipMemcpyAsync(f, h->dforces, sizeof(f), hipMemcpyDeviceToHost, h->stream));
    FC_HIP(h, hipMemcpyAsync(s, h->stat_out, sizeof(s), hipMemcpyDeviceToHost, h->stream));
    FC_HIP(h, hipStreamSynchronize(h->stream));
    const long long n = static_cast<long long>(s[2]);
    out->fluid_cells = n;
    out->avg_density = n ? s[0] / static_cast<double>(n) : 0.0;
    out->avg_speed = n ? s[1] / static_cast<double>(n) : 0.0;
    out->min_density = n ? s[4] : 0.0;
    out->max_density = n ? s[5] : 0.0;
    out->max_speed = s[6];
    out->total_mass = s[0];
    out->has_nan = s[3] > 0 ? 1 : 0;
    out->force_x = f[0];
    out->force_y = f[1];
    out->force_z = f[2];
    out->max_stress = f[3];
    return FC_LBM_OK;
}

int fc_lbm_probe_health(fc_lbm* h, int z, int* diverged) {
    if (!h || !diverged) return FC_LBM_EINVAL;
    *diverged = 0;
    int rc = CheckReady(h);
    if (rc) return rc;
    z = std::clamp(z, 0, h->nz - 1);
    FC_HIP(h, hipMemsetAsync(h->dprobe, 0, sizeof(int), h->stream));
    LaunchSurfaceForce(h);
    Dispatch(h, [&]<int Q, class S, class Lay>() {
        hipLaunchKernelGGL((KProbe<Q, S, Lay>), dim3(Blocks(static_cast<long long>(h->nx) * h->ny, 256)), dim3(256), 0,
                           h->stream, h->pops[0], h->pops[1], h->P, h->t, h->nx, h->ny, h->nz, h->pery, h->perz,
                           h->flags, z, h->dprobe);
    });
    if (h->nsurf > 0) {
        if (h->rbytes == 8)
            hipLaunchKernelGGL(KProbeSurface<double>, dim3(Blocks(h->nsurf, 256)), dim3(256), 0, h->stream, h->surf,
                               h->nsurf, static_cast<const double*>(h->sforce), h->nx, h->ny, z, h->dprobe);
        else
            hipLaunchKernelGGL(KProbeSurface<float>, dim3(Blocks(h->nsurf, 256)), dim3(256), 0, h->stream, h->surf,
                               h->nsurf, static_cast<const float*>(h->sforce), h->nx, h->ny, z, h->dprobe);
    }
    FC_HIP(h, hipGetLastError());
    FC_HIP(h, hipMemcpyAsync(diverged, h->dprobe, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    FC_HIP(h, hipStreamSynchronize(h->stream));
    return FC_LBM_OK;
}

int fc_lbm_has_diverged(fc_lbm* h, int* diverged) {
    if (!h || !diverged) return FC_LBM_EINVAL;
    *diverged = 0;
    if (!h->dflag) return Fail(h, FC_LBM_ESTATE, "handle was not created successfully");
    FC_HIP(h, hipMemcpyAsync(diverged, h->dflag, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    FC_HIP(h, hipStreamSynchronize(h->stream));
    return FC_LBM_OK;
}

int fc_lbm_get_populations(fc_lbm* h, void* out) {
    if (!h || !out) return FC_LBM_EINVAL;
    int rc = CheckReady(h);
    if (rc) return rc;
    const size_t bytes = static_cast<size_t>(h->Q) * h->N * h->rbytes;
    void* tmp = nullptr;
    FC_HIP(h, hipMalloc(&tmp, bytes));
    Dispatch(h, [&]<int Q, class S, class Lay>() {
        using R = typename S::R;
        hipLaunchKernelGGL((KGetPops<Q, S, Lay>), dim3(Blocks(h->N, 256)), dim3(256), 0, h->stream, h->pops[0],
                           h->pops[1], h->P, h->t, h->nx, h->ny, h->nz, h->pery, h->perz, h->dsolid,
                           static_cast<R*>(tmp));
    });
    hipError_t e = hipMemcpyAsync(out, tmp, bytes, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    DevFree(tmp);
    if (e != hipSuccess) return Fail(h, FC_LBM_EDEVICE, std::string("get_populations: ") + hipGetErrorString(e));
    return FC_LBM_OK;
}

int fc_lbm_set_populations(fc_lbm* h, const void* in) {
    if (!h || !in) return FC_LBM_EINVAL;
    if (!h->has_geometry) return Fail(h, FC_LBM_ESTATE, "no geometry: call fc_lbm_set_geometry first");
    const size_t bytes = static_cast<size_t>(h->Q) * h->N * h->rbytes;
    void* tmp = nullptr;
    FC_HIP(h, hipMalloc(&tmp, bytes));
    hipError_t e = hipMemcpyAsync(tmp, in, bytes, hipMemcpyHostToDevice, h->stream);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(h->stream);
        DevFree(tmp);
        return Fail(h, FC_LBM_EDEVICE, std::string("set_populations: ") + hipGetErrorString(e));
    }
    Dispatch(h, [&]<int Q, class S, class Lay>() {
        using R = typename S::R;
        hipLaunchKernelGGL((KSetPops<Q, S, Lay>), dim3(Blocks(h->N, 256)), dim3(256), 0, h->stream, h->pops[0],
                           h->pops[1], h->P, h->t, h->nx, h->ny, h->nz, h->pery, h->perz, static_cast<const R*>(tmp));
    });
    h->has_state = true;
    RefreshSide(h);
    e = hipStreamSynchronize(h->stream);
    DevFree(tmp);
    if (e != hipSuccess) return Fail(h, FC_LBM_EDEVICE, std::string("set_populations: ") + hipGetErrorString(e));
    return FC_LBM_OK;
}

int fc_lbm_sync(fc_lbm* h) {
    if (!h) return FC_LBM_EINVAL;
    FC_HIP(h, hipStreamSynchronize(h->stream));
    return FC_LBM_OK;
}

int fc_lbm_info(const fc_lbm* h, int* nx, int* ny, int* nz, int* q, int* storage, int* layout) {
    if (!h) return FC_LBM_EINVAL;
    if (nx) *nx = h->nx;
    if (ny) *ny = h->ny;
    if (nz) *nz = h->nz;
    if (q) *q = h->Q;
    if (storage) *storage = h->storage;
    if (layout) *layout = h->layout;
    return FC_LBM_OK;
}

long long fc_lbm_step_count(const fc_lbm* h) { return h ? h->t : -1; }

size_t fc_lbm_population_bytes(const fc_lbm* h) {
    if (!h) return 0;
    return PopElements(h) * h->tbytes * (h->layout == FC_LBM_LAYOUT_INPLACE ? 1 : 2);
}

void* fc_lbm_stream(const fc_lbm* h) { return h ? static_cast<void*>(h->stream) : nullptr; }

long long fc_lbm_surface_cells(const fc_lbm* h) { return h ? h->nsurf : -1; }

int fc_lbm_boundary_links(fc_lbm* h, long long max, long long* cell, int* dir, int* kind, float* q,
                          long long* count) {
    if (!h || !count) return FC_LBM_EINVAL;
    if (!h->has_geometry) return Fail(h, FC_LBM_ESTATE, "no geometry: call fc_lbm_set_geometry first");
    std::vector<long long> list(std::max(h->nrec, 1));
    const size_t rb = h->rbytes == 8 ? sizeof(BRec<double>) : sizeof(BRec<float>);
    std::vector<unsigned char> recs(rb * std::max(h->nrec, 1));
    FC_HIP(h, hipMemcpy(list.data(), h->nearlist, sizeof(long long) * h->nrec, hipMemcpyDeviceToHost));
    FC_HIP(h, hipMemcpy(recs.data(), h->recs, rb * h->nrec, hipMemcpyDeviceToHost));
    long long k = 0;
    for (int r = 0; r < h->nrec; ++r) {
        const long long n = list[r];
        const long long layer = static_cast<long long>(h->nx) * (h->ny + 2);
        const int x = static_cast<int>(n % h->nx);
        const int y = static_cast<int>((n / h->nx) % (h->ny + 2)) - 1;
        const int z = static_cast<int>(n / layer) - 1;
        const long long c = (static_cast<long long>(z) * h->ny + y) * h->nx + x;
        const uint32_t* kinds = reinterpret_cast<const uint32_t*>(recs.data() + rb * r);
        for (int d = 1; d < h->Q; ++d) {
            const unsigned kd = (kinds[d >> 3] >> ((d & 7) * 4)) & 15u;
            if (kd == kPull) continue;
            if (k < max) {
                if (cell) cell[k] = c;
                if (dir) dir[k] = d;
                if (kind) kind[k] = static_cast<int>(kd);
                if (q) {
                    if (h->rbytes == 8)
                        q[k] = static_cast<float>(reinterpret_cast<const BRec<double>*>(recs.data() + rb * r)->q[d]);
                    else
                        q[k] = reinterpret_cast<const BRec<float>*>(recs.data() + rb * r)->q[d];
                }
            }
            ++k;
        }
    }
    *count = k;
    return FC_LBM_OK;
}

}  // extern "C"
