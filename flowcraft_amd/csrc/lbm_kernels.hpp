// Device side of the Lattice Boltzmann engine (gfx950 / CDNA4).
//
// One fused kernel per substep: pull streaming, the boundary conditions,
// regularized collision with optional Smagorinsky eddy viscosity, and the
// store.  It is specialised at compile time on
//   Q       : 19 (D3Q19) or 15 (D3Q15)
//   S       : population storage (FP32, FP16 offset, BF16 offset, FP64)
//   INPLACE : two SoA copies (pull, ping-pong) or one SoA copy updated in
//             place with the esoteric-pull pairing (Lehmann et al. 2022).
//
// Per-cell arithmetic follows FlowCraft's LbmReference::Step
// (LbmReference.cpp:241-469) and the GLSL kernel (shaders/lbm_step.comp.glsl:
// 98-364) operation for operation: the library is built with
// -ffp-contract=off and IEEE division/sqrt, so the FP32 path reproduces the
// reference's rounding.  The macro field is NOT written every step (the GLSL
// kernel's macroOut); it is derived on demand from the current populations.
//
// HBM layout.  Blocked SoA with a one-cell ghost shell in y and z only: each padded
// row of a z layer is a run of blocks of 64 consecutive x cells, and a block holds the
// Q population planes of its 64 cells one after the other,
//   f[front + ((z+1)*(ny+2) + (y+1))*rs + (x/64)*Q*64 + d*64 + x%64],
//   rs = ceil(nx/64)*Q*64,
// for z in -1..nz (see Geo::Pop).  A 64-lane wave reading 64 consecutive x of one
// plane is one contiguous 256 B (FP32) / 128 B (16-bit) segment, and its Q planes lie
// in one Q*256 B (Q*128 B) region; everything a z layer's update touches lies within
// three consecutive layers, so the step kernel addresses it with 32-bit offsets from one
// per-workgroup base at any lattice size.  Ghost cells are never updated; in the
// in-place layout they hold the populations that leave the domain through a y/z wall,
// which is what makes wall reads race-free (see InPlace below).

#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "lattice.hpp"

namespace fc {

// Population stores stream through the cache hierarchy: a step writes every
// population once and nothing reads it again within the step.
#ifndef FC_LBM_NT_STORES
#define FC_LBM_NT_STORES 1
#endif
template <class T>
__device__ __forceinline__ void PutPop(T* p, T v) {
#if FC_LBM_NT_STORES
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}

// ---------------------------------------------------- layer-relative access
// The step kernel reads and writes populations through a buffer resource whose base
// is the workgroup's lower z layer (uniform, in scalar registers); each access adds a
// per-direction byte offset (a scalar, precomputed on the host) to a per-lane byte
// offset shared by every direction, so no vector address arithmetic is spent per
// population.  The resource's range is the full 32-bit offset space.
using BufRsrc = __amdgpu_buffer_rsrc_t;

__device__ __forceinline__ BufRsrc LayerRsrc(const void* base) {
    // word 3: 32-bit data format, no swizzle, no thread-id offset
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, -1, 0x00020000);
}

#if FC_LBM_NT_STORES
constexpr int kStoreAux = 2;  // non-temporal
#else
constexpr int kStoreAux = 0;
#endif
// Cache policy of the population loads: FC_LBM_LOAD_AUX (default 0) for the ping-pong
// layout and for lattices held in the last-level cache, FC_LBM_INPLACE_LOAD_AUX (default
// 2, non-temporal) for the streamed in-place step (see StepLoadAux).
#ifndef FC_LBM_LOAD_AUX
#define FC_LBM_LOAD_AUX 0
#endif
constexpr int kLoadAux = FC_LBM_LOAD_AUX;
#ifndef FC_LBM_INPLACE_LOAD_AUX
#define FC_LBM_INPLACE_LOAD_AUX 2
#endif

template <class T, int Aux = kLoadAux>
__device__ __forceinline__ T BufLoad(BufRsrc r, uint32_t vo, uint32_t so) {
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, Aux));
    else if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, Aux));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b16(r, vo, so, Aux));
}

template <class T, int Aux = kStoreAux>
__device__ __forceinline__ void BufStore(BufRsrc r, uint32_t vo, uint32_t so, T v) {
    if constexpr (sizeof(T) == 8)
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)), v), r, vo, so, Aux);
    else if constexpr (sizeof(T) == 4)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, vo, so, Aux);
    else
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(uint16_t, v), r, vo, so, Aux);
}

// Cache policy of the in-place layout's stores that land one cell off the block grid
// (odd directions with e_x != 0): each of their runs straddles two lines shared with the
// neighbouring wave.  Written through the cache (policy 0) the two halves of such a line
// meet in L2 before it goes to HBM: in place +1.2% (FP32, 31.8 -> 32.2 GLUPS) and +1%
// (FP16, 60.5 -> 61.1) on the 512x256x256 wing over non-temporal stores.  (All stores
// cached instead costs ping-pong 5% FP32 and 7% FP16.)
#ifndef FC_LBM_SHIFTED_STORE_AUX
#define FC_LBM_SHIFTED_STORE_AUX 0
#endif
constexpr int kShiftedStoreAux = FC_LBM_SHIFTED_STORE_AUX;

// Byte loads of per-cell arrays (flags) from a wave-uniform base: readfirstlane keeps
// the base in scalar registers, so the load issues as scalar base + 32-bit lane offset.
__device__ __forceinline__ uint64_t WaveUniform(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v >> 32));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

template <class T>
__device__ __forceinline__ T LoadU(uint64_t base, uint32_t vo) {
    using GlobalConstBytes = const char __attribute__((address_space(1)))*;
    const GlobalConstBytes p = reinterpret_cast<GlobalConstBytes>(WaveUniform(base));
    return *reinterpret_cast<const T __attribute__((address_space(1)))*>(p + vo);
}

// -------------------------------------------------------------- constants
template <class R>
__host__ __device__ constexpr R K(float f, double d) {
    return sizeof(R) == 4 ? static_cast<R>(f) : static_cast<R>(d);
}

template <int Q, class R>
__host__ __device__ constexpr R W(int d) {
    return sizeof(R) == 4 ? static_cast<R>(Lat<Q>::w[d]) : static_cast<R>(Lat<Q>::wd[d]);
}

// s_setreg operand for bits 6..7 of the MODE register (half/double denormal handling):
// hwreg id 1 (MODE) | offset 6 << 6 | (size 2 - 1) << 11.
constexpr int kHwRegModeFp16Denorm = 1 | (6 << 6) | (1 << 11);

// -------------------------------------------------------- storage formats
// Raw/Enc move between the stored element and the step's working value: the
// population f itself for FP32/FP64, the scaled offset g = (f - w) * S for the
// offset formats, whose step runs entirely on g (see CollideOffset).  Load/Store
// convert to and from f for readbacks, seeding and state dump/restore.
// FP32: the reference's float populations.
template <int Q>
struct StoreF32 {
    using T = float;
    using R = float;
    static constexpr bool kOffset = false;
    static constexpr bool kStochastic = false;
    static __device__ __forceinline__ R Load(T v, int) { return v; }
    static __device__ __forceinline__ T Store(R f, int) { return f; }
    static __device__ __forceinline__ R Raw(T v) { return v; }
    static __device__ __forceinline__ T Enc(R v) { return v; }
    static __device__ __forceinline__ void StepMode() {}
    static __device__ __forceinline__ T EncStep(R v) { return v; }
};

// FP16 offset: g = (f - w) * 4096 as IEEE binary16 (RNE); arithmetic in FP32.  Offsets
// that are tiny in the IEEE sense — below the smallest normal half 2^-14 after rounding
// to 11 significant bits with an unbounded exponent ("tininess after rounding") — are
// stored as a zero of the same sign: |f - w| < 1.5e-8, never a subnormal half, so the
// format does not depend on a device's half-precision denormal mode.  That is exactly
// what the hardware conversion produces with half denormals flushed (checked on
// MI355X over the values bracketing 2^-14, ties included): the step kernel sets that
// mode (StepMode) and converts with one instruction (EncStep); Enc spells the rule out
// for kernels that run in the default mode.
template <int Q>
struct StoreF16 {
    using T = _Float16;
    using R = float;
    static constexpr bool kOffset = true;
    static constexpr bool kStochastic = false;
    static constexpr float kScale = 4096.0f;
    static __device__ __forceinline__ R Raw(T v) { return static_cast<float>(v); }
    static __device__ __forceinline__ T Enc(R g) {
        const uint32_t b = __float_as_uint(g);
        const uint32_t a = b & 0x7FFFFFFFu;
        if (a < 0x38800000u) {  // |g| < 2^-14: round to 11 significant bits, unbounded exponent
            uint32_t hb = a >> 13;
            const uint32_t rest = a & 0x1FFFu;
            if (rest > 0x1000u || (rest == 0x1000u && (hb & 1u))) ++hb;
            if ((hb << 13) < 0x38800000u) return __builtin_bit_cast(T, static_cast<uint16_t>((b >> 16) & 0x8000u));
        }
        return static_cast<T>(g);
    }
    // Half denormals flushed (MODE.FP_DENORM for f16/f64 = 0) for the rest of the wave.
    static __device__ __forceinline__ void StepMode() { __builtin_amdgcn_s_setreg(kHwRegModeFp16Denorm, 0); }
    static __device__ __forceinline__ T EncStep(R g) { return static_cast<T>(g); }
    static __device__ __forceinline__ R Load(T v, int d) { return Lat<Q>::w[d] + Raw(v) * (1.0f / 4096.0f); }
    static __device__ __forceinline__ T Store(R f, int d) { return Enc((f - Lat<Q>::w[d]) * 4096.0f); }
};

// BF16 offset: g = f - w as bfloat16 (RNE, NaN kept quiet); arithmetic in FP32.
template <int Q>
struct StoreBF16 {
    using T = uint16_t;
    using R = float;
    static constexpr bool kOffset = true;
    static constexpr bool kStochastic = false;
    static constexpr float kScale = 1.0f;
    static __device__ __forceinline__ R Raw(T v) { return __uint_as_float(static_cast<uint32_t>(v) << 16); }
    static __device__ __forceinline__ T Enc(R g) {
        uint32_t b = __float_as_uint(g);
        if ((b & 0x7FFFFFFFu) > 0x7F800000u) return static_cast<T>((b >> 16) | 0x40u);
        b += 0x7FFFu + ((b >> 16) & 1u);
        return static_cast<T>(b >> 16);
    }
    static __device__ __forceinline__ R Load(T v, int d) { return Lat<Q>::w[d] + Raw(v); }
    static __device__ __forceinline__ T Store(R f, int d) { return Enc(f - Lat<Q>::w[d]); }
    static __device__ __forceinline__ void StepMode() {}
    static __device__ __forceinline__ T EncStep(R g) { return Enc(g); }
};

// FP64: the whole step in double (GPU twin of the FP64 oracle).
template <int Q>
struct StoreF64 {
    using T = double;
    using R = double;
    static constexpr bool kOffset = false;
    static constexpr bool kStochastic = false;
    static __device__ __forceinline__ R Load(T v, int) { return v; }
    static __device__ __forceinline__ T Store(R f, int) { return f; }
    static __device__ __forceinline__ R Raw(T v) { return v; }
    static __device__ __forceinline__ T Enc(R v) { return v; }
    static __device__ __forceinline__ void StepMode() {}
    static __device__ __forceinline__ T EncStep(R v) { return v; }
};

// Stochastic rounding of the step's stores (FC_LBM_STORE_F16_SR / _BF16_SR).  Rounding
// to nearest loses any change of a stored offset smaller than half a unit in the last
// place, every step, in the same direction: a slow drift (the outlet's density relaxation
// moves the outlet column by about one FP16 unit per step) stalls.  Stochastic rounding
// rounds up with probability equal to the discarded fraction, so the stored value is
// right on average and slow drifts survive.  The random bits are a pure function of
// (cell, step, direction) — the same for both layouts, any z-chunking and the CPU oracle:
// the cell's 32-bit hash (the "lowbias32" integer finaliser of C. Wellons' hash-prospector
// applied to cell*0x9E3779B1 ^ step*0x85EBCA77), advanced by d golden-ratio increments
// (a Weyl sequence) for direction d.  Seeding and state restore keep rounding to nearest.
__host__ __device__ __forceinline__ uint32_t SrHash(uint32_t cell, uint32_t step) {
    uint32_t h = (cell * 0x9E3779B1u) ^ (step * 0x85EBCA77u);
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return h;
}
constexpr uint32_t kSrWeyl = 0x9E3779B9u;

// FP16 offsets, stochastically rounded: the top 13 of the direction's random bits are
// added to the float's bit pattern below the half's last mantissa bit and the result is
// truncated (round toward zero, v_cvt_pkrtz_f16_f32, two populations per instruction) —
// the magnitude rounds up with probability (discarded bits) / 2^13.  With half
// denormals flushed (StepMode) a sum below 2^-14 in magnitude is stored as a zero of its
// sign; truncation never carries a value across 2^-14, so that is the whole rule.
template <int Q>
struct StoreF16SR : StoreF16<Q> {
    static constexpr bool kStochastic = true;
    static __device__ __forceinline__ float Dither(float g, uint32_t r) {
        return __uint_as_float(__float_as_uint(g) + (r >> 19));
    }
    static __device__ __forceinline__ void EncodeSR(const float (&fo)[Q], _Float16 (&enc)[Q], uint32_t h) {
        StaticFor<0, (Q + 1) / 2>([&](auto K) {
            constexpr int d = 2 * K;
            const float a = Dither(fo[d], h + static_cast<uint32_t>(d) * kSrWeyl);
            const float b = d + 1 < Q ? Dither(fo[d + 1 < Q ? d + 1 : d], h + static_cast<uint32_t>(d + 1) * kSrWeyl) : 0.0f;
            const uint32_t p = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(a, b));
            enc[d] = __builtin_bit_cast(_Float16, static_cast<uint16_t>(p & 0xFFFFu));
            if constexpr (d + 1 < Q) enc[d + 1] = __builtin_bit_cast(_Float16, static_cast<uint16_t>(p >> 16));
        });
    }
};

// BF16 offsets, stochastically rounded: the top 16 random bits added to the float's bit
// pattern, then its high half kept (truncation).
template <int Q>
struct StoreBF16SR : StoreBF16<Q> {
    static constexpr bool kStochastic = true;
    static __device__ __forceinline__ void EncodeSR(const float (&fo)[Q], uint16_t (&enc)[Q], uint32_t h) {
        StaticFor<0, Q>([&](auto I) {
            constexpr int d = I;
            const uint32_t r = h + static_cast<uint32_t>(d) * kSrWeyl;
            enc[d] = static_cast<uint16_t>((__float_as_uint(fo[d]) + (r >> 16)) >> 16);
        });
    }
};

// ------------------------------------------------------ boundary records
// One per fluid cell with a solid source.  kind: 4 bits per direction.
// side[d] (in-place layout only) holds the cell's own post-collision f_d of the
// previous step for q >= 1/2 Bouzidi links: in place, that slot is consumed and
// overwritten by the downstream neighbour during the step.
template <class R>
struct BRec {
    uint32_t kind[3];
    uint32_t pad;
    R q[19];
    R side[19];
};

__device__ __forceinline__ unsigned KindOf(const uint32_t k[3], int d) { return (k[d >> 3] >> ((d & 7) * 4)) & 15u; }

// ------------------------------------------------------------ step args
struct StepArgs {
    void* pop_src;   // ping-pong: state t;  in place: the single copy
    void* pop_dst;   // ping-pong: state t+1; in place: == pop_src
    const uint8_t* flags;
    const int* bidx;
    void* recs;      // BRec<R>*
    void* side_mom;  // R[2 parity][2 column][ny*nz][4]: moments of columns 1 and nx-2
    int* diverged;
    long long P;     // padded cells per plane set: nx*(ny+2)*(nz+2)
    long long layer_bytes;  // (ny+2)*rs*sizeof(T): one padded z layer of population blocks
    long long base_bytes;   // (PopFront - StepBias)*sizeof(T): workgroup base of layer z = 0
    uint32_t ld_off[19];    // bulk-pull byte offset of direction d from the lower layer's base
    uint32_t st_off[19];    // store byte offset of direction d (interior anchors in place)
    uint32_t prev_off[19];  // byte offset of the cell's own outgoing f_d of the previous step
    float inlet_f[19];      // inlet populations (working values: f, or the stored offsets)
    double inlet_d[19];     // the same for the FP64 step
    int nx, ny, nz;
    int z0, zc;
    int parity;      // step index & 1
    uint32_t step;   // step index (low 32 bits): the stochastic-rounding formats' random stream
    int reverse, fsy, fsz, pery, perz;
    float tau, u_in, cs;
    double tau_d, u_in_d, cs_d;
};

// Population blocks.  Runs of kBlk = 64 consecutive x cells (one wavefront) keep their Q
// populations together: element (plane pl, padded cell x, y, z) is at
//   front + ((z+1)*(ny+2) + (y+1)) * rs + (x >> 6) * Q*64 + pl*64 + (x & 63),
// rs = ceil(nx/64) * Q*64 (one padded row of blocks).  A wave's 19 (15) plane accesses
// of its own run then fall in one 19*64-element region instead of 19 planes a layer
// apart, which keeps HBM pages open: the bare D3Q19 pull pattern moves 5.89 TB/s with
// this layout against 5.57 TB/s on plain planes (both with non-temporal stores, MI355X).
// x may be -1 or nx (the in-place anchors of the inlet and outlet columns): x = -1 is
// the last slot of the previous block run, x = nx the next one — cells of the
// neighbouring row when nx is a multiple of 64, unused padding slots otherwise; either
// way slots that no cell owns during a step.  `front` (one block) and an equal margin
// after the last row keep every such slot, and every speculative load of the step
// kernel, inside the allocation.
constexpr int kBlk = 64;

__host__ __device__ inline long long PopRowStride(int nx, int Q) {
    return static_cast<long long>((nx + kBlk - 1) / kBlk) * Q * kBlk;
}

// Elements before row 0 of the padded layer z = -1: one padded row plus one block, so
// that the step kernel's per-workgroup base (see StepBias) stays inside the allocation.
__host__ __device__ inline long long PopFront(long long rs, int Q) { return rs + static_cast<long long>(Q) * kBlk; }

// The step kernel's scalar offsets reach one row and one cell below a cell's own row in
// the layer below; they are stored biased by StepBias elements (and the workgroup base
// lowered by as much) so that every offset is non-negative.
__host__ __device__ inline long long StepBias(long long rs) { return rs + kBlk; }

struct Geo {
    int nx, ny, nz, pery, perz;
    int sy, sz;  // padded strides of the per-cell arrays: sy = nx, sz = LS = nx*(ny+2)
    int Q;
    long long rs;  // population row stride, PopRowStride(nx, Q)
    // index of a padded cell in per-cell arrays (flags, boundary-record index); 64-bit, a
    // lattice may hold more than 2^31 padded cells
    __device__ __forceinline__ long long Idx(int x, int y, int z) const {
        return static_cast<long long>(z + 1) * sz + (y + 1) * sy + x;
    }
    // element index of population plane `pl` of padded cell (x, y, z) (x may be -1 or nx)
    __device__ __forceinline__ long long Pop(int pl, int x, int y, int z) const {
        return PopFront(rs, Q) + (static_cast<long long>(z + 1) * (ny + 2) + (y + 1)) * rs +
               static_cast<long long>(x >> 6) * (Q * kBlk) + pl * kBlk + (x & (kBlk - 1));
    }
    // neighbour coordinates with the wrap rules of Nbr
    __device__ __forceinline__ void NbrXYZ(int x, int y, int z, int dx, int dy, int dz, int& ax, int& ay,
                                           int& az) const {
        ax = x + dx;
        ay = y + dy;
        az = z + dz;
        ax = ax >= nx ? ax - nx : (ax < 0 ? ax + nx : ax);
        if (pery) ay = ay >= ny ? ay - ny : (ay < 0 ? ay + ny : ay);
        if (perz) az = az >= nz ? az - nz : (az < 0 ? az + nz : az);
    }
    // Neighbour index with the layout's wrap rules: x always wraps (only the
    // inlet/outlet columns ever reach across it, and they never read a slot);
    // y/z wrap when periodic, otherwise land in the ghost shell.
    __device__ __forceinline__ long long Nbr(int x, int y, int z, int dx, int dy, int dz) const {
        int ax = x + dx, ay = y + dy, az = z + dz;
        ax = ax >= nx ? ax - nx : (ax < 0 ? ax + nx : ax);
        if (pery) ay = ay >= ny ? ay - ny : (ay < 0 ? ay + ny : ay);
        if (perz) az = az >= nz ? az - nz : (az < 0 ? az + nz : az);
        return Idx(ax, ay, az);
    }
};

// Per-lane byte offsets of a cell within its z layer of population blocks (plane 0):
// c for the cell itself; m and p for the slot one cell to the left / right when that
// slot is in the neighbouring block (lanes 0 and 63 of a run), else equal to c.  The
// step kernel's scalar per-direction offsets include the in-block x shift; adding m or p
// in place of c moves the two edge lanes across the block boundary.
struct PopLane {
    uint32_t c, m, p;
};

template <int Q, class T>
__device__ __forceinline__ PopLane MakePopLane(int x, int y, long long rs) {
    const uint32_t c = static_cast<uint32_t>(((y + 1) * rs + (x >> 6) * (Q * kBlk) + (x & (kBlk - 1))) * sizeof(T));
    constexpr uint32_t jump = static_cast<uint32_t>((Q - 1) * kBlk * sizeof(T));
    const int xi = x & (kBlk - 1);
    return PopLane{c, xi == 0 ? c - jump : c, xi == kBlk - 1 ? c + jump : c};
}

template <int DX>
__device__ __forceinline__ uint32_t LaneSel(const PopLane& l) {
    if constexpr (DX < 0)
        return l.m;
    else if constexpr (DX > 0)
        return l.p;
    else
        return l.c;
}

__device__ __forceinline__ uint32_t LaneSelRT(const PopLane& l, int dx) { return dx < 0 ? l.m : (dx > 0 ? l.p : l.c); }

// -------------------------------------------------------------- layouts
// Ping-pong pull: read f_d(x - e_d) from the source copy, write f_d(x) into the
// destination copy.  A cell's own previous outgoing f_d(x) is src[d][x].
struct PingPong {
    static constexpr bool kInPlace = false;
};

// In-place esoteric pull.  Pair (a, b = a+1), e_b = -e_a, anchor A = x + e_a.
//   even step: in_a <- (b, x)   in_b <- (a, A)   out_a -> (a, A)   out_b -> (b, x)
//   odd step : in_a <- (a, x)   in_b <- (b, A)   out_a -> (b, A)   out_b -> (a, x)
//   rest     : (0, x) both ways.
// Each slot is read and then written by exactly one cell per step, so cells may
// run in any order.  The slot that holds a cell's own outgoing f_d of the
// previous step is the store location at the other parity; it is owned this step
// by the cell x + e_d, which is never active when the reference needs that value
// (a solid cell for bounce-back, a ghost cell for y/z walls).
struct InPlace {
    static constexpr bool kInPlace = true;
};

// Slot of population d read (load) / written (store) at a parity: its plane, and
// whether it sits at the cell itself or at the anchor of d's pair.
__host__ __device__ inline void LoadSlot(int d, int par, int& plane, bool& at_anchor) {
    if (d == 0) {
        plane = 0;
        at_anchor = false;
    } else if (d & 1) {
        plane = par == 0 ? d + 1 : d;
        at_anchor = false;
    } else {
        plane = par == 0 ? d - 1 : d;
        at_anchor = true;
    }
}

__host__ __device__ inline void StoreSlot(int d, int par, int& plane, bool& at_anchor) {
    if (d == 0) {
        plane = 0;
        at_anchor = false;
    } else if (d & 1) {
        plane = par == 0 ? d : d + 1;
        at_anchor = true;
    } else {
        plane = par == 0 ? d : d - 1;
        at_anchor = false;
    }
}

template <int Q>
__host__ __device__ constexpr int PairAnchor(int d) {
    return d == 0 ? 0 : ((d & 1) ? d : d - 1);
}

// Element index of an in-place slot of population d of cell (x,y,z) at a parity.  The
// anchor wraps across periodic y/z seams; in x it does not: the inlet and outlet
// columns, whose anchors would leave the row, use the neighbouring row's end cell
// (or the ghost row), a slot no other cell owns and that neither column reads back
// during a step.
template <int Q, bool kStore>
__device__ __forceinline__ long long SlotIndex(const Geo& g, int d, int par, int x, int y, int z) {
    int pl;
    bool at_anchor;
    if constexpr (kStore)
        StoreSlot(d, par, pl, at_anchor);
    else
        LoadSlot(d, par, pl, at_anchor);
    if (!at_anchor) return g.Pop(pl, x, y, z);
    const int a = PairAnchor<Q>(d);
    int ax, ay, az;
    g.NbrXYZ(x, y, z, Lat<Q>::ex[a], Lat<Q>::ey[a], Lat<Q>::ez[a], ax, ay, az);
    return g.Pop(pl, x + Lat<Q>::ex[a], ay, az);
}

// Element index holding the current post-collision population d of cell (x,y,z)
// after `t` steps, and the copy it is in.
template <int Q, class Lay>
__device__ __forceinline__ long long CurrentIndex(const Geo& g, long long t, int x, int y, int z, int d) {
    if constexpr (!Lay::kInPlace)
        return g.Pop(d, x, y, z);
    else
        return SlotIndex<Q, true>(g, d, static_cast<int>((t - 1) & 1), x, y, z);
}

template <class Lay>
__device__ __forceinline__ const void* CurrentCopy(const void* pops0, const void* pops1, long long t) {
    return (!Lay::kInPlace && (t & 1)) ? pops1 : pops0;
}

// Current post-collision population d of cell (x,y,z) after `t` steps.
template <int Q, class S, class Lay>
__device__ __forceinline__ typename S::R CurrentPop(const void* pops0, const void* pops1, long long t, const Geo& g,
                                                   int x, int y, int z, int d) {
    const auto* f = static_cast<const typename S::T*>(CurrentCopy<Lay>(pops0, pops1, t));
    return S::Load(f[CurrentIndex<Q, Lay>(g, t, x, y, z, d)], d);
}

// The same slot as the step's working value (S::Raw): f, or the offset g.
template <int Q, class S, class Lay>
__device__ __forceinline__ typename S::R CurrentRaw(const void* pops0, const void* pops1, long long t, const Geo& g,
                                                   int x, int y, int z, int d) {
    const auto* f = static_cast<const typename S::T*>(CurrentCopy<Lay>(pops0, pops1, t));
    return S::Raw(f[CurrentIndex<Q, Lay>(g, t, x, y, z, d)]);
}

// ------------------------------------------------- compile-time lattice sums
// The reference evaluates sums like ex*ux + ey*uy + ez*uz with ex in {-1,0,1}.  Terms
// with a zero coefficient add an exact +-0, and multiplying by +-1 is exact, so the
// sums below skip those terms and keep the remaining ones in the reference's order:
// for finite operands the result is identical, except possibly the sign of an exact
// zero.  This removes most of the collision's arithmetic without changing rounding.
template <int B, int E, class F>
__host__ __device__ __forceinline__ void StaticFor(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        StaticFor<B + 1, E>(f);
    }
}

template <int c, class R>
__host__ __device__ __forceinline__ void AccC(R& a, R v) {
    if constexpr (c == 1)
        a += v;
    else if constexpr (c == -1)
        a -= v;
}

template <int c, class R>
__host__ __device__ __forceinline__ R MulC(R v) {
    if constexpr (c == 1)
        return v;
    else
        return -v;
}

// c1*u + c2*v + c3*w, left to right, zero terms skipped.
template <int c1, int c2, int c3, class R>
__host__ __device__ __forceinline__ R Dot3(R u, R v, R w) {
    if constexpr (c1 != 0) {
        R r = MulC<c1>(u);
        if constexpr (c2 != 0) r = r + MulC<c2>(v);
        if constexpr (c3 != 0) r = r + MulC<c3>(w);
        return r;
    } else if constexpr (c2 != 0) {
        R r = MulC<c2>(v);
        if constexpr (c3 != 0) r = r + MulC<c3>(w);
        return r;
    } else if constexpr (c3 != 0) {
        return MulC<c3>(w);
    } else {
        return R(0);
    }
}

// Equilibrium of direction d (LbmReference.cpp:9-13): w*rho*(((1+3eu)+(4.5eu)eu)-1.5u2),
// u2 = ux*ux + uy*uy + uz*uz computed once by the caller.
template <int Q, int d, class R>
__host__ __device__ __forceinline__ R FeqD(R rho, R ux, R uy, R uz, R u2) {
    const R eu = Dot3<Lat<Q>::ex[d], Lat<Q>::ey[d], Lat<Q>::ez[d]>(ux, uy, uz);
    return W<Q, R>(d) * rho * (R(1) + R(3) * eu + R(4.5) * eu * eu - R(1.5) * u2);
}

template <int Q, class R>
__host__ __device__ __forceinline__ void FeqAll(R rho, R ux, R uy, R uz, R (&out)[Q]) {
    const R u2 = ux * ux + uy * uy + uz * uz;
    StaticFor<0, Q>([&](auto I) {
        constexpr int d = I;
        out[d] = FeqD<Q, d, R>(rho, ux, uy, uz, u2);
    });
}

// Moments (rho, rho*u) of a population set in index order (LbmReference.cpp:43-61, before
// the division by rho).
template <int Q, class R>
__device__ __forceinline__ void MomentSums(const R (&f)[Q], R& rho, R& jx, R& jy, R& jz) {
    rho = R(0);
    jx = R(0);
    jy = R(0);
    jz = R(0);
    StaticFor<0, Q>([&](auto I) {
        constexpr int d = I;
        rho += f[d];
        AccC<Lat<Q>::ex[d]>(jx, f[d]);
        AccC<Lat<Q>::ey[d]>(jy, f[d]);
        AccC<Lat<Q>::ez[d]>(jz, f[d]);
    });
}

// Regularized collision with Smagorinsky eddy viscosity (LbmReference.cpp:385-453):
// f holds the gathered populations on entry; fo receives the post-collision ones.
// Returns true when the cell diverged (non-finite, rho <= 0 or |u|^2 > 4).
template <int Q, class R>
__device__ __forceinline__ bool Collide(const R (&f)[Q], R (&fo)[Q], R tau, R cs) {
    R rho, ux, uy, uz;
    MomentSums<Q, R>(f, rho, ux, uy, uz);
    if (rho > K<R>(1e-6f, 1e-6)) {
        ux /= rho;
        uy /= rho;
        uz /= rho;
    }
    const bool bad = !__builtin_isfinite(rho) || !__builtin_isfinite(ux) || !__builtin_isfinite(uy) ||
                     !__builtin_isfinite(uz) || rho <= R(0) || (ux * ux + uy * uy + uz * uz) > R(4);
    R feq[Q];
    FeqAll<Q, R>(rho, ux, uy, uz, feq);
    R pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0;
    StaticFor<0, Q>([&](auto I) {
        constexpr int d = I;
        constexpr int cx = Lat<Q>::ex[d], cy = Lat<Q>::ey[d], cz = Lat<Q>::ez[d];
        const R fneq = f[d] - feq[d];
        if constexpr (cx != 0) pxx += fneq;
        if constexpr (cy != 0) pyy += fneq;
        if constexpr (cz != 0) pzz += fneq;
        AccC<cx * cy>(pxy, fneq);
        AccC<cx * cz>(pxz, fneq);
        AccC<cy * cz>(pyz, fneq);
    });
    const R piMag = sqrt(pxx * pxx + pyy * pyy + pzz * pzz + R(2) * (pxy * pxy + pxz * pxz + pyz * pyz));
    const R rmax = rho < K<R>(1e-6f, 1e-6) ? K<R>(1e-6f, 1e-6) : rho;
    const R kSmag = R(18) * K<R>(1.41421356f, 1.41421356);
    const R tauEff = R(0.5) * (tau + sqrt(tau * tau + kSmag * cs * cs * piMag / rmax));
    const R omega = R(1) / tauEff;
    const R keep = R(1) - omega;
    constexpr R kCs2 = R(1) / R(3);
    StaticFor<0, Q>([&](auto I) {
        constexpr int d = I;
        constexpr int cx = Lat<Q>::ex[d], cy = Lat<Q>::ey[d], cz = Lat<Q>::ez[d];
        constexpr R qxx = R(cx * cx) - kCs2, qyy = R(cy * cy) - kCs2, qzz = R(cz * cz) - kCs2;
        R sum = qxx * pxx;
        sum = sum + qyy * pyy;
        sum = sum + qzz * pzz;
        if constexpr (cx * cy != 0) sum = sum + R(2 * cx * cy) * pxy;
        if constexpr (cx * cz != 0) sum = sum + R(2 * cx * cz) * pxz;
        if constexpr (cy * cz != 0) sum = sum + R(2 * cy * cz) * pyz;
        const R reg = (W<Q, R>(d) * R(4.5)) * sum;
        fo[d] = feq[d] + keep * reg;
    });
    return bad;
}

// --------------------------------------------------- offset-space collision
// The same regularized BGK + Smagorinsky model for the offset storage formats,
// written on g_i = (f_i - w_i) * S in moment space: with rho = 1 + drho,
//   drho = sum g / S,  j = sum e g / S,  Pi_neq = sum e e g / S - j u - (drho/3) I,
//   T = j u + (1 - 1/tau_eff) Pi_neq,
//   g_i' = W_i (drho + 3 e.j + 4.5 Q_i : T),  Q_i = e e - I/3,  W_i = w_i S.
// Identical to the reference's population-space form in exact arithmetic (the
// lattice moments of f^eq are exact to second order for D3Q19 and D3Q15); it works
// on small offsets instead of O(w) populations.  Both halves run per opposite pair
// (i, i+1), e_{i+1} = -e_i: the moments from the pair sums g_i + g_{i+1} (even
// moments) and differences g_i - g_{i+1} (odd moments), and the reconstruction as
// g_i' = A + B, g_{i+1}' = A - B with the even part A = W drho + 4.5 W (e e : T -
// tr T / 3) and the odd part B = 3 W e.j, from moments scaled once per weight shell.
// That is about half the arithmetic of a per-direction evaluation.  The CPU oracle
// emulates exactly these operations for the offset formats.
template <int Q, class S>
__host__ __device__ __forceinline__ void ReconstructOffset(float drho, float jx, float jy, float jz, float txx, float tyy,
                                                  float tzz, float txy, float txz, float tyz, float (&go)[Q]) {
    const float tr = (txx + tyy) + tzz;
    {
        constexpr float W = Lat<Q>::w[0] * S::kScale;
        constexpr float W15 = 1.5f * W;
        go[0] = W * drho - W15 * tr;
    }
    StaticFor<0, (Q - 1) / 2>([&](auto K) {
        constexpr int d = 2 * K + 1;
        constexpr int cx = Lat<Q>::ex[d], cy = Lat<Q>::ey[d], cz = Lat<Q>::ez[d];
        constexpr float W = Lat<Q>::w[d] * S::kScale;
        constexpr float W15 = 1.5f * W, W3 = 3.0f * W, W45 = 4.5f * W, W9 = 9.0f * W;
        // even part: the terms shared by every pair of the shell are the same
        // expressions, evaluated once
        float a = W * drho - W15 * tr;
        if constexpr (cx != 0) a = a + W45 * txx;
        if constexpr (cy != 0) a = a + W45 * tyy;
        if constexpr (cz != 0) a = a + W45 * tzz;
        AccC<cx * cy>(a, W9 * txy);
        AccC<cx * cz>(a, W9 * txz);
        AccC<cy * cz>(a, W9 * tyz);
        const float b = Dot3<cx, cy, cz>(W3 * jx, W3 * jy, W3 * jz);
        go[d] = a + b;
        go[d + 1] = a - b;
    });
}

// Equilibrium offsets of (1 + drho, u).
template <int Q, class S>
__host__ __device__ __forceinline__ void EqOffset(float drho, float ux, float uy, float uz, float (&go)[Q]) {
    const float rho = 1.0f + drho;
    const float jx = rho * ux, jy = rho * uy, jz = rho * uz;
    ReconstructOffset<Q, S>(drho, jx, jy, jz, jx * ux, jy * uy, jz * uz, jx * uy, jx * uz, jy * uz, go);
}

template <int Q, class S>
__device__ __forceinline__ bool CollideOffset(const float (&g)[Q], float (&go)[Q], float tau, float cs) {
    constexpr float invS = 1.0f / S::kScale;
    // Sums over the pairs in index order.  The accumulators start at -0, the identity
    // of IEEE addition, so the first term costs nothing.
    float sg = g[0], Jx = -0.0f, Jy = -0.0f, Jz = -0.0f, gxx = -0.0f, gyy = -0.0f, gzz = -0.0f, gxy = -0.0f,
          gxz = -0.0f, gyz = -0.0f;
    StaticFor<0, (Q - 1) / 2>([&](auto K) {
        constexpr int d = 2 * K + 1;
        constexpr int cx = Lat<Q>::ex[d], cy = Lat<Q>::ey[d], cz = Lat<Q>::ez[d];
        const float ps = g[d] + g[d + 1];
        const float pd = g[d] - g[d + 1];
        sg += ps;
        AccC<cx>(Jx, pd);
        AccC<cy>(Jy, pd);
        AccC<cz>(Jz, pd);
        if constexpr (cx != 0) gxx += ps;
        if constexpr (cy != 0) gyy += ps;
        if constexpr (cz != 0) gzz += ps;
        AccC<cx * cy>(gxy, ps);
        AccC<cx * cz>(gxz, ps);
        AccC<cy * cz>(gyz, ps);
    });
    const float drho = sg * invS;
    const float jx = Jx * invS, jy = Jy * invS, jz = Jz * invS;
    const float rho = 1.0f + drho;
    const float r = rho > 1e-6f ? 1.0f / rho : 1.0f;
    const float ux = jx * r, uy = jy * r, uz = jz * r;
    const bool bad = !__builtin_isfinite(rho) || !__builtin_isfinite(ux) || !__builtin_isfinite(uy) ||
                     !__builtin_isfinite(uz) || rho <= 0.0f || (ux * ux + uy * uy + uz * uz) > 4.0f;
    constexpr float c3 = 1.0f / 3.0f;
    const float pxx = (gxx * invS - jx * ux) - drho * c3;
    const float pyy = (gyy * invS - jy * uy) - drho * c3;
    const float pzz = (gzz * invS - jz * uz) - drho * c3;
    const float pxy = gxy * invS - jx * uy;
    const float pxz = gxz * invS - jx * uz;
    const float pyz = gyz * invS - jy * uz;
    const float piMag = sqrtf(pxx * pxx + pyy * pyy + pzz * pzz + 2.0f * (pxy * pxy + pxz * pxz + pyz * pyz));
    const float rmax = rho < 1e-6f ? 1e-6f : rho;
    const float kSmag = 18.0f * 1.41421356f;
    const float tauEff = 0.5f * (tau + sqrtf(tau * tau + kSmag * cs * cs * piMag / rmax));
    const float keep = 1.0f - 1.0f / tauEff;
    ReconstructOffset<Q, S>(drho, jx, jy, jz, jx * ux + keep * pxx, jy * uy + keep * pyy, jz * uz + keep * pzz,
                            jx * uy + keep * pxy, jx * uz + keep * pxz, jy * uz + keep * pyz, go);
    return bad;
}

// ----------------------------------------------------------- step kernel
// Work decomposition: one workgroup = a 256-cell tile of one z layer, TileX cells
// along x by 256 / TileX rows, 256 threads = 4 waves = 4 population blocks, no integer
// division per thread.  On the 512x256x256 lattice (ping-pong, one cell per thread, in
// one run) 256x1 tiles measured 36.8 (FP32) / 72.2 (FP16) GLUPS, 128x2 37.3 / 70.5, 64x4
// 37.0 / 68.7: 128x2 for 4- and 8-byte populations, 256x1 for 2-byte ones, whose rows
// are half as long in bytes.  Workgroups are dealt to the 8 XCDs round-robin by
// hardware id; the id is remapped so that each XCD processes one contiguous range of
// tiles, keeping the cache lines shared by neighbouring tiles (x-shifted pulls, y/z
// neighbour rows) inside one XCD's L2 (without the remap: 36.4 / 61.2 GLUPS, 256x1 with
// the earlier two-cell 16-bit kernel).  FC_LBM_TILE_X=64|128|256 fixes the width for
// every format.
#ifndef FC_LBM_TILE_X
#define FC_LBM_TILE_X 0
#endif
#ifndef FC_LBM_THREADS_16
#define FC_LBM_THREADS_16 256
#endif
// Threads per workgroup of the step: 256, or FC_LBM_THREADS_16 for 2-byte populations.
// Whole 512-cell rows as one 512-thread tile (fewer x-shifted pulls re-fetching the
// neighbour tile's lines) measured slower on the 512x256x256 wing, same fields: FP16
// ping-pong 70.8 -> 63.9 GLUPS, in place 62.1 -> 60.3, BF16 ping-pong 69.4 -> 59.1.
__host__ __device__ constexpr int StepThreads(int tbytes) { return tbytes == 2 ? FC_LBM_THREADS_16 : 256; }
__host__ __device__ constexpr int TileX(int tbytes) {
    return FC_LBM_TILE_X > 0 ? FC_LBM_TILE_X : (tbytes == 2 ? StepThreads(2) : 128);
}
__host__ __device__ constexpr int TileY(int tbytes) { return StepThreads(tbytes) / TileX(tbytes); }
static_assert(TileX(2) % 64 == 0 && StepThreads(2) % TileX(2) == 0 && TileX(4) % 64 == 0 &&
                  StepThreads(4) % TileX(4) == 0,
              "a tile row is whole waves");

__device__ __forceinline__ unsigned XcdContiguous(unsigned hw, unsigned n) {
    const unsigned q = n >> 3, r = n & 7u, xcd = hw & 7u, slot = hw >> 3;
    return xcd < r ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
}

// Chunked remap (FC_LBM_XCD_REMAP=2, FC_LBM_XCD_CHUNK=C): each XCD receives runs of C
// consecutive tiles, and the eight XCDs take neighbouring runs, so the whole device works
// on one compact band of the lattice while tiles that share cache lines within a run stay
// in one L2.  Ids past the last whole group of 8*C map to themselves.  Measured on the
// 512x256x256 wing, D3Q19 ping-pong, FP32 / stochastically rounded FP16 / FP32 in place
// (GLUPS): contiguous ranges (the default) 37.1 / 69.2 / 32.1, C = 4 36.5 / 65.2 / 29.8,
// C = 16 35.0 / 58.2 / 30.9, hardware order (FC_LBM_XCD_REMAP=0) 34.3 / 66.0 / 27.5.
template <unsigned C>
__device__ __forceinline__ unsigned XcdChunked(unsigned hw, unsigned n) {
    const unsigned full = n / (8u * C) * (8u * C);
    if (hw >= full) return hw;
    const unsigned xcd = hw & 7u, slot = hw >> 3;
    return ((slot / C) * 8u + xcd) * C + slot % C;
}

// Stored form of a cell's post-collision working values (inside the step kernel, which
// has set the storage format's StepMode).
// `h`: the cell's stochastic-rounding hash (SrHash), used by the *SR formats only.
template <int Q, class S>
__device__ __forceinline__ void EncodeCell(const typename S::R (&fo)[Q], typename S::T (&enc)[Q], uint32_t h) {
    if constexpr (S::kStochastic) {
        S::EncodeSR(fo, enc, h);
    } else {
        StaticFor<0, Q>([&](auto I) {
            constexpr int d = I;
            enc[d] = S::EncStep(fo[d]);
        });
    }
}

template <class S>
__device__ __forceinline__ uint32_t CellSrHash(const StepArgs& a, int x, int y, int z) {
    if constexpr (S::kStochastic)
        // the global cell index (z*ny + y)*nx + x, exact below 2^32 cells (unsigned wrap)
        return SrHash((static_cast<uint32_t>(z) * static_cast<uint32_t>(a.ny) + static_cast<uint32_t>(y)) *
                              static_cast<uint32_t>(a.nx) +
                          static_cast<uint32_t>(x),
                      a.step);
    else
        return 0u;
}

template <int Q, class S>
__device__ __forceinline__ void WriteMoments(typename S::R* side_mom, const typename S::T (&enc)[Q], int par,
                                             int col, int plane_yz, int zy) {
    using R = typename S::R;
    R v[Q];
    StaticFor<0, Q>([&](auto I) {
        constexpr int d = I;
        v[d] = S::Raw(enc[d]);
    });
    R m0, m1, m2, m3;
    MomentSums<Q, R>(v, m0, m1, m2, m3);
    R* m = side_mom + ((static_cast<long long>(par) * 2 + col) * plane_yz + zy) * 4;
    m[0] = m0;
    m[1] = m1;
    m[2] = m2;
    m[3] = m3;
}

// x shift (-1, 0, +1) of the slot that population d is loaded from / stored to,
// relative to the cell: the pull source x - e_d (ping-pong), or the pair anchor (in
// place: loads of even d and stores of odd d sit at the anchor x + e_{pair}).  The
// own-previous slot (prev_off) of the in-place layout is a store slot.
template <int Q, class Lay>
__host__ __device__ constexpr int LoadDx(int d) {
    if constexpr (!Lay::kInPlace)
        return -Lat<Q>::ex[d];
    else
        return (d != 0 && (d & 1) == 0) ? -Lat<Q>::ex[d] : 0;
}

template <int Q, class Lay>
__host__ __device__ constexpr int StoreDx(int d) {
    if constexpr (!Lay::kInPlace)
        return 0;
    else
        return (d & 1) ? Lat<Q>::ex[d] : 0;
}

// Bulk pull of one cell, issued before its flag byte is known.  `src` addresses the
// workgroup's lower z layer; `ln` holds the cell's byte offsets within a layer, shared by
// every direction.  The offsets are those of an interior cell and stay inside the
// padded allocation for every cell of the lattice; values for cells that take another
// branch are discarded.
template <int Q, class S, class Lay, int LA>
__device__ __forceinline__ void SpecLoad(const StepArgs& a, BufRsrc src, const PopLane& ln, typename S::R (&fs)[Q]) {
    using T = typename S::T;
    StaticFor<0, Q>([&](auto I) {
        constexpr int d = I;
        fs[d] = S::Raw(BufLoad<T, LA>(src, LaneSel<LoadDx<Q, Lay>(d)>(ln), a.ld_off[d]));
    });
}

// Stores of a cell's post-collision populations (and, in place, the side data that
// other cells read next step).  `seam`: a cell on a periodic y/z edge, whose anchors
// wrap (only instantiated for periodic lattices).  SA: cache policy of the stores that
// stay on the block grid (kStoreAux streams them; 0 keeps them in the caches, for
// lattices whose populations fit in the last-level cache).
template <int Q, class S, class Lay, bool PER, int SA>
__device__ __forceinline__ void StoreCell(const StepArgs& a, const Geo& g, BufRsrc dst_layer, const PopLane& ln, int x,
                                          int y, int z, bool seam, const typename S::R (&fo)[Q], BRec<typename S::R>* rec) {
    using T = typename S::T;
    using R = typename S::R;
    T enc[Q];
    EncodeCell<Q, S>(fo, enc, CellSrHash<S>(a, x, y, z));
    auto store_all = [&]() {
        StaticFor<0, Q>([&](auto I) {
            constexpr int d = I;
            constexpr int dx = StoreDx<Q, Lay>(d);
            constexpr int aux = dx != 0 ? kShiftedStoreAux : SA;
            BufStore<T, aux>(dst_layer, LaneSel<dx>(ln), a.st_off[d], enc[d]);
        });
    };
    if constexpr (PER && Lay::kInPlace) {
        if (seam) {
            T* dst = static_cast<T*>(a.pop_dst);
            StaticFor<0, Q>([&](auto I) {
                constexpr int d = I;
                if constexpr (SA == 0)
                    dst[SlotIndex<Q, true>(g, d, a.parity, x, y, z)] = enc[d];
                else
                    PutPop(&dst[SlotIndex<Q, true>(g, d, a.parity, x, y, z)], enc[d]);
            });
        } else {
            store_all();
        }
    } else {
        store_all();
    }
    if constexpr (Lay::kInPlace) {
        if (rec != nullptr) {
            const uint32_t kinds[3] = {rec->kind[0], rec->kind[1], rec->kind[2]};
            StaticFor<1, Q>([&](auto I) {
                constexpr int d = I;
                if (KindOf(kinds, d) == kBouzidiHi) rec->side[d] = S::Raw(enc[d]);
            });
        }
        if (x == 1 || x == a.nx - 2)
            WriteMoments<Q, S>(static_cast<R*>(a.side_mom), enc, a.parity, x == 1 ? 0 : 1, a.ny * a.nz, z * a.ny + y);
    }
}

// Everything after the bulk pull for one cell: branch on the cell type, patch wall and
// solid directions, collide, store.
template <int Q, class S, class Lay, bool PER, int SA, int LA>
__device__ __forceinline__ void UpdateCell(const StepArgs& a, const Geo& g, BufRsrc src_layer, BufRsrc dst_layer,
                                           int x, int y, int z, const typename S::R (&fs)[Q], uint8_t fl) {
    using T = typename S::T;
    using R = typename S::R;
    constexpr bool IP = Lay::kInPlace;
    const int par = a.parity;
    const PopLane ln = MakePopLane<Q, T>(x, y, g.rs);
    const bool seam = PER && ((a.pery && (y == 0 || y == a.ny - 1)) || (a.perz && (z == 0 || z == a.nz - 1)));
    const int inletX = a.reverse ? a.nx - 1 : 0;
    const int outletX = a.reverse ? 0 : a.nx - 1;

    if (fl & kSolid) {
        // Solid cell: populations w_i (LbmReference.cpp:264-267).  In place, solid
        // cells are inactive; only the outlet-column moments they feed are written.
        R fo[Q];
        StaticFor<0, Q>([&](auto I) {
            constexpr int d = I;
            if constexpr (S::kOffset)
                fo[d] = R(0);
            else
                fo[d] = W<Q, R>(d);
        });
        if constexpr (!IP) {
            StaticFor<0, Q>([&](auto I) {
                constexpr int d = I;
                BufStore<T, SA>(dst_layer, ln.c, a.st_off[d], S::Enc(fo[d]));
            });
        } else {
            if (x == 1 || x == a.nx - 2) {
                T enc[Q];
                EncodeCell<Q, S>(fo, enc, CellSrHash<S>(a, x, y, z));
                WriteMoments<Q, S>(static_cast<R*>(a.side_mom), enc, par, x == 1 ? 0 : 1, a.ny * a.nz, z * a.ny + y);
            }
        }
        return;
    }

    R fo[Q];
    BRec<R>* rec = nullptr;
    if (x == inletX) {
        // Equilibrium inlet (LbmReference.cpp:269-273), evaluated once per launch.
        StaticFor<0, Q>([&](auto I) {
            constexpr int d = I;
            if constexpr (sizeof(R) == 4)
                fo[d] = a.inlet_f[d];
            else
                fo[d] = a.inlet_d[d];
        });
    } else if (x == outletX) {
        // Zero-gradient outlet from the neighbour's previous post-collision moments,
        // density relaxed toward 1 (LbmReference.cpp:275-327).
        const int nbx = a.reverse ? x + 1 : x - 1;
        R nr, nu, nv, nw;
        if constexpr (!IP) {
            R v[Q];
            const uint32_t nvo = nbx < x ? ln.m - static_cast<uint32_t>(sizeof(T)) : ln.p + static_cast<uint32_t>(sizeof(T));
            StaticFor<0, Q>([&](auto I) {
                constexpr int d = I;
                v[d] = S::Raw(BufLoad<T, LA>(src_layer, nvo, a.st_off[d]));
            });
            MomentSums<Q, R>(v, nr, nu, nv, nw);
        } else {
            const int col = nbx == 1 ? 0 : 1;
            const R* m = static_cast<const R*>(a.side_mom) +
                         ((static_cast<long long>(par ^ 1) * 2 + col) * (a.ny * a.nz) + z * a.ny + y) * 4;
            nr = m[0];
            nu = m[1];
            nv = m[2];
            nw = m[3];
        }
        if constexpr (S::kOffset) {
            // sums of offsets: rho_n - 1 and j_n in population units
            constexpr float invS = 1.0f / S::kScale;
            const float dn = nr * invS;
            const float rn = 1.0f + dn;
            float ux = nu * invS, uy = nv * invS, uz = nw * invS, drho = 0.0f;
            if (rn > 1e-6f) {
                const float r = 1.0f / rn;
                ux *= r;
                uy *= r;
                uz *= r;
                drho = 0.999f * dn;
            }
            EqOffset<Q, S>(drho, ux, uy, uz, fo);
        } else {
            if (nr > K<R>(1e-6f, 1e-6)) {
                nu /= nr;
                nv /= nr;
                nw /= nr;
            }
            const R rho = nr > K<R>(1e-6f, 1e-6) ? R(1) + K<R>(0.999f, 0.999) * (nr - R(1)) : R(1);
            FeqAll<Q, R>(rho, nu, nv, nw, fo);
        }
    } else {
        // -------------------------------------------------------------- gather
        // f starts as the bulk pull (issued before the flag byte is known); cells on a wall
        // or next to a solid then correct the affected directions.
        R f[Q];
        StaticFor<0, Q>([&](auto I) {
            constexpr int d = I;
            f[d] = fs[d];
        });
        const bool near = (fl & kNearSolid) != 0;
        const bool edge = (y == 0) | (y == a.ny - 1) | (z == 0) | (z == a.nz - 1);
        if (edge | near) {
            // The cell's own previous outgoing populations (every value a wall or a solid link
            // may read) and a seam cell's wrapped pulls are loaded in one batch before any of
            // them is used.  Loaded inside the per-direction branches below, each cost the wave
            // a memory round trip of its own (a wall cell five in a row): FP16 ping-pong on the
            // 512x256x256 wing 68.1 -> 69.1 GLUPS, the periodic 128x64x64 sphere 19.3 -> 18.8 us
            // per step (FP32).  The link q values stay in the branches: loading them in the batch too
            // holds 18 more registers (FP16 in place 62 -> 57 GLUPS).
            R own[Q];
            if constexpr (!IP) {
                StaticFor<1, Q>([&](auto I) {
                    constexpr int j = I;
                    own[j] = S::Raw(BufLoad<T, LA>(src_layer, ln.c, a.st_off[j]));
                });
            } else {
                auto own_slots = [&]() {
                    StaticFor<1, Q>([&](auto I) {
                        constexpr int j = I;
                        own[j] = S::Raw(BufLoad<T, LA>(dst_layer, LaneSel<(j & 1) ? Lat<Q>::ex[j] : 0>(ln), a.prev_off[j]));
                    });
                };
                if constexpr (PER) {
                    if (seam) {
                        const T* dst = static_cast<const T*>(a.pop_dst);
                        StaticFor<1, Q>([&](auto I) {
                            constexpr int j = I;
                            own[j] = S::Raw(dst[SlotIndex<Q, true>(g, j, par ^ 1, x, y, z)]);
                        });
                    } else {
                        own_slots();
                    }
                } else {
                    own_slots();
                }
            }
            // A seam cell reloads every pull through the wrap rules: the wrapped directions
            // read the cell across the seam, the others the slot the bulk pull read (the
            // same value: nothing writes it during the step before this cell does).
            if constexpr (PER) {
                if (seam) {
                    StaticFor<1, Q>([&](auto I) {
                        constexpr int d = I;
                        if constexpr (!IP) {
                            int sx, sy2, sz2;
                            g.NbrXYZ(x, y, z, -Lat<Q>::ex[d], -Lat<Q>::ey[d], -Lat<Q>::ez[d], sx, sy2, sz2);
                            f[d] = S::Raw(static_cast<const T*>(a.pop_src)[g.Pop(d, sx, sy2, sz2)]);
                        } else {
                            f[d] = S::Raw(static_cast<const T*>(a.pop_dst)[SlotIndex<Q, false>(g, d, par, x, y, z)]);
                        }
                    });
                }
            }
            uint32_t kinds[3] = {0u, 0u, 0u};
            if (near) {
                rec = static_cast<BRec<R>*>(a.recs) + a.bidx[g.Idx(x, y, z)];
                kinds[0] = rec->kind[0];
                kinds[1] = rec->kind[1];
                kinds[2] = rec->kind[2];
            }
            // Walls and solid sources, from the cell's own previous outgoing populations
            // (LbmReference.cpp:100-122, 336-381), in the reference's direction order.
            StaticFor<1, Q>([&](auto I) {
                constexpr int d = I;
                const int sy_ = y - Lat<Q>::ey[d], sz_ = z - Lat<Q>::ez[d];
                const bool outY = !a.pery && (sy_ < 0 || sy_ >= a.ny);
                const bool outZ = !a.perz && (sz_ < 0 || sz_ >= a.nz);
                if (outY || outZ) {
                    const bool ok = (!outY || a.fsy) && (!outZ || a.fsz);
                    if (ok) {
                        if (outY && outZ)
                            f[d] = own[MirrorZ<Q>(MirrorY<Q>(d))];
                        else if (outY)
                            f[d] = own[MirrorY<Q>(d)];
                        else
                            f[d] = own[MirrorZ<Q>(d)];
                    } else {
                        f[d] = own[Opp(d)];
                    }
                    return;
                }
                const unsigned kd = KindOf(kinds, d);
                if (kd == kPull) return;
                const R plain = own[Opp(d)];
                if (kd == kPlainBB) {
                    f[d] = plain;
                } else if (kd == kBouzidiLo) {
                    const R q = rec->q[d];
                    f[d] = R(2) * q * plain + (R(1) - R(2) * q) * f[Opp(d)];
                } else {
                    const R q = rec->q[d];
                    const R inv2q = R(1) / (R(2) * q);
                    R fi;
                    if constexpr (!IP)
                        fi = own[d];
                    else
                        fi = rec->side[d];
                    f[d] = inv2q * plain + (R(1) - inv2q) * fi;
                }
            });
        }
        // ------------------------------------------------------------- collide
        const R tau = sizeof(R) == 4 ? static_cast<R>(a.tau) : static_cast<R>(a.tau_d);
        const R cs = sizeof(R) == 4 ? static_cast<R>(a.cs) : static_cast<R>(a.cs_d);
        bool bad;
        if constexpr (S::kOffset)
            bad = CollideOffset<Q, S>(f, fo, tau, cs);
        else
            bad = Collide<Q, R>(f, fo, tau, cs);
        // At most one store per wave, from its first diverged lane, and none once the
        // sticky flag is set: a lattice that has blown up everywhere would otherwise
        // serialise one same-address store per wave (half a million per step at 33.5M
        // cells) in one L2 channel.
        const uint64_t bad_lanes = __builtin_amdgcn_ballot_w64(bad);
        if (bad_lanes != 0 && __lane_id() == static_cast<unsigned>(__builtin_ctzll(bad_lanes)) &&
            __hip_atomic_load(a.diverged, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
            __hip_atomic_store(a.diverged, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // One store sequence for every non-solid cell: inlet and outlet lanes write together
    // with the fluid lanes of their wave, so no cache line is written in two pieces.
    StoreCell<Q, S, Lay, PER, SA>(a, g, dst_layer, ln, x, y, z, seam, fo, rec);
}

#ifndef FC_LBM_PIN_PULLS
#define FC_LBM_PIN_PULLS 1
#endif

// One workgroup = a kTileX by kTileY*CPT tile of one z layer; each thread updates CPT
// cells kTileY rows apart.  The bulk pulls of all CPT cells are issued first.
// The per-direction scalar offsets hold ~100 SGPRs, which caps the step at 7 waves per
// SIMD.  Forcing 8 (FC_LBM_WAVES_PER_EU=8: 78 SGPRs, the rest spilled to VGPR lanes)
// measured slower on the 512x256x256 wing (FP32 ping-pong 37.06 -> 36.9 GLUPS, FP16
// 70.1 -> 68.2), so no limit is set by default.  FC_LBM_WAVES_PER_EU=7 (the periodic
// FP32 ping-pong build 73 -> 70 VGPRs, 6 -> 7 waves) changed nothing beyond run-to-run
// spread: periodic 128x64x64 sphere FP32 18.8-19.1 -> 18.6-18.7 us, FP16 15.6 -> 15.9 us.
#ifndef FC_LBM_WAVES_PER_EU
#define FC_LBM_WAVES_PER_EU 0
#endif
#if FC_LBM_WAVES_PER_EU > 0
#define FC_LBM_STEP_ATTR __attribute__((amdgpu_waves_per_eu(FC_LBM_WAVES_PER_EU)))
#else
#define FC_LBM_STEP_ATTR
#endif
// Load cache policy of the step.  Non-temporal loads measured faster for the streamed
// in-place step and slower for ping-pong (512x256x256 wing, identical fields: FP32 in
// place 32.1 -> 32.6 GLUPS, stochastically rounded FP16 in place 62.0 -> 66.8; ping-pong
// FP32 37.2 -> 35.5, FP16 68.4 -> 66.8, and 37.2 -> 36.95 with only the pulls without
// an x shift non-temporal; periodic 128x64x64 sphere held in the cache 18.8 -> 22.5 us).
template <class Lay, bool RESIDENT>
__host__ __device__ constexpr int StepLoadAux() {
    return Lay::kInPlace && !RESIDENT ? FC_LBM_INPLACE_LOAD_AUX : kLoadAux;
}
// RESIDENT: the populations fit in the last-level cache, so the stores stay cached for
// the next step's pulls instead of streaming to HBM.
template <int Q, class S, class Lay, int CPT, bool PER, bool RESIDENT = false>
__global__ __launch_bounds__(StepThreads(sizeof(typename S::T))) FC_LBM_STEP_ATTR void KStep(const StepArgs a) {
    constexpr int SA = RESIDENT ? 0 : kStoreAux;
    constexpr int LA = StepLoadAux<Lay, RESIDENT>();
    using R = typename S::R;
    using T = typename S::T;
    S::StepMode();
#ifndef FC_LBM_XCD_REMAP
#define FC_LBM_XCD_REMAP 1
#endif
#ifndef FC_LBM_XCD_CHUNK
#define FC_LBM_XCD_CHUNK 4
#endif
    const unsigned tile = FC_LBM_XCD_REMAP == 1   ? XcdContiguous(blockIdx.x, gridDim.x)
                          : FC_LBM_XCD_REMAP == 2 ? XcdChunked<FC_LBM_XCD_CHUNK>(blockIdx.x, gridDim.x)
                                                  : blockIdx.x;
    constexpr int kTileX = TileX(sizeof(T)), kTileY = TileY(sizeof(T));
    const unsigned tiles_x = static_cast<unsigned>((a.nx + kTileX - 1) / kTileX);
    const unsigned tiles_y = static_cast<unsigned>((a.ny + kTileY * CPT - 1) / (kTileY * CPT));
    const unsigned tx = tile % tiles_x;
    const unsigned rest = tile / tiles_x;
    const unsigned ty = rest % tiles_y;
    const int x = static_cast<int>(tx * kTileX + (threadIdx.x & (kTileX - 1)));
    const int y0 = static_cast<int>(ty * kTileY * CPT + (threadIdx.x / kTileX));
    const int z = a.z0 + static_cast<int>(rest / tiles_y);
    if (x >= a.nx) return;
    const Geo g{a.nx, a.ny, a.nz, a.pery, a.perz, a.nx, a.nx * (a.ny + 2), Q, PopRowStride(a.nx, Q)};
    // workgroup bases: padded layer z (the lower neighbour of the cells' own layer z+1)
    const long long zb = a.base_bytes + static_cast<long long>(z) * a.layer_bytes;
    const BufRsrc src_layer = LayerRsrc(static_cast<const char*>(Lay::kInPlace ? a.pop_dst : a.pop_src) + zb);
    const BufRsrc dst_layer = LayerRsrc(static_cast<char*>(a.pop_dst) + zb);
    R fs[CPT][Q];
    uint8_t fl[CPT];
    // Rows past the lattice edge (a ragged last tile) load row ny-1 and are not updated.
    const uint64_t flag_layer = reinterpret_cast<uint64_t>(a.flags) + static_cast<uint64_t>(z + 1) * g.sz;
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int y = min(y0 + k * kTileY, a.ny - 1);
        SpecLoad<Q, S, Lay, LA>(a, src_layer, MakePopLane<Q, T>(x, y, g.rs), fs[k]);
    }
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int y = min(y0 + k * kTileY, a.ny - 1);
        fl[k] = LoadU<uint8_t>(flag_layer, static_cast<uint32_t>((y + 1) * g.sy + x));
    }
#if FC_LBM_PIN_PULLS
    // Pin the bulk pulls ahead of the branches on the flag byte.  Their values are used
    // only inside the fluid branch, so the compiler would otherwise sink each pull into
    // it, behind the flag's round trip (and, with two cells per thread, behind the other
    // cell's update).  An empty asm that takes every pulled value keeps all loads of the
    // thread in flight together: one memory round trip per wave instead of two or more.
#pragma unroll
    for (int k = 0; k < CPT; ++k)
#pragma unroll
        for (int d = 0; d < Q; ++d) asm volatile("" ::"v"(fs[k][d]));
#endif
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int y = y0 + k * kTileY;
        if (y < a.ny) UpdateCell<Q, S, Lay, PER, SA, LA>(a, g, src_layer, dst_layer, x, y, z, fs[k], fl[k]);
    }
}

}  // namespace fc
