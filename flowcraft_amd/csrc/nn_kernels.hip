// libfc_nn.so — hand-written BF16 MFMA kernels for the warm-start surrogate
// (a 3D convolutional encoder-decoder), forward and backward, plus Adam.
//
// Tensors are channels-last bf16: [voxel][channel] with voxel = ((n*D + z)*H + y)*W + x.
// Convolution weights are [cout][27][cin] bf16 (k = tap*cin + ci contiguous per output
// channel), taps ordered tz-major: tap = (kz*3 + ky)*3 + kx; padding 1.
//
// Convolution = implicit GEMM on v_mfma_f32_16x16x32_bf16 (gfx950):
//   Out[v][n] = sum_k A[v][k] * B[k][n],  A = gathered input, B[k][n] = Wk[n][k].
// Lane layout of the instruction (checked on the device): lane l holds A row l%16,
// k = 8*(l/16)..+7; B column l%16, same k; result rows 4*(l/16)+r, column l%16.
// Every 8-run of k stays inside one tap because channel counts are multiples of 8, so
// a lane's A fragment is one 16-byte load (or zeros outside the volume).
//
// Gather modes:
//   kFwd   : input voxel i = o*s + t - 1 per axis (stride s), zero outside
//   kUpcat : stride 1; channels [0,c1) from a half-resolution tensor at i/2 (nearest
//            upsample), channels [c1,cin) from a full-resolution tensor at i
//   kAdj   : adjoint of a stride-s kFwd conv (data gradient): o = (i + 1 - t)/s when
//            divisible and inside the forward output; weights pre-transposed
//            to [cin][27][cout]
// All reductions are in a fixed order (split-K partials summed in split order), so
// results are deterministic.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>
#include <type_traits>

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16;

namespace {

constexpr int kFwd = 0, kUpcat = 1, kAdj = 2;
constexpr float kLeak = 0.05f;

__device__ __forceinline__ float Bf2F(u16 h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }
__device__ __forceinline__ u16 F2Bf(float f) {
    uint32_t b = __float_as_uint(f);
    if ((b & 0x7FFFFFFFu) > 0x7F800000u) return static_cast<u16>((b >> 16) | 0x40u);
    b += 0x7FFFu + ((b >> 16) & 1u);
    return static_cast<u16>(b >> 16);
}

// Two floats to two bf16 (round to nearest even) in one word, low half first: one
// v_cvt_pk_bf16_f32 on gfx950 (F2Bf spells the same rounding out for finite values).
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t PackBf2(float lo, float hi) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2_t{lo, hi}, bf16x2_t));
}

struct ConvArgs {
    const u16* in;    // kFwd/kAdj: input; kUpcat: full-resolution part
    const u16* in2;   // kUpcat: half-resolution part
    const u16* w;     // [n_out][27][k_in]
    const float* bias;  // [n_out] or null
    u16* out;         // [V_out][n_out] bf16, or
    float* outf;      // [V_out][n_out] fp32 (when non-null)
    int nb;
    int di, hi, wi;   // gather-source spatial dims (kUpcat: the full-resolution dims)
    int dd, hd, wd;   // destination spatial dims
    int kin;          // channels gathered per tap
    int c1;           // kUpcat: channels taken from in2
    int nout;         // output channels
    int stride;
    int act;          // 1 = leaky ReLU on the output
};

// Gathers the 8 values A[v][k0..k0+7] for destination voxel (n, z, y, x).
template <int MODE>
__device__ __forceinline__ bf16x8 GatherA(const ConvArgs& a, int n, int z, int y, int x, int k0, bool vok) {
    bf16x8 r;
    uint4 raw = make_uint4(0u, 0u, 0u, 0u);
    const int K = 27 * a.kin;
    if (vok && k0 < K) {
        const int t = k0 / a.kin;
        const int ci = k0 - t * a.kin;
        const int kz = t / 9, ky = (t / 3) % 3, kx = t % 3;
        int iz, iy, ix;
        bool ok;
        if (MODE == kAdj) {
            const int s = a.stride;
            const int pz = z + 1 - kz, py = y + 1 - ky, px = x + 1 - kx;
            ok = pz >= 0 && py >= 0 && px >= 0 && (pz % s) == 0 && (py % s) == 0 && (px % s) == 0;
            iz = pz / s;
            iy = py / s;
            ix = px / s;
            ok = ok && iz < a.di && iy < a.hi && ix < a.wi;
        } else {
            const int s = MODE == kUpcat ? 1 : a.stride;
            iz = z * s + kz - 1;
            iy = y * s + ky - 1;
            ix = x * s + kx - 1;
            ok = iz >= 0 && iy >= 0 && ix >= 0 && iz < a.di && iy < a.hi && ix < a.wi;
        }
        if (ok) {
            if (MODE == kUpcat && ci < a.c1) {
                const int h2 = a.hi >> 1, w2 = a.wi >> 1, d2 = a.di >> 1;
                const long long vi = ((static_cast<long long>(n) * d2 + (iz >> 1)) * h2 + (iy >> 1)) * w2 + (ix >> 1);
                raw = *reinterpret_cast<const uint4*>(a.in2 + vi * a.c1 + ci);
            } else {
                const int cs = MODE == kUpcat ? a.kin - a.c1 : a.kin;
                const int cc = MODE == kUpcat ? ci - a.c1 : ci;
                const long long vi = ((static_cast<long long>(n) * a.di + iz) * a.hi + iy) * a.wi + ix;
                raw = *reinterpret_cast<const uint4*>(a.in + vi * cs + cc);
            }
        }
    }
    __builtin_memcpy(&r, &raw, 16);
    return r;
}

// Bias, activation and store of output channels c0 .. c0+3 of voxel v (those below a.nout),
// the four results a lane holds when the weights are the MFMA's first operand (rows = output
// channels): one 8-byte (bf16) or 16-byte (fp32) store when a.nout is a multiple of 4.
__device__ __forceinline__ void StoreC4(const ConvArgs& a, long long v, int c0, const f32x4& acc) {
    if (c0 >= a.nout) return;
    float val[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        val[q] = acc[q] + (a.bias && c0 + q < a.nout ? a.bias[c0 + q] : 0.f);
        if (a.act == 1) val[q] = val[q] > 0.f ? val[q] : kLeak * val[q];
    }
    const long long o = v * a.nout + c0;
    if ((a.nout & 3) == 0) {  // then c0 + 3 < nout, and o is a multiple of 4
        if (a.outf) {
            *reinterpret_cast<float4*>(a.outf + o) = make_float4(val[0], val[1], val[2], val[3]);
        } else {
            *reinterpret_cast<uint2*>(a.out + o) = make_uint2(PackBf2(val[0], val[1]), PackBf2(val[2], val[3]));
        }
    } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (c0 + q >= a.nout) continue;
            if (a.outf)
                a.outf[o + q] = val[q];
            else
                a.out[o + q] = F2Bf(val[q]);
        }
    }
}

// One wave computes (16*TM) destination voxels x (16*TN) output channels; a block is
// 4 waves along the voxel axis.
template <int MODE, int TM, int TN>
__global__ __launch_bounds__(256) void KConv(const ConvArgs a) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const long long V = static_cast<long long>(a.nb) * a.dd * a.hd * a.wd;
    const long long vbase = (static_cast<long long>(blockIdx.x) * 4 + wave) * (16 * TM);
    const int nbase = blockIdx.y * (16 * TN);
    const int K = 27 * a.kin;
    const int nchunks = (K + 31) / 32;
    const int row = lane & 15, kq = (lane >> 4) * 8;

    int vn[TM], vz[TM], vy[TM], vx[TM];
    bool vok[TM];
#pragma unroll
    for (int m = 0; m < TM; ++m) {
        long long v = vbase + m * 16 + row;
        vok[m] = v < V;
        if (!vok[m]) v = 0;
        vx[m] = static_cast<int>(v % a.wd);
        long long r = v / a.wd;
        vy[m] = static_cast<int>(r % a.hd);
        r /= a.hd;
        vz[m] = static_cast<int>(r % a.dd);
        vn[m] = static_cast<int>(r / a.dd);
    }
    f32x4 acc[TM][TN];
#pragma unroll
    for (int m = 0; m < TM; ++m)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[m][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    // K chunk kc's operands; those of chunk kc+1 are loaded before chunk kc's MFMAs, so the
    // gathers' latency overlaps one chunk of matrix work
    auto load = [&](int kc, bf16x8 (&bfr)[TN], bf16x8 (&afr)[TM]) {
        const int k0 = kc * 32 + kq;
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            uint4 raw = make_uint4(0u, 0u, 0u, 0u);
            const int n = nbase + j * 16 + row;
            if (k0 < K && n < a.nout) raw = *reinterpret_cast<const uint4*>(a.w + static_cast<long long>(n) * K + k0);
            __builtin_memcpy(&bfr[j], &raw, 16);
        }
#pragma unroll
        for (int m = 0; m < TM; ++m) afr[m] = GatherA<MODE>(a, vn[m], vz[m], vy[m], vx[m], k0, vok[m]);
    };
    bf16x8 bcur[TN], acur[TM];
    load(0, bcur, acur);
    for (int kc = 0; kc < nchunks; ++kc) {
        bf16x8 bnext[TN], anext[TM];
        if (kc + 1 < nchunks) load(kc + 1, bnext, anext);
#pragma unroll
        for (int m = 0; m < TM; ++m)
#pragma unroll
            for (int j = 0; j < TN; ++j) acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bcur[j], acur[m], acc[m][j], 0, 0, 0);
        if (kc + 1 < nchunks) {
#pragma unroll
            for (int j = 0; j < TN; ++j) bcur[j] = bnext[j];
#pragma unroll
            for (int m = 0; m < TM; ++m) acur[m] = anext[m];
        }
    }
    // epilogue: the weights are the MFMA's first operand, so lane (row, g) holds output
    // channels 16j + 4g .. +3 of voxel vbase + 16m + row
#pragma unroll
    for (int m = 0; m < TM; ++m) {
        const long long v = vbase + m * 16 + row;
        if (v >= V) continue;
#pragma unroll
        for (int j = 0; j < TN; ++j) StoreC4(a, v, nbase + j * 16 + (lane >> 4) * 4, acc[m][j]);
    }
}

// Adjoint (data gradient) of a stride-2 convolution, split by output parity.  Along one
// axis an even output coordinate c receives only tap 1 (from input c/2) and an odd one taps
// 0 and 2 (from (c+1)/2 and (c-1)/2), so a voxel has 1 to 8 valid taps of 27 (3.4 on
// average); KConv's gather visits all 27 and zeroes the rest.  Wave w of a workgroup takes
// parity class w = (pz, py, px); a class's voxels are enumerated at half resolution, so each
// wave's K loop visits only its class's valid taps, 32 gathered channels per step (kin % 32
// == 0).  The eight waves cover the same 16*TM half-resolution positions, so on even extents
// a workgroup writes whole 2 x 2 x 2 blocks of output voxels and no output line is left
// partly written for another workgroup.  The weights are the MFMA's first operand: a lane
// holds four consecutive output channels of one voxel and stores them at once (StoreC4).
// blockIdx.z selects 16*TN output channels.
template <int TM, int TN>
__global__ __launch_bounds__(512) void KConvAdj2(const ConvArgs a) {
    const int lane = threadIdx.x & 63;
    const int cls = threadIdx.x >> 6;
    const int px = cls & 1, py = (cls >> 1) & 1, pz = cls >> 2;
    const int wx = (a.wd - px + 1) / 2, wy = (a.hd - py + 1) / 2, wz = (a.dd - pz + 1) / 2;
    const long long Vc = static_cast<long long>(a.nb) * wz * wy * wx;
    const long long vbase = static_cast<long long>(blockIdx.x) * (16 * TM);
    if (vbase >= Vc) return;  // wave-uniform: classes with odd extents hold fewer voxels
    const int nbase = blockIdx.z * (16 * TN);
    const int K = 27 * a.kin;
    const int row = lane & 15, kq = (lane >> 4) * 8;
    // the output voxel of class index v
    auto coords = [&](long long v, int& n, int& z, int& y, int& x) {
        x = 2 * static_cast<int>(v % wx) + px;
        long long r = v / wx;
        y = 2 * static_cast<int>(r % wy) + py;
        r /= wy;
        z = 2 * static_cast<int>(r % wz) + pz;
        n = static_cast<int>(r / wz);
    };
    // An output voxel (z, y, x) = (2zc + pz, 2yc + py, 2xc + px) reads, through tap k along an
    // axis, the input at (c + 1 - k) / 2: c/2 = zc for an even c (k = 1), and zc + 1 (k = 0)
    // or zc (k = 2) for an odd one.  So a tap's input voxel is (zc + dz, yc + dy, xc + dx)
    // with d in {0, 1}: the lane's coarse index plus an offset the whole wave shares.
    int vn[TM], vz[TM], vy[TM], vx[TM];
    long long vin[TM];
    bool vok[TM];
#pragma unroll
    for (int m = 0; m < TM; ++m) {
        const long long v = vbase + m * 16 + row;
        vok[m] = v < Vc;
        coords(vok[m] ? v : 0, vn[m], vz[m], vy[m], vx[m]);
        vin[m] = ((static_cast<long long>(vn[m]) * a.di + (vz[m] >> 1)) * a.hi + (vy[m] >> 1)) * a.wi + (vx[m] >> 1);
    }
    f32x4 acc[TM][TN];
#pragma unroll
    for (int m = 0; m < TM; ++m)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[m][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int nkx = px ? 2 : 1, nky = py ? 2 : 1, nkz = pz ? 2 : 1;
    const int ntaps = nkx * nky * nkz, kblocks = a.kin / 32;
    for (int ti = 0; ti < ntaps; ++ti) {
        const int jx = ti % nkx, jy = (ti / nkx) % nky, jz = ti / (nkx * nky);
        const int kx = px ? 2 * jx : 1, ky = py ? 2 * jy : 1, kz = pz ? 2 * jz : 1;
        const int t = (kz * 3 + ky) * 3 + kx;
        const int dx = kx == 0, dy = ky == 0, dz = kz == 0;
        const long long toff = (static_cast<long long>(dz) * a.hi + dy) * a.wi + dx;
        bool tok[TM];
#pragma unroll
        for (int m = 0; m < TM; ++m)
            tok[m] = vok[m] && (vz[m] >> 1) + dz < a.di && (vy[m] >> 1) + dy < a.hi && (vx[m] >> 1) + dx < a.wi;
        for (int kc = 0; kc < kblocks; ++kc) {
            const int k0 = t * a.kin + kc * 32 + kq;
            bf16x8 wfr[TN];
#pragma unroll
            for (int j = 0; j < TN; ++j) {
                uint4 raw = make_uint4(0u, 0u, 0u, 0u);
                const int n = nbase + j * 16 + row;
                if (n < a.nout) raw = *reinterpret_cast<const uint4*>(a.w + static_cast<long long>(n) * K + k0);
                __builtin_memcpy(&wfr[j], &raw, 16);
            }
#pragma unroll
            for (int m = 0; m < TM; ++m) {
                uint4 raw = make_uint4(0u, 0u, 0u, 0u);
                if (tok[m]) raw = *reinterpret_cast<const uint4*>(a.in + (vin[m] + toff) * a.kin + kc * 32 + kq);
                bf16x8 gfr;
                __builtin_memcpy(&gfr, &raw, 16);
#pragma unroll
                for (int j = 0; j < TN; ++j)
                    acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wfr[j], gfr, acc[m][j], 0, 0, 0);
            }
        }
    }
    // epilogue: lane holds output channels 4*(lane/16)+r of class voxel vbase + 16m + lane%16
#pragma unroll
    for (int m = 0; m < TM; ++m) {
        if (!vok[m]) continue;
        const long long vo = ((static_cast<long long>(vn[m]) * a.dd + vz[m]) * a.hd + vy[m]) * a.wd + vx[m];
#pragma unroll
        for (int j = 0; j < TN; ++j) StoreC4(a, vo, nbase + j * 16 + 4 * (lane >> 4), acc[m][j]);
    }
}

// Weight gradient of a kFwd/kUpcat conv: dW[co][k] = sum_v dZ[v][co] * A[v][k], split over
// voxel ranges (blockIdx.z) into fp32 partials part[split][co][K].  32 voxels per step,
// staged transposed in LDS so fragments are single 16-byte reads.
struct WgradArgs {
    ConvArgs g;       // gather description of the forward conv (MODE, dims, kin, c1, stride)
    const u16* dz;    // [V][cout]
    int cout;
    long long vper;   // voxels per split
    float* part;      // [splits][cout][27*kin]
};

// Block tile TCO output channels x TK = 4096 / TCO k-columns (16 x 256, 32 x 128 or
// 64 x 64, matched to the layer's channel count so narrow layers waste no MFMA work), 32
// voxels per step; each of the 4 waves owns TCO/16 x (TK/4)/16 MFMA tiles (4 in all).
template <int MODE, int TCO>
__global__ __launch_bounds__(256) void KWgrad(const WgradArgs w) {
    constexpr int VS = 32, LD = 40;  // voxels per step, padded LDS row (bf16 elements)
    constexpr int TK = 4096 / TCO;
    constexpr int WI = TCO == 16 ? 1 : 2;  // MFMA tiles per wave along co
    constexpr int WJ = 4 / WI;             // ... along k
    static_assert(TCO == 16 || TCO == 32 || TCO == 64, "co tile");
    __shared__ __attribute__((aligned(16))) u16 sA[TCO * LD];  // [co][v]
    __shared__ __attribute__((aligned(16))) u16 sB[TK * LD];   // [k][v]
    const ConvArgs& a = w.g;
    const int K = 27 * a.kin;
    const int co0 = blockIdx.y * TCO, k0b = blockIdx.x * TK;
    const long long V = static_cast<long long>(a.nb) * a.dd * a.hd * a.wd;
    const long long vbeg = static_cast<long long>(blockIdx.z) * w.vper;
    const long long vend = vbeg + w.vper < V ? vbeg + w.vper : V;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // wave tile: rows [wm, wm + 16*WI) of co, columns [wn, wn + 16*WJ) of k
    const int wm = TCO == 64 ? (wave & 1) * 32 : 0;
    const int wn = TCO == 64 ? (wave >> 1) * 32 : wave * 16 * WJ;
    f32x4 acc[WI][WJ];
    for (int i = 0; i < WI; ++i)
        for (int j = 0; j < WJ; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (long long v0 = vbeg; v0 < vend; v0 += VS) {
        // stage dZ[v0..v0+32][co0..co0+TCO] transposed
        for (int e = threadIdx.x; e < VS * TCO; e += 256) {
            const int vv = e / TCO, cc = e % TCO;
            const long long v = v0 + vv;
            u16 val = 0;
            if (v < vend && co0 + cc < w.cout) val = w.dz[v * w.cout + co0 + cc];
            sA[cc * LD + vv] = val;
        }
        // stage A[v][k0b..k0b+TK] (gathered input) transposed, voxel-fastest: lanes
        // 32 apart in e share a k-run, consecutive lanes write consecutive voxels of one
        // LDS row (no bank conflicts), and each thread keeps one voxel for all its runs.
        {
            const int vv = threadIdx.x % VS;
            const long long v = v0 + vv;
            const bool vok = v < vend;
            long long vc = vok ? v : 0;
            const int x = static_cast<int>(vc % a.wd);
            long long r = vc / a.wd;
            const int y = static_cast<int>(r % a.hd);
            r /= a.hd;
            const int z = static_cast<int>(r % a.dd);
            const int n = static_cast<int>(r / a.dd);
            for (int kk = (threadIdx.x / VS) * 8; kk < TK; kk += (256 / VS) * 8) {
                const bf16x8 g8 = GatherA<MODE>(a, n, z, y, x, k0b + kk, vok);
                u16 tmp[8];
                __builtin_memcpy(tmp, &g8, 16);
                for (int q = 0; q < 8; ++q) sB[(kk + q) * LD + vv] = tmp[q];
            }
        }
        __syncthreads();
        const int row = lane & 15, kq = (lane >> 4) * 8;
        bf16x8 af[WI], bf[WJ];
        for (int i = 0; i < WI; ++i) __builtin_memcpy(&af[i], &sA[(wm + i * 16 + row) * LD + kq], 16);
        for (int j = 0; j < WJ; ++j) __builtin_memcpy(&bf[j], &sB[(wn + j * 16 + row) * LD + kq], 16);
        for (int i = 0; i < WI; ++i)
            for (int j = 0; j < WJ; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bf[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }
    float* part = w.part + static_cast<long long>(blockIdx.z) * w.cout * K;
    for (int i = 0; i < WI; ++i)
        for (int j = 0; j < WJ; ++j) {
            const int k = k0b + wn + j * 16 + (lane & 15);
            for (int r = 0; r < 4; ++r) {
                const int co = co0 + wm + i * 16 + (lane >> 4) * 4 + r;
                if (co < w.cout && k < K) part[static_cast<long long>(co) * K + k] = acc[i][j][r];
            }
        }
}

// out[i] = sum_s part[s][i], optionally += into out.  A block owns 16 outputs; its 16
// thread groups each sum every 16th split, then the 16 group sums are added in group
// order — a fixed order, so results are reproducible run to run.
constexpr int kRedOut = 16, kRedGroups = 16;
__global__ __launch_bounds__(256) void KReduceSplits(const float* __restrict__ part, int splits, long long n,
                                                     float* __restrict__ out, int accumulate) {
    const int io = threadIdx.x % kRedOut, g = threadIdx.x / kRedOut;
    const long long i = static_cast<long long>(blockIdx.x) * kRedOut + io;
    float s = 0.f;
    if (i < n)
        for (int k = g; k < splits; k += kRedGroups) s += part[static_cast<long long>(k) * n + i];
    __shared__ float red[kRedGroups][kRedOut];
    red[g][io] = s;
    __syncthreads();
    if (g == 0 && i < n) {
        float t = accumulate ? out[i] : 0.f;
        for (int k = 0; k < kRedGroups; ++k) t += red[k][io];
        out[i] = t;
    }
}

// Per-channel sums over voxels of a [V][C] bf16 tensor: part[block][c] (fixed grid).  A
// thread reads 8 consecutive channels of a voxel with one 16-byte load.
constexpr int kColBlocks = 1024;  // partial sums per channel (fc_nn_col_blocks)
__global__ __launch_bounds__(256) void KColSumPartial(const u16* __restrict__ x, long long V, int C,
                                                      float* __restrict__ part) {
    const int octs = C / 8;                  // 16-byte groups per voxel
    const int lanes = 256 / octs;            // voxels in flight per block
    const int o = threadIdx.x % octs, sub = threadIdx.x / octs;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (long long v = static_cast<long long>(blockIdx.x) * lanes + sub; v < V;
         v += static_cast<long long>(kColBlocks) * lanes) {
        const uint4 raw = *reinterpret_cast<const uint4*>(x + v * C + o * 8);
        const uint32_t w4[4] = {raw.x, raw.y, raw.z, raw.w};
        for (int q = 0; q < 4; ++q) {
            s[2 * q] += __uint_as_float(w4[q] << 16);
            s[2 * q + 1] += __uint_as_float(w4[q] & 0xFFFF0000u);
        }
    }
    __shared__ float red[256 * 8];
    for (int q = 0; q < 8; ++q) red[threadIdx.x * 8 + q] = s[q];
    __syncthreads();
    if (threadIdx.x < C) {
        const int oc = threadIdx.x / 8, q = threadIdx.x % 8;
        float t = 0.f;
        for (int k = 0; k < lanes; ++k) t += red[(k * octs + oc) * 8 + q];
        part[blockIdx.x * C + threadIdx.x] = t;
    }
}

// dZ = dY * leaky'(Y) fused with the per-channel sums of dZ (the bias gradient): the thread
// mapping, grid and summation order of KColSumPartial, which reads back the stored bf16 dZ,
// so dZ and the partial sums are bit-identical to the two kernels run apart.  Each thread
// keeps the loads of kLcUnroll voxels in flight and accumulates them in voxel order.
constexpr int kLcUnroll = 8;
__global__ __launch_bounds__(256) void KLeakyColSum(const u16* __restrict__ y, const u16* __restrict__ dy,
                                                    u16* __restrict__ dz, long long V, int C,
                                                    float* __restrict__ part) {
    const int octs = C / 8;
    const int lanes = 256 / octs;
    const int o = threadIdx.x % octs, sub = threadIdx.x / octs;
    const long long stride = static_cast<long long>(kColBlocks) * lanes;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (long long v0 = static_cast<long long>(blockIdx.x) * lanes + sub; v0 < V; v0 += kLcUnroll * stride) {
        uint4 ry[kLcUnroll], rg[kLcUnroll];
#pragma unroll
        for (int u = 0; u < kLcUnroll; ++u) {
            const long long v = v0 + u * stride;
            if (v < V) {
                ry[u] = *reinterpret_cast<const uint4*>(y + v * C + o * 8);
                rg[u] = *reinterpret_cast<const uint4*>(dy + v * C + o * 8);
            }
        }
#pragma unroll
        for (int u = 0; u < kLcUnroll; ++u) {
            const long long v = v0 + u * stride;
            if (v >= V) break;
            u16 yv[8], gv[8], out[8];
            __builtin_memcpy(yv, &ry[u], 16);
            __builtin_memcpy(gv, &rg[u], 16);
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const float g = Bf2F(gv[q]);
                out[q] = F2Bf(Bf2F(yv[q]) > 0.f ? g : kLeak * g);
                s[q] += Bf2F(out[q]);
            }
            uint4 w;
            __builtin_memcpy(&w, out, 16);
            *reinterpret_cast<uint4*>(dz + v * C + o * 8) = w;
        }
    }
    __shared__ float red[256 * 8];
    for (int q = 0; q < 8; ++q) red[threadIdx.x * 8 + q] = s[q];
    __syncthreads();
    if (threadIdx.x < C) {
        const int oc = threadIdx.x / 8, q = threadIdx.x % 8;
        float t = 0.f;
        for (int k = 0; k < lanes; ++k) t += red[(k * octs + oc) * 8 + q];
        part[blockIdx.x * C + threadIdx.x] = t;
    }
}

// dZ = dY * leaky'(Y) (Y is the activated output; its sign is the pre-activation's).
__global__ void KLeakyBackward(const u16* __restrict__ y, const u16* __restrict__ dy, u16* __restrict__ dz,
                               long long n) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = Bf2F(dy[i]);
    dz[i] = F2Bf(Bf2F(y[i]) > 0.f ? g : kLeak * g);
}

// The same, 8 elements per thread (n % 8 == 0): one 16-byte load of each input and one store.
__global__ void KLeakyBackward8(const u16* __restrict__ y, const u16* __restrict__ dy, u16* __restrict__ dz,
                                long long n8) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n8) return;
    u16 yv[8], gv[8], out[8];
    const uint4 ry = reinterpret_cast<const uint4*>(y)[i], rg = reinterpret_cast<const uint4*>(dy)[i];
    __builtin_memcpy(yv, &ry, 16);
    __builtin_memcpy(gv, &rg, 16);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const float g = Bf2F(gv[q]);
        out[q] = F2Bf(Bf2F(yv[q]) > 0.f ? g : kLeak * g);
    }
    uint4 ro;
    __builtin_memcpy(&ro, out, 16);
    reinterpret_cast<uint4*>(dz)[i] = ro;
}

// Wt[ci][t][co] = W[co][t][ci]
// With flip, Wt[ci][26 - t][co] = W[co][t][ci]: the weights of the forward convolution of
// dZ that equals the stride-1 adjoint (tap (kz, ky, kx) mirrored to (2-kz, 2-ky, 2-kx)).
__global__ void KTransposeW(const u16* __restrict__ w, u16* __restrict__ wt, int cout, int cin, int flip) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long n = static_cast<long long>(cout) * 27 * cin;
    if (i >= n) return;
    const int ci = static_cast<int>(i % cin);
    const int t = static_cast<int>((i / cin) % 27);
    const int co = static_cast<int>(i / (27LL * cin));
    wt[(static_cast<long long>(ci) * 27 + (flip ? 26 - t : t)) * cout + co] = w[i];
}

// Split the data gradient of a kUpcat conv input ([V][c1+c2] at full resolution):
// channels >= c1 are added into dskip ([V][c2]); channels < c1 are summed over the 8
// children of each coarse voxel into dcoarse ([V/8][c1]).
__global__ void KUpcatSplit(const u16* __restrict__ dcat, int c1, int c2, int nb, int d, int h, int w,
                            u16* __restrict__ dskip, u16* __restrict__ dcoarse) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int C = c1 + c2;
    const int d2 = d >> 1, h2 = h >> 1, w2 = w >> 1;
    const long long Vc = static_cast<long long>(nb) * d2 * h2 * w2;
    const long long V = static_cast<long long>(nb) * d * h * w;
    if (i < V * c2) {
        const long long v = i / c2;
        const int c = static_cast<int>(i % c2);
        dskip[i] = F2Bf(Bf2F(dskip[i]) + Bf2F(dcat[v * C + c1 + c]));
    }
    if (i < Vc * c1) {
        const long long vc = i / c1;
        const int c = static_cast<int>(i % c1);
        const int x2 = static_cast<int>(vc % w2);
        long long r = vc / w2;
        const int y2 = static_cast<int>(r % h2);
        r /= h2;
        const int z2 = static_cast<int>(r % d2);
        const int n = static_cast<int>(r / d2);
        float s = 0.f;
        for (int q = 0; q < 8; ++q) {
            const int z = 2 * z2 + (q >> 2), y = 2 * y2 + ((q >> 1) & 1), x = 2 * x2 + (q & 1);
            const long long v = ((static_cast<long long>(n) * d + z) * h + y) * w + x;
            s += Bf2F(dcat[v * C + c]);
        }
        dcoarse[i] = F2Bf(s);
    }
}

// The same split, 8 channels per thread (c1 % 8 == 0 and c2 % 8 == 0), summing the 8
// children in the same order.
__global__ void KUpcatSplit8(const u16* __restrict__ dcat, int c1, int c2, int nb, int d, int h, int w,
                             u16* __restrict__ dskip, u16* __restrict__ dcoarse) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int C = c1 + c2, g1 = c1 / 8, g2 = c2 / 8;
    const int d2 = d >> 1, h2 = h >> 1, w2 = w >> 1;
    const long long Vc = static_cast<long long>(nb) * d2 * h2 * w2;
    const long long V = static_cast<long long>(nb) * d * h * w;
    if (i < V * g2) {
        const long long v = i / g2;
        const int c = static_cast<int>(i % g2) * 8;
        u16 a[8], b[8];
        const uint4 ra = reinterpret_cast<const uint4*>(dskip)[i];
        const uint4 rb = *reinterpret_cast<const uint4*>(dcat + v * C + c1 + c);
        __builtin_memcpy(a, &ra, 16);
        __builtin_memcpy(b, &rb, 16);
#pragma unroll
        for (int q = 0; q < 8; ++q) a[q] = F2Bf(Bf2F(a[q]) + Bf2F(b[q]));
        uint4 ro;
        __builtin_memcpy(&ro, a, 16);
        reinterpret_cast<uint4*>(dskip)[i] = ro;
    }
    if (i < Vc * g1) {
        const long long vc = i / g1;
        const int c = static_cast<int>(i % g1) * 8;
        const int x2 = static_cast<int>(vc % w2);
        long long r = vc / w2;
        const int y2 = static_cast<int>(r % h2);
        r /= h2;
        const int z2 = static_cast<int>(r % d2);
        const int n = static_cast<int>(r / d2);
        float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int q = 0; q < 8; ++q) {
            const int z = 2 * z2 + (q >> 2), y = 2 * y2 + ((q >> 1) & 1), x = 2 * x2 + (q & 1);
            const long long v = ((static_cast<long long>(n) * d + z) * h + y) * w + x;
            u16 b[8];
            const uint4 rb = *reinterpret_cast<const uint4*>(dcat + v * C + c);
            __builtin_memcpy(b, &rb, 16);
#pragma unroll
            for (int e = 0; e < 8; ++e) acc[e] += Bf2F(b[e]);
        }
        u16 o[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) o[e] = F2Bf(acc[e]);
        uint4 ro;
        __builtin_memcpy(&ro, o, 16);
        reinterpret_cast<uint4*>(dcoarse)[i] = ro;
    }
}

// Output head (1x1 conv, cin -> 4, fp32) fused with the masked MSE loss:
//   pred = Wh h + bh;  loss = sum_v m_v * sum_c (pred_c - tgt_c)^2 / norm
// writes dh (bf16, d loss / d h), per-block partials of dWh [4][cin], dbh [4] and loss.
constexpr int kHeadBlocks = 512;
template <int cin>
__global__ __launch_bounds__(256) void KHeadLoss(const u16* __restrict__ h, const float* __restrict__ wh,
                                                 const float* __restrict__ bh, const float* __restrict__ tgt,
                                                 const float* __restrict__ wmask, long long V, float inv_norm,
                                                 u16* __restrict__ dh, float* __restrict__ part, float* pred_out) {
    // part[block][4*cin + 4 + 1]; a voxel's cin bf16 activations are cin/8 16-byte loads, its
    // targets and prediction one float4, its activation gradient cin/8 16-byte stores
    static_assert(cin % 8 == 0, "16-byte activation runs");
    constexpr int P = 4 * cin + 5;
    float gw[4 * cin];
    float gb[4] = {0, 0, 0, 0};
    float ls = 0.f;
#pragma unroll
    for (int k = 0; k < 4 * cin; ++k) gw[k] = 0.f;
    for (long long v = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x; v < V;
         v += static_cast<long long>(kHeadBlocks) * 256) {
        float hv[cin];
#pragma unroll
        for (int r = 0; r < cin / 8; ++r) {
            const uint4 raw = *reinterpret_cast<const uint4*>(h + v * cin + 8 * r);
            const uint32_t w4[4] = {raw.x, raw.y, raw.z, raw.w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                hv[8 * r + 2 * q] = Bf2F(static_cast<u16>(w4[q] & 0xFFFFu));
                hv[8 * r + 2 * q + 1] = Bf2F(static_cast<u16>(w4[q] >> 16));
            }
        }
        const float4 tv = tgt ? *reinterpret_cast<const float4*>(tgt + v * 4) : make_float4(0.f, 0.f, 0.f, 0.f);
        const float tg[4] = {tv.x, tv.y, tv.z, tv.w};
        const float m = wmask ? wmask[v] : 1.f;
        float g[4], pr[4];
#pragma unroll
        for (int o = 0; o < 4; ++o) {
            float p = bh[o];
#pragma unroll
            for (int c = 0; c < cin; ++c) p += wh[o * cin + c] * hv[c];
            pr[o] = p;
            const float e = p - tg[o];
            ls += m * e * e * inv_norm;
            g[o] = 2.f * m * e * inv_norm;
            gb[o] += g[o];
#pragma unroll
            for (int c = 0; c < cin; ++c) gw[o * cin + c] += g[o] * hv[c];
        }
        if (pred_out) *reinterpret_cast<float4*>(pred_out + v * 4) = make_float4(pr[0], pr[1], pr[2], pr[3]);
        if (dh) {
#pragma unroll
            for (int r = 0; r < cin / 8; ++r) {
                uint32_t w4[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    float s0 = 0.f, s1 = 0.f;
#pragma unroll
                    for (int o = 0; o < 4; ++o) {
                        s0 += g[o] * wh[o * cin + 8 * r + 2 * q];
                        s1 += g[o] * wh[o * cin + 8 * r + 2 * q + 1];
                    }
                    w4[q] = PackBf2(s0, s1);
                }
                *reinterpret_cast<uint4*>(dh + v * cin + 8 * r) = make_uint4(w4[0], w4[1], w4[2], w4[3]);
            }
        }
    }
    // block partials, in a fixed order: a butterfly over each wave's 64 lanes, then the four
    // waves' sums in wave order
    __shared__ float red[4][P];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < P; ++k) {
        float val = k < 4 * cin ? gw[k] : (k < 4 * cin + 4 ? gb[k - 4 * cin] : ls);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) val += __shfl_xor(val, off, 64);
        if (lane == 0) red[wave][k] = val;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < P; k += 256)
        part[static_cast<long long>(blockIdx.x) * P + k] = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
}

// Adam (Kingma & Ba 2015) on fp32 master weights, refreshing the bf16 copy.
__global__ void KAdam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                      u16* __restrict__ pbf, long long n, float lr, float b1, float b2, float eps, float bc1,
                      float bc2, float gscale) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] * gscale;
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float upd = lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    const float pi = p[i] - upd;
    p[i] = pi;
    if (pbf) pbf[i] = F2Bf(pi);
}

// y += x (bf16 tensors, fp32 add): accumulation of activation gradients.
__global__ void KAddBf16(const u16* __restrict__ x, u16* __restrict__ y, long long n) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = F2Bf(Bf2F(y[i]) + Bf2F(x[i]));
}

// The same, 8 elements per thread (n % 8 == 0): one 16-byte load of each input and one store.
__global__ void KAddBf16x8(const u16* __restrict__ x, u16* __restrict__ y, long long n8) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n8) return;
    u16 xv[8], yv[8];
    const uint4 rx = reinterpret_cast<const uint4*>(x)[i], ry = reinterpret_cast<const uint4*>(y)[i];
    __builtin_memcpy(xv, &rx, 16);
    __builtin_memcpy(yv, &ry, 16);
#pragma unroll
    for (int q = 0; q < 8; ++q) yv[q] = F2Bf(Bf2F(yv[q]) + Bf2F(xv[q]));
    uint4 w;
    __builtin_memcpy(&w, yv, 16);
    reinterpret_cast<uint4*>(y)[i] = w;
}

__global__ void KF32ToBf16(const float* __restrict__ x, u16* __restrict__ y, long long n) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = F2Bf(x[i]);
}

__global__ void KBf16ToF32(const u16* __restrict__ x, float* __restrict__ y, long long n) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] = Bf2F(x[i]);
}

// Surrogate input features from a voxel grid: [V][8] bf16 =
//   (solid, clip(sdf,+-16)/16, u_in/0.1, (x+.5)/W, (y+.5)/H, (z+.5)/D, 0, 0)
__global__ void KFeatures(const uint8_t* __restrict__ solid, const float* __restrict__ sdf, float u_in, int nb, int d,
                          int h, int w, u16* __restrict__ out) {
    const long long v = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long V = static_cast<long long>(nb) * d * h * w;
    if (v >= V) return;
    const int x = static_cast<int>(v % w);
    const int y = static_cast<int>((v / w) % h);
    const int z = static_cast<int>((v / (static_cast<long long>(w) * h)) % d);
    float s = sdf[v];
    s = s != s ? 0.f : (s > 16.f ? 16.f : (s < -16.f ? -16.f : s));
    const float f[8] = {solid[v] ? 1.f : 0.f, s / 16.f, u_in / 0.1f, (x + 0.5f) / w, (y + 0.5f) / h,
                        (z + 0.5f) / d, 0.f, 0.f};
    for (int c = 0; c < 8; ++c) out[v * 8 + c] = F2Bf(f[c]);
}

// Training targets from an LBM field (4 floats per cell): ((rho-1)*3/u^2, u/u_in);
// solid cells get 0 and weight 0.
__global__ void KTargets(const float* __restrict__ field, const uint8_t* __restrict__ solid, float u_in, long long V,
                         float* __restrict__ tgt, float* __restrict__ wmask) {
    const long long v = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const bool s = solid[v] != 0;
    const float inv = 1.f / u_in;
    tgt[v * 4 + 0] = s ? 0.f : (field[v * 4 + 0] - 1.f) * 3.f * inv * inv;
    tgt[v * 4 + 1] = s ? 0.f : field[v * 4 + 1] * inv;
    tgt[v * 4 + 2] = s ? 0.f : field[v * 4 + 2] * inv;
    tgt[v * 4 + 3] = s ? 0.f : field[v * 4 + 3] * inv;
    wmask[v] = s ? 0.f : 1.f;
}

// Inverse of KTargets: a surrogate prediction -> LBM seed field (rho, ux, uy, uz).
__global__ void KPredToField(const float* __restrict__ pred, float u_in, long long V, float* __restrict__ field) {
    const long long v = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= V) return;
    field[v * 4 + 0] = 1.f + pred[v * 4 + 0] * u_in * u_in / 3.f;
    field[v * 4 + 1] = pred[v * 4 + 1] * u_in;
    field[v * 4 + 2] = pred[v * 4 + 2] * u_in;
    field[v * 4 + 3] = pred[v * 4 + 3] * u_in;
}

unsigned Blk(long long n, int b) { return static_cast<unsigned>((n + b - 1) / b); }

// LDS-tiled convolution for the stride-1 modes (forward, upsample-concat forward, and
// the adjoint of a stride-1 convolution).  A workgroup computes a 16 (x) x 4 (y) x 4 (z)
// block of output voxels (wave w = z-slice w, 4 MFMA row tiles of 16 x-voxels each) for
// 16*TN output channels.  For each chunk of up to 8*RMAX input channels it stages the
// 18 x 6 x 6 halo of the input in LDS once (zeros outside the volume; for upsample-concat
// the nearest-upsampled coarse channels and the skip channels side by side), then runs
// the implicit GEMM of the chunk: every input voxel is read from global memory once per
// chunk instead of once per tap.  With 16 output channels the chunk's weights are staged
// in LDS too (each wave would otherwise fetch all of them from L2, every K step a round
// trip: 0.37 of the 0.99 ms of d0's matrix loop on MI355X).
#ifndef FC_NN_CONV_TILED
#define FC_NN_CONV_TILED 1
#endif
constexpr bool kConvTiled = FC_NN_CONV_TILED;
#ifndef FC_NN_CONV_PERSISTENT
#define FC_NN_CONV_PERSISTENT 1
#endif
constexpr bool kConvPersistent = FC_NN_CONV_PERSISTENT;
constexpr int kTX = 16, kTY = 4, kTZ = 4;
constexpr int kHX = kTX + 2, kHY = kTY + 2, kHZ = kTZ + 2, kHV = kHX * kHY * kHZ;  // 648 halo voxels
// LDS row stride (bf16) of a halo voxel for chunks of up to 8*RMAX channels, padded
// against bank conflicts (80 B for 32 channels, 48 B for 8 or 16)
__host__ __device__ constexpr int HaloStride(int rmax) { return rmax >= 3 ? 8 * rmax + 8 : 24; }

// Stages the halo of channels [cb, cb + 8*R8) into sX: one 16-byte run per thread iteration
// (zeros outside the volume; for upsample-concat the nearest-upsampled coarse channels and
// the skip channels side by side).
// The 8 gathered channels [ci, ci+8) of input voxel (n, iz, iy, ix) of a stride-1 gather
// (zeros outside the volume; for upsample-concat the coarse channels at the half-resolution
// voxel, the skip channels at the full-resolution one).
template <int MODE>
__device__ __forceinline__ uint4 InputRun(const ConvArgs& a, int n, int iz, int iy, int ix, int ci) {
    uint4 raw = make_uint4(0u, 0u, 0u, 0u);
    if (ix >= 0 && iy >= 0 && iz >= 0 && ix < a.wi && iy < a.hi && iz < a.di) {
        if (MODE == kUpcat && ci < a.c1) {
            const int h2 = a.hi >> 1, w2 = a.wi >> 1, d2 = a.di >> 1;
            const long long vi = ((static_cast<long long>(n) * d2 + (iz >> 1)) * h2 + (iy >> 1)) * w2 + (ix >> 1);
            raw = *reinterpret_cast<const uint4*>(a.in2 + vi * a.c1 + ci);
        } else {
            const int cs = MODE == kUpcat ? a.kin - a.c1 : a.kin;
            const int c0 = MODE == kUpcat ? ci - a.c1 : ci;
            const long long vi = ((static_cast<long long>(n) * a.di + iz) * a.hi + iy) * a.wi + ix;
            raw = *reinterpret_cast<const uint4*>(a.in + vi * cs + c0);
        }
    }
    return raw;
}

template <int MODE, int R8, int CS>
__device__ __forceinline__ void StageHalo(const ConvArgs& a, u16* sX, int n, int x0, int y0, int z0, int cb) {
    for (int e = threadIdx.x; e < kHV * R8; e += 256) {
        const int hv = e / R8, r8 = e - hv * R8;
        const int hx = hv % kHX, hy = (hv / kHX) % kHY, hz = hv / (kHX * kHY);
        const uint4 raw = InputRun<MODE>(a, n, z0 - 1 + hz, y0 - 1 + hy, x0 - 1 + hx, cb + r8 * 8);
        *reinterpret_cast<uint4*>(&sX[hv * CS + r8 * 8]) = raw;
    }
}

// Stages the weights of channels [cb, cb + 8*R8) of the 16 output channels into sW as
// [tap][8-channel run][output channel][8]: a lane group's B fragment (16 output channels
// of one run of one tap) is then 256 contiguous bytes.
template <int R8>
__device__ __forceinline__ void StageWeights16(const ConvArgs& a, u16* sW, int K, int cb) {
    for (int e = threadIdx.x; e < 16 * 27 * R8; e += 256) {
        const int nn = e & 15, tc = e >> 4;
        const int t = tc / R8, c8 = tc - t * R8;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (nn < a.nout) v = *reinterpret_cast<const uint4*>(a.w + static_cast<long long>(nn) * K + t * a.kin + cb + c8 * 8);
        *reinterpret_cast<uint4*>(&sW[e * 8]) = v;
    }
}

#ifndef FC_NN_CHUNK_UNROLL
#define FC_NN_CHUNK_UNROLL 1
#endif
// The implicit GEMM of one staged chunk of 8*R8 channels, for wide output tiles (TN = 4,
// whose 16*TN accumulators leave no room for the six-row schedule below): its 27*R8 (tap, 8-channel run)
// pairs, four per 32-wide MFMA K step (lane group g takes run 4s + g).  The chunk width is
// a compile-time constant, so the run -> (tap, channel) arithmetic is shifts and constant
// divisions; the operands of step s+1 (weights from global memory, halo fragments from
// LDS) are loaded before the MFMAs of step s (software pipelining), so their latency
// overlaps one step of matrix work.
template <int MODE, int TN, int R8, int CS>
__device__ __forceinline__ void ChunkGemmRuns(const ConvArgs& a, const u16* sX, int wave, int row, int grp, int nbase,
                                          int K, int cb, f32x4 (&acc)[4][TN]) {
    constexpr int kRuns = 27 * R8;
    constexpr int kSteps = (kRuns + 3) / 4;
    const u16* wrow[TN];
    bool wok[TN];
#pragma unroll
    for (int j = 0; j < TN; ++j) {
        const int nn = nbase + j * 16 + row;
        wok[j] = nn < a.nout;
        wrow[j] = a.w + static_cast<long long>(wok[j] ? nn : 0) * K + cb;
    }
    const int lbase = (wave * kHY * kHX + row) * CS;  // the lane's halo row at tap (0,0,0), m = 0
    // step s's operands: the weight fragments of its runs and the four halo fragments
    auto load = [&](int s, uint4 (&bv)[TN], uint4 (&av)[4]) {
        const int r = 4 * s + grp;
        const bool rok = r < kRuns;
        const int t = r / R8, c8 = r - t * R8;
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            bv[j] = make_uint4(0u, 0u, 0u, 0u);
            if (rok && wok[j]) bv[j] = *reinterpret_cast<const uint4*>(wrow[j] + t * a.kin + c8 * 8);
        }
        const int kz = t / 9, ky = (t / 3) % 3, kx = t % 3;
        // halo offsets of the tap: forward/upcat read o + k - 1, the adjoint o + 1 - k
        const int oz = MODE == kAdj ? 2 - kz : kz, oy = MODE == kAdj ? 2 - ky : ky, ox = MODE == kAdj ? 2 - kx : kx;
        const int tap = lbase + ((oz * kHY + oy) * kHX + ox) * CS + c8 * 8;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            av[m] = make_uint4(0u, 0u, 0u, 0u);
            if (rok) av[m] = *reinterpret_cast<const uint4*>(&sX[tap + m * kHX * CS]);
        }
    };
    uint4 bcur[TN], acur[4];
    load(0, bcur, acur);
#pragma unroll FC_NN_CHUNK_UNROLL
    for (int s = 0; s < kSteps; ++s) {
        uint4 bnext[TN], anext[4];
        if (s + 1 < kSteps) load(s + 1, bnext, anext);  // issued ahead of this step's MFMAs
        bf16x8 bfr[TN];
#pragma unroll
        for (int j = 0; j < TN; ++j) __builtin_memcpy(&bfr[j], &bcur[j], 16);
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            bf16x8 afr;
            __builtin_memcpy(&afr, &acur[m], 16);
#pragma unroll
            for (int j = 0; j < TN; ++j) acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfr[j], afr, acc[m][j], 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < TN; ++j) bcur[j] = bnext[j];
#pragma unroll
        for (int m = 0; m < 4; ++m) acur[m] = anext[m];
    }
}

// The implicit GEMM of one staged chunk of 8*R8 channels for 16 output channels, weights
// in LDS (LDSW: sW, as staged by StageWeights16) or read from global memory (L2).  Which
// is faster depends on the LDS left for the halo: the weights of a 32-channel chunk (27.6
// KB) would cut a CU to two workgroups, and on d0 (48 -> 16 channels, 16.8M voxels) that
// costs more (2.33 ms) than the global reads do (1.84 ms); with 8- or 16-channel chunks
// three or four workgroups still fit (e0: 0.75 -> 0.61 ms).  The wave's four output rows (y = y0 + m) read
// halo rows m + oy(ky), so the three ky taps of one (kz, kx, 8-channel run) combination
// share six halo fragments: each K step takes four such combinations (lane group g takes
// combination 4s + g), loads the six halo fragments once and the weight fragments of the
// three ky taps, and runs 12 MFMAs (acc[m] += A[m + oy(ky)] * B[ky]).  The chunk width is
// a compile-time constant, so the combination -> (kz, kx, run) arithmetic is shifts and
// constant divisions; the operands of step s+1 are loaded before the MFMAs of step s.
// sW holds WR runs per tap ([tap][WR runs][16 output channels][8]); the chunk's runs
// start at run wrun0 (WR = R8, wrun0 = 0 when the chunk's own weights are staged).
template <int MODE, int R8, int CS, bool LDSW, int WR = R8>
__device__ __forceinline__ void ChunkGemmRows16(const ConvArgs& a, const u16* sX, const u16* sW, int wave, int row,
                                                int grp, int K, int cb, f32x4 (&acc)[4][1], int wrun0 = 0) {
    constexpr int kCombos = 9 * R8;
    constexpr int kSteps = (kCombos + 3) / 4;
    const bool wok = row < a.nout;
    const u16* wrow = a.w + static_cast<long long>(wok ? row : 0) * K + cb;
    const int lbase = (wave * kHY * kHX + row) * CS;  // the lane's halo row at offset (0,0,0), m = 0
    // step s's operands: the weight fragments of the three ky taps and six halo rows
    auto load = [&](int s, uint4 (&bv)[3], uint4 (&av)[6]) {
        const int c = 4 * s + grp;
        const bool cok = c < kCombos;
        const int kzx = c / R8, c8 = c - kzx * R8;
        const int kz = kzx / 3, kx = kzx - kz * 3;
#pragma unroll
        for (int ky = 0; ky < 3; ++ky) {
            const int t = (kz * 3 + ky) * 3 + kx;
            bv[ky] = make_uint4(0u, 0u, 0u, 0u);
            if constexpr (LDSW) {
                if (cok) bv[ky] = *reinterpret_cast<const uint4*>(&sW[((t * WR + wrun0 + c8) * 16 + row) * 8]);
            } else {
                if (cok && wok) bv[ky] = *reinterpret_cast<const uint4*>(wrow + t * a.kin + c8 * 8);
            }
        }
        // halo offsets: forward/upcat read o + k - 1, the adjoint o + 1 - k
        const int oz = MODE == kAdj ? 2 - kz : kz, ox = MODE == kAdj ? 2 - kx : kx;
        const int base = lbase + (oz * kHY * kHX + ox) * CS + c8 * 8;
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            av[r] = make_uint4(0u, 0u, 0u, 0u);
            if (cok) av[r] = *reinterpret_cast<const uint4*>(&sX[base + r * kHX * CS]);
        }
    };
    uint4 bcur[3], acur[6];
    load(0, bcur, acur);
#pragma unroll 1
    for (int s = 0; s < kSteps; ++s) {
        uint4 bnext[3], anext[6];
        if (s + 1 < kSteps) load(s + 1, bnext, anext);  // issued ahead of this step's MFMAs
        bf16x8 afr[6];
#pragma unroll
        for (int r = 0; r < 6; ++r) __builtin_memcpy(&afr[r], &acur[r], 16);
#pragma unroll
        for (int ky = 0; ky < 3; ++ky) {
            const int oy = MODE == kAdj ? 2 - ky : ky;
            bf16x8 bfr;
            __builtin_memcpy(&bfr, &bcur[ky], 16);
#pragma unroll
            for (int m = 0; m < 4; ++m)
                acc[m][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfr, afr[m + oy], acc[m][0], 0, 0, 0);
        }
#pragma unroll
        for (int ky = 0; ky < 3; ++ky) bcur[ky] = bnext[ky];
#pragma unroll
        for (int r = 0; r < 6; ++r) acur[r] = anext[r];
    }
}

// Weights staged in LDS: 16 output channels and chunks of at most 16 channels.
__host__ __device__ constexpr bool LdsWeights(int tn, int rmax) { return tn == 1 && rmax <= 2; }

// One chunk: stage the halo (and the weights, when they go through LDS), then its GEMM.
template <int MODE, int TN, int R8, int CS, bool LDSW>
__device__ __forceinline__ void ConvChunk(const ConvArgs& a, u16* sX, u16* sW, int n, int x0, int y0, int z0, int cb,
                                          int wave, int row, int grp, int nbase, int K, f32x4 (&acc)[4][TN]) {
    StageHalo<MODE, R8, CS>(a, sX, n, x0, y0, z0, cb);
    if constexpr (LDSW) StageWeights16<R8>(a, sW, K, cb);
    __syncthreads();
    if constexpr (TN == 1)
        ChunkGemmRows16<MODE, R8, CS, LDSW>(a, sX, sW, wave, row, grp, K, cb, acc);
    else
        ChunkGemmRuns<MODE, TN, R8, CS>(a, sX, wave, row, grp, nbase, K, cb, acc);
    __syncthreads();
}

// Chunks of up to 8*RMAX input channels.  With 16 output channels the register budget is
// set to three waves per SIMD (170 VGPRs), as many as the LDS footprint allows.
template <int MODE, int TN, int RMAX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(TN == 1 ? 3 : 1))) void KConvTile(const ConvArgs a) {
    constexpr int CS = HaloStride(RMAX);
    constexpr bool LDSW = LdsWeights(TN, RMAX);
    constexpr int kW = LDSW ? 16 * 27 * 8 * RMAX : 0;  // staged weights (bf16)
    __shared__ __attribute__((aligned(16))) u16 sX[kHV * CS + kW];
    u16* sW = sX + kHV * CS;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tiles_x = (a.wd + kTX - 1) / kTX, tiles_y = (a.hd + kTY - 1) / kTY, tiles_z = (a.dd + kTZ - 1) / kTZ;
    int b = blockIdx.x;
    const int tx = b % tiles_x;
    b /= tiles_x;
    const int ty = b % tiles_y;
    b /= tiles_y;
    const int tz = b % tiles_z;
    const int n = b / tiles_z;
    const int x0 = tx * kTX, y0 = ty * kTY, z0 = tz * kTZ;
    const int nbase = blockIdx.y * (16 * TN);
    const int row = lane & 15, grp = lane >> 4;
    f32x4 acc[4][TN];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[m][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int K = 27 * a.kin;
    for (int cb = 0; cb < a.kin; cb += 8 * RMAX) {
        const int r8 = (a.kin - cb < 8 * RMAX ? a.kin - cb : 8 * RMAX) / 8;  // runs in this chunk
        // a ragged last chunk of r8 < RMAX runs takes its own instantiation
        constexpr int R2 = RMAX < 2 ? RMAX : 2, R3 = RMAX < 3 ? RMAX : 3;
        if (r8 >= RMAX)
            ConvChunk<MODE, TN, RMAX, CS, LDSW>(a, sX, sW, n, x0, y0, z0, cb, wave, row, grp, nbase, K, acc);
        else if (r8 == 1)
            ConvChunk<MODE, TN, 1, CS, LDSW>(a, sX, sW, n, x0, y0, z0, cb, wave, row, grp, nbase, K, acc);
        else if (r8 == 2)
            ConvChunk<MODE, TN, R2, CS, LDSW>(a, sX, sW, n, x0, y0, z0, cb, wave, row, grp, nbase, K, acc);
        else
            ConvChunk<MODE, TN, R3, CS, LDSW>(a, sX, sW, n, x0, y0, z0, cb, wave, row, grp, nbase, K, acc);
    }
    // epilogue: the weights are the MFMA's first operand, so lane (row, grp) holds output
    // channels 16j + 4grp .. +3 of voxel x0 + row
    const int z = z0 + wave, x = x0 + row;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int y = y0 + m;
        if (x >= a.wd || y >= a.hd || z >= a.dd) continue;
        const long long v = ((static_cast<long long>(n) * a.dd + z) * a.hd + y) * a.wd + x;
#pragma unroll
        for (int j = 0; j < TN; ++j) StoreC4(a, v, nbase + j * 16 + 4 * grp, acc[m][j]);
    }
}

// Persistent convolution for the network's stride-1 layers of 16 to 64 output channels
// (e0: 8 inputs -> 16; e1b: 32 -> 32; e2b: 64 -> 64; d2: 192 = 128 upsampled + 64 skip ->
// 64; d1: 96 = 64 upsampled + 32 skip -> 32; d0: 48 = 32 upsampled + 16 skip -> 16) and
// the stride-1 adjoints run as forward convolutions, forward and upsample-concat modes,
// CO = 16*NT output channels per workgroup (blockIdx.y selects the block of CO channels;
// a.nout must be a multiple of CO, which the launcher checks: bias reads and output stores
// are not bounded by a.nout).  A workgroup loops over 16 (x) x 4 (y) x 4*ZW (z) output
// tiles; wave w owns z slices w*ZW .. w*ZW+ZW-1, each as four MFMA tiles of 16 x-voxels
// (one per y row) by the CO output channels.  Input channels go in chunks of RC 8-channel
// runs (RC = RCH, 2 or 3, unless the layer has fewer runs; 48 bytes per halo voxel hold two
// runs and a pad, or three runs, either way a stride at which the 16 lanes of a fragment
// read hit distinct banks).  Each chunk's halo (18 x 6 x (4*ZW+2) voxels)
// and weights are loaded into registers while the previous chunk's matrix loop runs and
// written to LDS between two barriers; the matrix loop reads only LDS.  (Weights read from
// global memory inside the loop, issued after the halo prefetch, make every wait on them a
// wait on the prefetch too: memory counters retire in order.)  A layer with one chunk
// stages its weights once.
// Matrix loop: K step s gives lane group g the combination c = 4s + g of (kz, kx, run);
// the wave's four output rows y = m read halo rows m + ky, so the three ky taps share six
// halo fragments per z slice (12*NT MFMAs per slice and step, each weight fragment feeding
// 4*ZW of them).  Combinations past the chunk's 9*RC (only in the last step) read a valid
// halo address and zero weights, so the loop has no branches: with 16 output channels the
// weights come from a zero block in LDS (no extra instructions: d0 0.97 against 1.00 ms
// with selects), with 32 the fragments are replaced by zeros (the block would take LDS
// that the second workgroup per CU needs).  The instruction's operands are the weights
// (rows = output channels) and the input (columns = voxels), so a lane ends with 4
// consecutive output channels of one voxel per 16-channel tile and the epilogue writes 8
// bytes (bf16) or 16 bytes (fp32) per lane and tile.
template <int MODE, int NRUNS, int ZW, int NT, int RCH = 2>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void KConvTileP(const ConvArgs a) {
    static_assert(MODE == kFwd || MODE == kUpcat, "forward gathers only");
    constexpr int CO = 16 * NT;                        // output channels of the workgroup
    static_assert(RCH == 2 || RCH == 3, "chunks of 2 or 3 runs");
    constexpr int RC = NRUNS < RCH ? NRUNS : RCH;      // 8-channel runs per chunk
    // bf16 per halo voxel: 48 B for 2 or 3 runs, 16 B for one (16 lanes then read 256
    // contiguous bytes; the smaller halo leaves LDS for more workgroups per CU)
    constexpr int CS = RC == 1 ? 8 : 24;
    constexpr int TZ = 4 * ZW, HZ = TZ + 2, HV = kHX * kHY * HZ;
    static_assert(NRUNS % RC == 0, "whole chunks");
    constexpr int kChunks = NRUNS / RC;
    constexpr int kPV = (HV + 255) / 256;              // halo voxels per thread (RC runs each)
    constexpr int WE = 27 * RC * CO;                   // weight fragments of a chunk (16 B each)
    constexpr int kPW = (WE + 255) / 256;
    constexpr int kCombos = 9 * RC, kSteps = (kCombos + 3) / 4;
    constexpr int kPlane = kHY * kHX * CS;             // one halo z layer (bf16)
    constexpr int kTapW = RC * CO * 8;                 // one tap's weights (bf16)
    constexpr bool ZB = NT == 1;                       // zero weights from an LDS block
    constexpr int kZeroW = WE * 8;                     // the zero block: reached at offsets
    constexpr int kZeroLen = ZB ? 6 * kTapW + CO * 8 : 0;  // 0, 3, 6 taps (+ CO rows)
    __shared__ __attribute__((aligned(16))) u16 sX[HV * CS];
    __shared__ __attribute__((aligned(16))) u16 sW[WE * 8 + kZeroLen];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, grp = lane >> 4;
    const int K = 27 * a.kin;
    const int nbase = blockIdx.y * CO;
    const int tiles_x = (a.wd + kTX - 1) / kTX, tiles_y = (a.hd + kTY - 1) / kTY, tiles_z = (a.dd + TZ - 1) / TZ;
    const int ntiles = a.nb * tiles_z * tiles_y * tiles_x;
    int tile = blockIdx.x;
    if (tile >= ntiles) return;  // whole workgroup, before any barrier
    auto decode = [&](int t, int& n, int& x0, int& y0, int& z0) {
        x0 = (t % tiles_x) * kTX;
        t /= tiles_x;
        y0 = (t % tiles_y) * kTY;
        t /= tiles_y;
        z0 = (t % tiles_z) * TZ;
        n = t / tiles_z;
    };
    // Halo prefetch: a thread takes whole halo voxels and loads their RC runs with one address
    // computation per voxel (the runs of a chunk are contiguous channels; for upsample-concat
    // a run comes from the coarse or the skip tensor, the same for every lane).
    uint4 pf[kPV][RC], pw[kPW];
    auto prefetch_halo = [&](int t, int c) {
        int n, x0, y0, z0;
        decode(t, n, x0, y0, z0);
#pragma unroll
        for (int j = 0; j < kPV; ++j) {
            const int hv = threadIdx.x + 256 * j;
#pragma unroll
            for (int r = 0; r < RC; ++r) pf[j][r] = make_uint4(0u, 0u, 0u, 0u);
            if (hv < HV) {
                const int hx = hv % kHX, hy = (hv / kHX) % kHY, hz = hv / (kHX * kHY);
                const int ix = x0 - 1 + hx, iy = y0 - 1 + hy, iz = z0 - 1 + hz;
                if (ix >= 0 && iy >= 0 && iz >= 0 && ix < a.wi && iy < a.hi && iz < a.di) {
                    const int cs = MODE == kUpcat ? a.kin - a.c1 : a.kin;
                    const u16* fine = a.in + (((static_cast<long long>(n) * a.di + iz) * a.hi + iy) * a.wi + ix) * cs;
                    const u16* coarse = nullptr;
                    if constexpr (MODE == kUpcat)
                        coarse = a.in2 + (((static_cast<long long>(n) * (a.di >> 1) + (iz >> 1)) * (a.hi >> 1) + (iy >> 1)) *
                                              (a.wi >> 1) + (ix >> 1)) * a.c1;
#pragma unroll
                    for (int r = 0; r < RC; ++r) {
                        const int ci = 8 * RC * c + 8 * r;
                        if (MODE == kUpcat && ci < a.c1)
                            pf[j][r] = *reinterpret_cast<const uint4*>(coarse + ci);
                        else
                            pf[j][r] = *reinterpret_cast<const uint4*>(fine + (MODE == kUpcat ? ci - a.c1 : ci));
                    }
                }
            }
        }
    };
    // sW = [tap][run][output channel][8 channels]
    auto prefetch_weights = [&](int c) {
#pragma unroll
        for (int j = 0; j < kPW; ++j) {
            const int e = threadIdx.x + 256 * j;
            pw[j] = make_uint4(0u, 0u, 0u, 0u);
            if (e < WE) {
                const int co = nbase + e % CO, tr = e / CO;
                const int t = tr / RC, r = tr - t * RC;
                if (co < a.nout) pw[j] = *reinterpret_cast<const uint4*>(a.w + static_cast<long long>(co) * K + t * a.kin +
                                                                         8 * RC * c + r * 8);
            }
        }
    };
    auto commit_halo = [&]() {
#pragma unroll
        for (int j = 0; j < kPV; ++j) {
            const int hv = threadIdx.x + 256 * j;
            if (hv < HV) {
#pragma unroll
                for (int r = 0; r < RC; ++r) *reinterpret_cast<uint4*>(&sX[hv * CS + r * 8]) = pf[j][r];
            }
        }
    };
    auto commit_weights = [&]() {
#pragma unroll
        for (int j = 0; j < kPW; ++j) {
            const int e = threadIdx.x + 256 * j;
            if (e < WE) *reinterpret_cast<uint4*>(&sW[e * 8]) = pw[j];
        }
    };
    // LDS element offsets of step s for this lane: its halo fragment at (slice 0, row 0) and
    // its weight fragment at ky = 0, tile 0; past the chunk's combinations, combination 0's
    // halo and the zero block (ZB) or combination 0's weights, replaced by zeros in load_w.
    auto step_offsets = [&](int s, int& xo, int& wo, bool& ok) {
        const int c = 4 * s + grp;
        ok = c < kCombos;
        const int cc = ok ? c : 0;
        const int kzx = cc / RC, run = cc - kzx * RC;
        const int kz = kzx / 3, kx = kzx - kz * 3;
        xo = (wave * ZW + kz) * kPlane + (row + kx) * CS + run * 8;
        wo = (ok || !ZB ? ((kz * 9 + kx) * RC + run) * CO * 8 : kZeroW) + row * 8;
    };
    auto load_w = [&](int wo, bool ok, bf16x8 (&w)[3][NT]) {
#pragma unroll
        for (int ky = 0; ky < 3; ++ky)
#pragma unroll
            for (int jt = 0; jt < NT; ++jt) {
                uint4 v = *reinterpret_cast<const uint4*>(&sW[wo + ky * 3 * kTapW + jt * 128]);
                if (!ZB && !ok) v = make_uint4(0u, 0u, 0u, 0u);
                __builtin_memcpy(&w[ky][jt], &v, 16);
            }
    };
    auto load_x = [&](int xo, int j, bf16x8 (&x)[6]) {
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            const uint4 v = *reinterpret_cast<const uint4*>(&sX[xo + j * kPlane + r * kHX * CS]);
            __builtin_memcpy(&x[r], &v, 16);
        }
    };
    // the zero block: written once, never overwritten by the chunk weights
    if constexpr (ZB)
        for (int e = threadIdx.x; e < kZeroLen / 8; e += 256)
            *reinterpret_cast<uint4*>(&sW[kZeroW + e * 8]) = make_uint4(0u, 0u, 0u, 0u);
    float bias[NT][4];
#pragma unroll
    for (int jt = 0; jt < NT; ++jt)
#pragma unroll
        for (int q = 0; q < 4; ++q) bias[jt][q] = a.bias ? a.bias[nbase + 16 * jt + 4 * grp + q] : 0.f;
    prefetch_halo(tile, 0);
    prefetch_weights(0);
    if constexpr (kChunks == 1) {
        commit_weights();  // read after the first barrier below
    }
    f32x4 acc[ZW][4][NT];
    for (;;) {
#pragma unroll
        for (int j = 0; j < ZW; ++j)
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int jt = 0; jt < NT; ++jt) acc[j][m][jt] = f32x4{0.f, 0.f, 0.f, 0.f};
        for (int c = 0; c < kChunks; ++c) {
            __syncthreads();  // the previous chunk's matrix loop is done with sX and sW
            commit_halo();
            if constexpr (kChunks > 1) commit_weights();
            __syncthreads();
            // the next chunk: this tile's next one, else the next tile's first
            const int nt = c + 1 < kChunks ? tile : tile + static_cast<int>(gridDim.x);
            if (nt < ntiles) {
                prefetch_halo(nt, c + 1 < kChunks ? c + 1 : 0);
                if constexpr (kChunks > 1) prefetch_weights(c + 1 < kChunks ? c + 1 : 0);
            }
            // matrix loop, unrolled: the operands of the next (step, slice) are read before this
            // one's MFMAs
            int xo, wo;
            bool ok;
            step_offsets(0, xo, wo, ok);
            bf16x8 wc[3][NT], xc[6];
            load_w(wo, ok, wc);
            load_x(xo, 0, xc);
#pragma unroll
            for (int s = 0; s < kSteps; ++s) {
#pragma unroll
                for (int j = 0; j < ZW; ++j) {
                    const bool last = j + 1 == ZW;
                    bf16x8 wn[3][NT], xn[6];
                    int xo2 = xo, wo2 = wo;
                    bool ok2 = ok;
                    if (!last) {
                        load_x(xo, j + 1, xn);
                    } else if (s + 1 < kSteps) {
                        step_offsets(s + 1, xo2, wo2, ok2);
                        load_w(wo2, ok2, wn);
                        load_x(xo2, 0, xn);
                    }
#pragma unroll
                    for (int ky = 0; ky < 3; ++ky)
#pragma unroll
                        for (int m = 0; m < 4; ++m)
#pragma unroll
                            for (int jt = 0; jt < NT; ++jt)
                                acc[j][m][jt] =
                                    __builtin_amdgcn_mfma_f32_16x16x32_bf16(wc[ky][jt], xc[m + ky], acc[j][m][jt], 0, 0, 0);
                    if (!last || s + 1 < kSteps) {
#pragma unroll
                        for (int r = 0; r < 6; ++r) xc[r] = xn[r];
                    }
                    if (last && s + 1 < kSteps) {
#pragma unroll
                        for (int ky = 0; ky < 3; ++ky)
#pragma unroll
                            for (int jt = 0; jt < NT; ++jt) wc[ky][jt] = wn[ky][jt];
                        xo = xo2;
                        wo = wo2;
                        ok = ok2;
                    }
                }
            }
        }
        // epilogue: lane holds output channels 16*jt + 4*grp .. +3 of voxel x0 + row
        int n, x0, y0, z0;
        decode(tile, n, x0, y0, z0);
        const int x = x0 + row;
#pragma unroll
        for (int j = 0; j < ZW; ++j) {
            const int z = z0 + wave * ZW + j;
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int y = y0 + m;
                if (x >= a.wd || y >= a.hd || z >= a.dd) continue;
                const long long v = ((static_cast<long long>(n) * a.dd + z) * a.hd + y) * a.wd + x;
#pragma unroll
                for (int jt = 0; jt < NT; ++jt) {
                    float val[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        val[q] = acc[j][m][jt][q] + bias[jt][q];
                        if (a.act == 1) val[q] = val[q] > 0.f ? val[q] : kLeak * val[q];
                    }
                    const long long o = v * a.nout + nbase + 16 * jt + 4 * grp;
                    if (a.outf) {
                        *reinterpret_cast<float4*>(a.outf + o) = make_float4(val[0], val[1], val[2], val[3]);
                    } else {
                        *reinterpret_cast<uint2*>(a.out + o) = make_uint2(PackBf2(val[0], val[1]), PackBf2(val[2], val[3]));
                    }
                }
            }
        }
        tile += gridDim.x;
        if (tile >= ntiles) break;
    }
}


// LDS-tiled forward convolution with stride 2 (the encoder's downsampling layers).  A
// workgroup computes a 16 (x) x 4 (y) x 4 (z) block of output voxels (wave w = z-slice w,
// 4 MFMA row tiles of 16 x-voxels, one per y row) for 16*TN output channels, one 8-channel
// run of the input at a time.  Per run it holds in LDS the 33 x 9 x 9 input halo of the
// block, 16 bytes per voxel, each halo row split by x parity (even x at slots 0..16, odd x
// at 17..32): the three kx taps of 16 consecutive output voxels (input x = 2x + kx - 1) then
// read 16 consecutive slots, 256 contiguous bytes and no bank conflicts, where the unsplit
// row would be read at a 32-byte stride; and the run's weights as [tap][output channel][8].
// The next run's halo and weights are loaded into registers while the current run's matrix
// loop runs (which reads only LDS, so no wait on those loads stalls it) and written to LDS
// between two barriers.  The weights are the MFMA's first operand, so a lane ends with 4
// consecutive output channels of one voxel (8- or 16-byte stores).  The implicit GEMM of a run takes four taps per 32-wide K step (lane
// group g takes tap 4s + g; the 28th is zero).  Every input voxel is read from global
// memory about 1.45 times per run (the halo overlap) instead of once per tap.
constexpr int kS2X = 2 * kTX + 1, kS2Y = 2 * kTY + 1, kS2Z = 2 * kTZ + 1;  // 33 x 9 x 9 input halo
constexpr int kS2V = kS2X * kS2Y * kS2Z;
__device__ __forceinline__ int S2Slot(int hx) { return (hx & 1) ? (kS2X + 1) / 2 + (hx >> 1) : (hx >> 1); }

template <int TN>
__global__ __launch_bounds__(256) void KConvTileS2(const ConvArgs a) {
    constexpr int kPF = (kS2V + 255) / 256;  // halo prefetch registers (16 B each)
    constexpr int WE = 27 * 16 * TN;         // weight fragments of a run (16 B each)
    constexpr int kPW = (WE + 255) / 256;
    __shared__ __attribute__((aligned(16))) u16 sX[kS2V * 8];
    __shared__ __attribute__((aligned(16))) u16 sW[WE * 8];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tiles_x = (a.wd + kTX - 1) / kTX, tiles_y = (a.hd + kTY - 1) / kTY, tiles_z = (a.dd + kTZ - 1) / kTZ;
    int b = blockIdx.x;
    const int tx = b % tiles_x;
    b /= tiles_x;
    const int ty = b % tiles_y;
    b /= tiles_y;
    const int tz = b % tiles_z;
    const int n = b / tiles_z;
    const int x0 = tx * kTX, y0 = ty * kTY, z0 = tz * kTZ;
    const int nbase = blockIdx.y * (16 * TN);
    const int row = lane & 15, grp = lane >> 4;
    const int K = 27 * a.kin;
    uint4 pf[kPF], pw[kPW];
    auto prefetch = [&](int cb) {
#pragma unroll
        for (int j = 0; j < kPF; ++j) {
            const int hv = threadIdx.x + 256 * j;
            pf[j] = make_uint4(0u, 0u, 0u, 0u);
            if (hv < kS2V) {
                const int hx = hv % kS2X, hy = (hv / kS2X) % kS2Y, hz = hv / (kS2X * kS2Y);
                const int ix = 2 * x0 - 1 + hx, iy = 2 * y0 - 1 + hy, iz = 2 * z0 - 1 + hz;
                if (ix >= 0 && iy >= 0 && iz >= 0 && ix < a.wi && iy < a.hi && iz < a.di) {
                    const long long vi = ((static_cast<long long>(n) * a.di + iz) * a.hi + iy) * a.wi + ix;
                    pf[j] = *reinterpret_cast<const uint4*>(a.in + vi * a.kin + cb);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < kPW; ++j) {
            const int e = threadIdx.x + 256 * j;  // e = t * 16TN + co
            pw[j] = make_uint4(0u, 0u, 0u, 0u);
            if (e < WE) {
                const int t = e / (16 * TN), co = nbase + (e - t * (16 * TN));
                if (co < a.nout) pw[j] = *reinterpret_cast<const uint4*>(a.w + static_cast<long long>(co) * K + t * a.kin + cb);
            }
        }
    };
    auto commit = [&]() {
#pragma unroll
        for (int j = 0; j < kPF; ++j) {
            const int hv = threadIdx.x + 256 * j;
            if (hv < kS2V) {
                const int hx = hv % kS2X, hyz = hv / kS2X;
                *reinterpret_cast<uint4*>(&sX[(hyz * kS2X + S2Slot(hx)) * 8]) = pf[j];
            }
        }
#pragma unroll
        for (int j = 0; j < kPW; ++j) {
            const int e = threadIdx.x + 256 * j;
            if (e < WE) *reinterpret_cast<uint4*>(&sW[e * 8]) = pw[j];
        }
    };
    f32x4 acc[4][TN];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int j = 0; j < TN; ++j) acc[m][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    constexpr int kSteps = 7;  // 27 taps, four per step
    // step s's operands: weights of tap 4s + grp for the TN channel tiles, and the halo
    // fragments of the wave's four y rows (zeros for the 28th tap)
    auto load = [&](int s, uint4 (&bv)[TN], uint4 (&av)[4]) {
        const int t = 4 * s + grp;
        const bool tok = t < 27;
        const int tt = tok ? t : 0;
        const int kz = tt / 9, ky = (tt / 3) % 3, kx = tt % 3;
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            bv[j] = make_uint4(0u, 0u, 0u, 0u);
            if (tok) bv[j] = *reinterpret_cast<const uint4*>(&sW[((tt * TN + j) * 16 + row) * 8]);
        }
        const int slot = kx == 1 ? (kS2X + 1) / 2 + row : row + (kx >> 1);
        const int base = ((2 * wave + kz) * kS2Y + ky) * kS2X + slot;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            av[m] = make_uint4(0u, 0u, 0u, 0u);
            if (tok) av[m] = *reinterpret_cast<const uint4*>(&sX[(base + 2 * m * kS2X) * 8]);
        }
    };
    prefetch(0);
    for (int cb = 0; cb < a.kin; cb += 8) {
        if (cb > 0) __syncthreads();  // the previous run's matrix loop is done with sX and sW
        commit();
        __syncthreads();
        if (cb + 8 < a.kin) prefetch(cb + 8);  // in flight during this run's matrix loop
        uint4 bcur[TN], acur[4];
        load(0, bcur, acur);
#pragma unroll
        for (int s = 0; s < kSteps; ++s) {
            uint4 bnext[TN], anext[4];
            if (s + 1 < kSteps) load(s + 1, bnext, anext);
            bf16x8 bfr[TN];
#pragma unroll
            for (int j = 0; j < TN; ++j) __builtin_memcpy(&bfr[j], &bcur[j], 16);
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                bf16x8 afr;
                __builtin_memcpy(&afr, &acur[m], 16);
#pragma unroll
                for (int j = 0; j < TN; ++j)
                    acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bfr[j], afr, acc[m][j], 0, 0, 0);
            }
            if (s + 1 < kSteps) {
#pragma unroll
                for (int j = 0; j < TN; ++j) bcur[j] = bnext[j];
#pragma unroll
                for (int m = 0; m < 4; ++m) acur[m] = anext[m];
            }
        }
    }
    // epilogue: lane (row, grp) holds output channels 16j + 4grp .. +3 of voxel x0 + row
    const int z = z0 + wave, x = x0 + row;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int y = y0 + m;
        if (x >= a.wd || y >= a.hd || z >= a.dd) continue;
        const long long v = ((static_cast<long long>(n) * a.dd + z) * a.hd + y) * a.wd + x;
#pragma unroll
        for (int j = 0; j < TN; ++j) StoreC4(a, v, nbase + j * 16 + 4 * grp, acc[m][j]);
    }
}

// Workgroups of a persistent launch: the CUs times the kernel's resident workgroups per CU.
template <class K>
unsigned PersistentGrid(K kernel, long long work) {
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0) != hipSuccess || per_cu <= 0) per_cu = 1;
    const long long g = static_cast<long long>(cus) * per_cu;
    return static_cast<unsigned>(work < g ? work : g);
}

template <int MODE, int TN>
void LaunchConvTile(const ConvArgs& a, dim3 grid, hipStream_t s) {
    // chunk width: 8 input channels -> one run, 16 or 24 -> two, otherwise four
    if (a.kin <= 8)
        hipLaunchKernelGGL((KConvTile<MODE, TN, 1>), grid, dim3(256), 0, s, a);
    else if (a.kin <= 24)
        hipLaunchKernelGGL((KConvTile<MODE, TN, 2>), grid, dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((KConvTile<MODE, TN, 4>), grid, dim3(256), 0, s, a);
}

// 8 x 8 transpose of 16-bit values: in[j] holds elements (j, 0..7) two to a word, low half
// first; out[c] receives elements (0..7, c) in the same packing.
__device__ __forceinline__ void Transpose8x8(const uint4* in, uint4 (&out)[8]) {
    uint32_t w[8][4];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        w[j][0] = in[j].x;
        w[j][1] = in[j].y;
        w[j][2] = in[j].z;
        w[j][3] = in[j].w;
    }
    uint32_t o[8][4];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t lo = w[2 * k][m], hi = w[2 * k + 1][m];
            o[2 * m][k] = (lo & 0xffffu) | (hi << 16);
            o[2 * m + 1][k] = (lo >> 16) | (hi & 0xffff0000u);
        }
#pragma unroll
    for (int c = 0; c < 8; ++c) out[c] = make_uint4(o[c][0], o[c][1], o[c][2], o[c][3]);
}

// Weight gradient of a stride-1 convolution (forward or upsample-concat), LDS-tiled:
//   dW[co][tap][ci] = sum_v dZ[v][co] * X[v + off(tap)][ci]
// as MFMAs with M = 16 output channels, N = 16 input channels, K = 32 voxels.  A workgroup
// owns one chunk of 16 input channels and TCO output channels and walks its share of the
// 16 x 4 x 4 voxel tiles (one split).  Per tile it stages the dZ tile transposed (sZ:
// [co][voxel]) and the 18 x 6 x 6 input halo transposed, in three copies shifted by
// kx = 0, 1, 2 along x (sX: [kx][ci][hz][hy][x]), so that the B fragment of any tap — 8
// consecutive x-voxels of one channel — is one aligned 16-byte LDS read.  The dZ fragments
// of the tile's 8 K blocks stay in registers for all 27 taps; wave w accumulates taps w,
// w+4, ... over the whole split and writes them to the split's partial sums.
template <int MODE, int TCO>
__global__ __launch_bounds__(256) void KWgradTile(const WgradArgs w) {
    constexpr int TI = TCO / 16;               // output-channel MFMA tiles
    constexpr int CIS = kHZ * kHY * kTX + 8;   // one channel of one shifted copy (bf16), padded
    constexpr int ZS = kTX * kTY * kTZ + 8;    // one output channel of the dZ tile, padded
    constexpr int kTaps = 7;                   // taps per wave (27 over 4 waves)
    __shared__ __attribute__((aligned(16))) u16 sX[3 * 16 * CIS];
    __shared__ __attribute__((aligned(16))) u16 sZ[TCO * ZS];
    const ConvArgs& a = w.g;
    const int K = 27 * a.kin;
    const int cb = blockIdx.x * 16, co0 = blockIdx.y * TCO;
    const int runs = a.kin - cb >= 16 ? 2 : 1;  // 8-channel runs in this chunk
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, grp = lane >> 4;
    const int tiles_x = (a.wd + kTX - 1) / kTX, tiles_y = (a.hd + kTY - 1) / kTY, tiles_z = (a.dd + kTZ - 1) / kTZ;
    const long long ntiles = static_cast<long long>(a.nb) * tiles_z * tiles_y * tiles_x;
    const long long tps = (ntiles + gridDim.z - 1) / gridDim.z;
    const long long t0 = static_cast<long long>(blockIdx.z) * tps;
    const long long t1 = t0 + tps < ntiles ? t0 + tps : ntiles;
    if (runs == 1)  // channels 8..15 of the chunk stay zero
        for (int e = threadIdx.x; e < 3 * 16 * CIS / 8; e += 256)
            reinterpret_cast<uint4*>(sX)[e] = make_uint4(0u, 0u, 0u, 0u);
    f32x4 acc[kTaps][TI];
#pragma unroll
    for (int tt = 0; tt < kTaps; ++tt)
#pragma unroll
        for (int i = 0; i < TI; ++i) acc[tt][i] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (long long tile = t0; tile < t1; ++tile) {
        long long b = tile;
        const int x0 = static_cast<int>(b % tiles_x) * kTX;
        b /= tiles_x;
        const int y0 = static_cast<int>(b % tiles_y) * kTY;
        b /= tiles_y;
        const int z0 = static_cast<int>(b % tiles_z) * kTZ;
        const int n = static_cast<int>(b / tiles_z);
        // Staging, 16-byte LDS rows: a halo item (hz, hy, half xh of the 16 x positions, an
        // 8-channel run) loads the 10 voxels x0-1+8xh .. +9, which feed 8 consecutive positions
        // of all three x-shifted copies; a dZ item loads 8 consecutive voxels of one 8-channel
        // run.  Each 8 x 8 block (voxels x channels) is transposed in registers and stored as
        // 8 rows of one channel.  Halo and dZ items share one index space.
        const int halo_items = kHZ * kHY * 2 * runs;
        const int z_items = (kTX * kTY * kTZ / 8) * TI * 2;
        for (int e = threadIdx.x; e < halo_items + z_items; e += 256) {
            if (e < halo_items) {
                const int r8 = e % runs;
                int it = e / runs;
                const int xh = it & 1;
                it >>= 1;
                const int hy = it % kHY, hz = it / kHY;
                uint4 raw[10];
#pragma unroll
                for (int j = 0; j < 10; ++j)
                    raw[j] = InputRun<MODE>(a, n, z0 - 1 + hz, y0 - 1 + hy, x0 - 1 + 8 * xh + j, cb + r8 * 8);
                u16* dst = sX + r8 * 8 * CIS + (hz * kHY + hy) * kTX + 8 * xh;
#pragma unroll
                for (int kx = 0; kx < 3; ++kx) {
                    uint4 t[8];
                    Transpose8x8(raw + kx, t);
#pragma unroll
                    for (int c = 0; c < 8; ++c) *reinterpret_cast<uint4*>(dst + (kx * 16 + c) * CIS) = t[c];
                }
            } else {
                const int ez = e - halo_items;
                const int g = ez / (TI * 2), q = ez - g * (TI * 2);
                const int v0 = 8 * g;
                const int xb = x0 + (v0 & (kTX - 1)), y = y0 + (v0 / kTX) % kTY, z = z0 + v0 / (kTX * kTY);
                const bool row_in = y < a.hd && z < a.dd && co0 + q * 8 < w.cout;
                const long long vrow = ((static_cast<long long>(n) * a.dd + z) * a.hd + y) * a.wd;
                uint4 raw[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    raw[j] = make_uint4(0u, 0u, 0u, 0u);
                    if (row_in && xb + j < a.wd)
                        raw[j] = *reinterpret_cast<const uint4*>(w.dz + (vrow + xb + j) * w.cout + co0 + q * 8);
                }
                uint4 t[8];
                Transpose8x8(raw, t);
#pragma unroll
                for (int c = 0; c < 8; ++c) *reinterpret_cast<uint4*>(&sZ[(q * 8 + c) * ZS + v0]) = t[c];
            }
        }
        __syncthreads();
        // K block kb: voxels 32kb..32kb+31 (tile rows 2kb, 2kb+1 of 16 x-voxels); lane
        // group g holds voxels 32kb + 8g .. +7: row 2kb + g/2, x = 8 (g & 1) ..
        bf16x8 af[TI][8];
#pragma unroll
        for (int i = 0; i < TI; ++i)
#pragma unroll
            for (int kb = 0; kb < 8; ++kb)
                __builtin_memcpy(&af[i][kb], &sZ[(i * 16 + row) * ZS + 32 * kb + 8 * grp], 16);
#pragma unroll
        for (int tt = 0; tt < kTaps; ++tt) {
            const int t = wave + 4 * tt;
            if (t >= 27) break;
            const int kz = t / 9, ky = (t / 3) % 3, kx = t % 3;
            const u16* xb = sX + (kx * 16 + row) * CIS + 8 * (grp & 1);
#pragma unroll
            for (int kb = 0; kb < 8; ++kb) {
                const int rv = 2 * kb + (grp >> 1);  // tile row: z = rv / 4, y = rv % 4
                bf16x8 bfr;
                __builtin_memcpy(&bfr, xb + (((rv >> 2) + kz) * kHY + (rv & 3) + ky) * kTX, 16);
#pragma unroll
                for (int i = 0; i < TI; ++i)
                    acc[tt][i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i][kb], bfr, acc[tt][i], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // partial sums of this split: lane holds rows co = 4*grp + r, column ci = row
    float* part = w.part + static_cast<long long>(blockIdx.z) * w.cout * K;
#pragma unroll
    for (int tt = 0; tt < kTaps; ++tt) {
        const int t = wave + 4 * tt;
        if (t >= 27) break;
        if (row >= runs * 8) continue;
#pragma unroll
        for (int i = 0; i < TI; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int co = co0 + i * 16 + 4 * grp + r;
                if (co < w.cout) part[static_cast<long long>(co) * K + t * a.kin + cb + row] = acc[tt][i][r];
            }
    }
}

template <int MODE>
void LaunchConv(const ConvArgs& a, hipStream_t s) {
    const long long V = static_cast<long long>(a.nb) * a.dd * a.hd * a.wd;
    if ((MODE == kUpcat || a.stride == 1) && kConvTiled) {
        const unsigned tiles = static_cast<unsigned>(a.nb) * ((a.dd + kTZ - 1) / kTZ) * ((a.hd + kTY - 1) / kTY) *
                               ((a.wd + kTX - 1) / kTX);
        if (tiles == 0) return;
        bool done = false;
        if constexpr (MODE != kAdj && kConvPersistent) {
            // the persistent kernel for the network's layers of 16 to 64 output channels, in
            // blocks of 16 or 32 on grid.y; the workgroups of one channel block at most one
            // per output tile (16 x 4 x 4*ZW voxels)
            auto persistent = [&](auto kern, int zw, int co) {
                if (a.nout % co != 0) return;  // a whole number of channel blocks only
                const unsigned ptiles = static_cast<unsigned>(a.nb) * ((a.dd + 4 * zw - 1) / (4 * zw)) *
                                        ((a.hd + kTY - 1) / kTY) * ((a.wd + kTX - 1) / kTX);
                const unsigned blocks = Blk(a.nout, co);
                const unsigned g = PersistentGrid(kern, static_cast<long long>(ptiles) * blocks) / blocks;
                hipLaunchKernelGGL(kern, dim3(g > 0 ? g : 1, blocks), dim3(256), 0, s, a);
                done = true;
            };
            // z depth per tile (ZW) as measured fastest per layer at batch 4 x 256x128x128
            // (the forward-mode shapes with 16 to 128 input channels also serve the stride-1
            // adjoints, run as forward convolutions with mirrored weights)
            if (a.nout == 16 && a.kin == 8)
                persistent(KConvTileP<MODE, 1, 2, 1>, 2, 16);
            else if (a.nout == 16 && a.kin == 48)  // two chunks of 24 channels: 3.6% of the K steps
                persistent(KConvTileP<MODE, 6, 2, 1, 3>, 2, 16);  // padded instead of 10% with 16
            else if (MODE == kUpcat && a.nout == 32 && a.kin == 96)
                persistent(KConvTileP<MODE, 12, 2, 2>, 2, 32);
            else if (MODE == kUpcat && a.nout == 64 && a.kin == 192)
                persistent(KConvTileP<MODE, 24, 2, 2>, 2, 32);
            else if (MODE == kFwd && a.nout == 48 && a.kin == 16)
                persistent(KConvTileP<MODE, 2, 1, 3>, 1, 48);
            else if (MODE == kFwd && a.nout % 32 == 0 && a.kin == 32)
                persistent(KConvTileP<MODE, 4, 1, 2>, 1, 32);
            else if (MODE == kFwd && a.nout % 32 == 0 && a.kin == 64)
                persistent(KConvTileP<MODE, 8, 1, 2>, 1, 32);
            else if (MODE == kFwd && a.nout % 32 == 0 && a.kin == 128)
                persistent(KConvTileP<MODE, 16, 1, 2>, 1, 32);
        }
        if (done) {
        } else if (a.nout <= 16)
            LaunchConvTile<MODE, 1>(a, dim3(tiles, 1), s);
        else if (a.nout <= 32)
            LaunchConvTile<MODE, 2>(a, dim3(tiles, 1), s);
        else if (a.nout <= 48)  // the upsample-concat adjoints (48 = 32 + 16 channels)
            LaunchConvTile<MODE, 3>(a, dim3(tiles, 1), s);
        else
            LaunchConvTile<MODE, 4>(a, dim3(tiles, Blk(a.nout, 64)), s);
        return;
    }
    // stride 2: the tiled kernel re-stages the halo for every 8-channel run, which pays for
    // the 16-channel layer (e1: 0.37 -> 0.27 ms at batch 4 x 256x128x128) but not for 32
    // (e2: 0.13 -> 0.16 ms); wider inputs take the direct gather below
    if (MODE == kFwd && a.stride == 2 && a.kin <= 16 && kConvTiled) {
        const unsigned tiles = static_cast<unsigned>(a.nb) * ((a.dd + kTZ - 1) / kTZ) * ((a.hd + kTY - 1) / kTY) *
                               ((a.wd + kTX - 1) / kTX);
        if (tiles == 0) return;
        if (a.nout <= 16)
            hipLaunchKernelGGL((KConvTileS2<1>), dim3(tiles, 1), dim3(256), 0, s, a);
        else if (a.nout <= 32)
            hipLaunchKernelGGL((KConvTileS2<2>), dim3(tiles, 1), dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL((KConvTileS2<4>), dim3(tiles, Blk(a.nout, 64)), dim3(256), 0, s, a);
        return;
    }
    if (MODE == kAdj && a.stride == 2 && a.kin % 32 == 0) {
        // class 0 (all coordinates even) is the largest parity class
        const long long V0 = static_cast<long long>(a.nb) * ((a.dd + 1) / 2) * ((a.hd + 1) / 2) * ((a.wd + 1) / 2);
        if (a.nout <= 16)
            hipLaunchKernelGGL((KConvAdj2<4, 1>), dim3(Blk(V0, 64), 1, 1), dim3(512), 0, s, a);
        else if (a.nout <= 32)
            hipLaunchKernelGGL((KConvAdj2<2, 2>), dim3(Blk(V0, 32), 1, 1), dim3(512), 0, s, a);
        else
            hipLaunchKernelGGL((KConvAdj2<2, 4>), dim3(Blk(V0, 32), 1, Blk(a.nout, 64)), dim3(512), 0, s, a);
        return;
    }
    if (a.nout <= 16) {
        hipLaunchKernelGGL((KConv<MODE, 4, 1>), dim3(Blk(V, 256), 1), dim3(256), 0, s, a);
    } else if (a.nout <= 32) {
        hipLaunchKernelGGL((KConv<MODE, 2, 2>), dim3(Blk(V, 128), 1), dim3(256), 0, s, a);
    } else {
        hipLaunchKernelGGL((KConv<MODE, 2, 4>), dim3(Blk(V, 128), Blk(a.nout, 64)), dim3(256), 0, s, a);
    }
}

}  // namespace

extern "C" {

// Convolution (mode 0 forward, 1 upsample-concat forward, 2 adjoint/data-gradient).
// Shapes: see ConvArgs.  Returns 0 or a HIP error code.
int fc_nn_conv(int mode, const void* in, const void* in2, const void* w, const float* bias, void* out, float* outf,
               int nb, int di, int hi, int wi, int dd, int hd, int wd, int kin, int c1, int nout, int stride, int act,
               void* stream) {
    if (kin % 8 != 0 || (mode == kUpcat && (c1 % 8 != 0 || (kin - c1) % 8 != 0))) return -1;
    ConvArgs a{static_cast<const u16*>(in), static_cast<const u16*>(in2), static_cast<const u16*>(w), bias,
               static_cast<u16*>(out), outf, nb, di, hi, wi, dd, hd, wd, kin, c1, nout, stride, act};
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (mode == kFwd)
        LaunchConv<kFwd>(a, s);
    else if (mode == kUpcat)
        LaunchConv<kUpcat>(a, s);
    else
        LaunchConv<kAdj>(a, s);
    return static_cast<int>(hipGetLastError());
}

// Weight gradient (fp32, [cout][27*kin]) of a mode-0/1 conv into dw (overwrite, or add
// when accumulate); ws must hold splits*cout*27*kin floats.
int fc_nn_wgrad(int mode, const void* in, const void* in2, const void* dz, float* dw, float* ws, int splits,
                int accumulate, int nb, int di, int hi, int wi, int dd, int hd, int wd, int kin, int c1, int cout,
                int stride, void* stream) {
    if (kin % 8 != 0) return -1;
    WgradArgs w{};
    w.g = ConvArgs{static_cast<const u16*>(in), static_cast<const u16*>(in2), nullptr, nullptr, nullptr, nullptr,
                   nb, di, hi, wi, dd, hd, wd, kin, c1, cout, stride, 0};
    w.dz = static_cast<const u16*>(dz);
    w.cout = cout;
    const long long V = static_cast<long long>(nb) * dd * hd * wd;
    long long vper = (V + splits - 1) / splits;
    vper = (vper + 31) / 32 * 32;
    w.vper = vper;
    w.part = ws;
    const int K = 27 * kin;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if ((mode == kUpcat || stride == 1) && (kin == 8 || kin % 16 == 0) && cout % 16 == 0) {
        // stride 1: the LDS-tiled kernel
        auto launch_tile = [&](auto tco) {
            constexpr int TCO = decltype(tco)::value;
            const dim3 grid(Blk(kin, 16), cout / TCO, splits);
            if (mode == kUpcat)
                hipLaunchKernelGGL((KWgradTile<kUpcat, TCO>), grid, dim3(256), 0, s, w);
            else
                hipLaunchKernelGGL((KWgradTile<kFwd, TCO>), grid, dim3(256), 0, s, w);
        };
        if (cout % 32 == 0)
            launch_tile(std::integral_constant<int, 32>{});
        else
            launch_tile(std::integral_constant<int, 16>{});
        const long long n = static_cast<long long>(cout) * K;
        hipLaunchKernelGGL(KReduceSplits, dim3(Blk(n, kRedOut)), dim3(256), 0, s, ws, splits, n, dw, accumulate);
        return static_cast<int>(hipGetLastError());
    }
    auto launch = [&](auto tco) {
        constexpr int TCO = decltype(tco)::value;
        const dim3 grid(Blk(K, 4096 / TCO), Blk(cout, TCO), splits);
        if (mode == kUpcat)
            hipLaunchKernelGGL((KWgrad<kUpcat, TCO>), grid, dim3(256), 0, s, w);
        else
            hipLaunchKernelGGL((KWgrad<kFwd, TCO>), grid, dim3(256), 0, s, w);
    };
    if (cout <= 16)
        launch(std::integral_constant<int, 16>{});
    else if (cout <= 32)
        launch(std::integral_constant<int, 32>{});
    else
        launch(std::integral_constant<int, 64>{});
    const long long n = static_cast<long long>(cout) * K;
    hipLaunchKernelGGL(KReduceSplits, dim3(Blk(n, kRedOut)), dim3(256), 0, s, ws, splits, n, dw, accumulate);
    return static_cast<int>(hipGetLastError());
}

// Per-channel sums (bias gradient) of a [V][C] bf16 tensor into out[C] (fp32); ws holds
// 256*C floats.  C is a multiple of 8 and C/8 divides 256.
int fc_nn_colsum(const void* x, long long V, int C, float* out, float* ws, int accumulate, void* stream) {
    if (C <= 0 || C % 8 != 0 || 256 % (C / 8) != 0) return -1;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(KColSumPartial, dim3(kColBlocks), dim3(256), 0, s, static_cast<const u16*>(x), V, C, ws);
    hipLaunchKernelGGL(KReduceSplits, dim3(Blk(C, kRedOut)), dim3(256), 0, s, ws, kColBlocks, static_cast<long long>(C),
                       out, accumulate);
    return static_cast<int>(hipGetLastError());
}

// Number of per-block partial sums behind fc_nn_colsum and fc_nn_leaky_backward_colsum: their
// workspace holds fc_nn_col_blocks() * C floats.
int fc_nn_col_blocks() { return kColBlocks; }

int fc_nn_leaky_backward_colsum(const void* y, const void* dy, void* dz, long long V, int C, float* out, float* ws,
                                int accumulate, void* stream) {
    if (C <= 0 || C % 8 != 0 || 256 % (C / 8) != 0) return -1;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(KLeakyColSum, dim3(kColBlocks), dim3(256), 0, s, static_cast<const u16*>(y),
                       static_cast<const u16*>(dy), static_cast<u16*>(dz), V, C, ws);
    hipLaunchKernelGGL(KReduceSplits, dim3(Blk(C, kRedOut)), dim3(256), 0, s, ws, kColBlocks, static_cast<long long>(C),
                       out, accumulate);
    return static_cast<int>(hipGetLastError());
}

int fc_nn_leaky_backward(const void* y, const void* dy, void* dz, long long n, void* stream) {
    if (n % 8 == 0)
        hipLaunchKernelGGL(KLeakyBackward8, dim3(Blk(n / 8, 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                           static_cast<const u16*>(y), static_cast<const u16*>(dy), static_cast<u16*>(dz), n / 8);
    else
        hipLaunchKernelGGL(KLeakyBackward, dim3(Blk(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                           static_cast<const u16*>(y), static_cast<const u16*>(dy), static_cast<u16*>(dz), n);
    return static_cast<int>(hipGetLastError());
}

int fc_nn_transpose_w(const void* w, void* wt, int cout, int cin, void* stream) {
    const long long n = static_cast<long long>(cout) * 27 * cin;
    hipLaunchKernelGGL(KTransposeW, dim3(Blk(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const u16*>(w), static_cast<u16*>(wt), cout, cin, 0);
    return static_cast<int>(hipGetLastError());
}

// Wt[ci][26 - t][co] = W[co][t][ci]: a mode-0 (forward, stride 1) convolution of dZ with Wt
// is the stride-1 adjoint, and runs on the forward kernels.
int fc_nn_flip_transpose_w(const void* w, void* wt, int cout, int cin, void* stream) {
    const long long n = static_cast<long long>(cout) * 27 * cin;
    hipLaunchKernelGGL(KTransposeW, dim3(Blk(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const u16*>(w), static_cast<u16*>(wt), cout, cin, 1);
    return static_cast<int>(hipGetLastError());
}

int fc_nn_upcat_split(const void* dcat, int c1, int c2, int nb, int d, int h, int w, void* dskip, void* dcoarse,
                      void* stream) {
    const long long V = static_cast<long long>(nb) * d * h * w;
    if (c1 % 8 == 0 && c2 % 8 == 0) {
        const long long n8 = V * c2 > V / 8 * c1 ? V * c2 / 8 : V / 8 * c1 / 8;
        hipLaunchKernelGGL(KUpcatSplit8, dim3(Blk(n8, 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                           static_cast<const u16*>(dcat), c1, c2, nb, d, h, w, static_cast<u16*>(dskip),
                           static_cast<u16*>(dcoarse));
        return static_cast<int>(hipGetLastError());
    }
    const long long n = V * c2 > V / 8 * c1 ? V * c2 : V / 8 * c1;
    hipLaunchKernelGGL(KUpcatSplit, dim3(Blk(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const u16*>(dcat), c1, c2, nb, d, h, w, static_cast<u16*>(dskip),
                       static_cast<u16*>(dcoarse));
    return static_cast<int>(hipGetLastError());
}

// Head + loss.  Results: dwh[4*cin], dbh[4], loss (fp32, device) via ws (kHeadBlocks*(4cin+5)).
int fc_nn_head_loss(const void* h, int cin, const float* wh, const float* bh, const float* tgt, const float* wmask,
                    long long V, float inv_norm, void* dh, float* ws, float* out_grads_and_loss, float* pred_out,
                    void* stream) {
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (cin == 16)
        hipLaunchKernelGGL(KHeadLoss<16>, dim3(kHeadBlocks), dim3(256), 0, s, static_cast<const u16*>(h), wh, bh, tgt,
                           wmask, V, inv_norm, static_cast<u16*>(dh), ws, pred_out);
    else if (cin == 8)
        hipLaunchKernelGGL(KHeadLoss<8>, dim3(kHeadBlocks), dim3(256), 0, s, static_cast<const u16*>(h), wh, bh, tgt,
                           wmask, V, inv_norm, static_cast<u16*>(dh), ws, pred_out);
    else
        return -1;
    const long long P = 4LL * cin + 5;
    hipLaunchKernelGGL(KReduceSplits, dim3(Blk(P, kRedOut)), dim3(256), 0, s, ws, kHeadBlocks, P, out_grads_and_loss, 0);
    return static_cast<int>(hipGetLastError());
}

int fc_nn_adam(float* p, const float* g, float* m, float* v, void* pbf, long long n, float lr, float b1, float b2,
               float eps, int step, float gscale, void* stream) {
    const float bc1 = 1.f - powf(b1, static_cast<float>(step));
    const float bc2 = 1.f - powf(b2, static_cast<float>(step));
    hipLaunchKernelGGL(KAdam, dim3(Blk(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), p, g, m, v,
                       static_cast<u16*>(pbf), n, lr, b1, b2, eps, bc1, bc2, gscale);
    return static_cast<int>(hipGetLastError());
}

int fc_nn_add_bf16(const void* x, void* y, long long n, void* stream) {
    if (n % 8 == 0 && reinterpret_cast<uintptr_t>(x) % 16 == 0 && reinterpret_cast<uintptr_t>(y) % 16 == 0)
        hipLaunchKernelGGL(KAddBf16x8, dim3(Blk(n / 8, 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                           static_cast<const u16*>(x), static_cast<u16*>(y), n / 8);
    else
        hipLaunchKernelGGL(KAddBf16, dim3(Blk(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                           static_cast<const u16*>(x), static_cast<u16*>(y), n);
    return static_cast<int>(hipGetLastError());
}

int fc_nn_f32_to_bf16(const float* x, void* y, long long n, void* stream) {
    hipLaunchKernelGGL(KF32ToBf16, dim3(Blk(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), x,
                       static_cast<u16*>(y), n);
    return static_cast<int>(hipGetLastError());
}

int fc_nn_bf16_to_f32(const void* x, float* y, long long n, void* stream) {
    hipLaunchKernelGGL(KBf16ToF32, dim3(Blk(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const u16*>(x), y, n);
    return static_cast<int>(hipGetLastError());
}

int fc_nn_features(const uint8_t* solid, const float* sdf, float u_in, int nb, int d, int h, int w, void* out,
                   void* stream) {
    const long long V = static_cast<long long>(nb) * d * h * w;
    hipLaunchKernelGGL(KFeatures, dim3(Blk(V, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), solid, sdf, u_in,
                       nb, d, h, w, static_cast<u16*>(out));
    return static_cast<int>(hipGetLastError());
}

int fc_nn_targets(const float* field, const uint8_t* solid, float u_in, long long V, float* tgt, float* wmask,
                  void* stream) {
    hipLaunchKernelGGL(KTargets, dim3(Blk(V, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), field, solid, u_in,
                       V, tgt, wmask);
    return static_cast<int>(hipGetLastError());
}

int fc_nn_pred_to_field(const float* pred, float u_in, long long V, float* field, void* stream) {
    hipLaunchKernelGGL(KPredToField, dim3(Blk(V, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), pred, u_in, V,
                       field);
    return static_cast<int>(hipGetLastError());
}

}  // extern "C"
