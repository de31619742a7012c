// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// Clean-room CPU restatement of FlowCraft's LbmReference (the CPU twin of the
// GPU LbmSolver step).  It exists to CHECK the HIP engine in flowcraft_amd/:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// it, and never as the thing measured or shipped.
//
// Templated on the arithmetic type:
//   Real = float  : parity twin of LbmReference (same operations in the same
//                   order, compiled with -ffp-contract=off so no FMA contraction)
//   Real = double : the FP64 oracle every GPU result states its error against.
// For Real=float the population store can also be emulated in the engine's
// reduced-precision formats (FP16 / BF16 offsets from the rest weights), so the
// GPU's reduced-storage kernels have an exact CPU twin as well.
//
// Followed, function by function (paths relative to the FlowCraft repo):
//   lattice tables ............ FlowCraft/LatticeModel.h:20-64
//   Equilibrium ............... FlowCraft/LbmReference.cpp:9-13
//   ParallelForZ .............. FlowCraft/LbmReference.cpp:24-39
//   MacroscopicFrom ........... FlowCraft/LbmReference.cpp:43-61
//   InterpolatedBounceBack .... FlowCraft/LbmReference.cpp:100-122
//   SolidCellForce ............ FlowCraft/LbmReference.cpp:127-148
//   Init (mirror search) ...... FlowCraft/LbmReference.cpp:150-190
//   Reset / ResetFromField .... FlowCraft/LbmReference.cpp:192-239
//   Step ...................... FlowCraft/LbmReference.cpp:241-469
//   ReadbackField ............. FlowCraft/LbmReference.cpp:475-500
//   ComputeStats .............. FlowCraft/LbmReference.cpp:502-585
//   ComputeAerodynamicForces .. FlowCraft/FieldOverlay.cpp:469-621 (on a readback field;
//                               cell centres as VoxelGrid::CellCenter, VoxelGrid.h:49-51)
// Extensions with no reference counterpart (pinned only by the FP64 build and
// the invariants in tests/): periodic y/z wrap, FP16/BF16 offset storage (rounded to
// nearest or stochastically).  The
// offset formats run the engine's offset-space form of the same collision
// (StepOffset below), operation for operation, on the stored offsets
// g = (f - w) * S; their state arrays hold g, and readbacks decode f = w + g / S.
//
// Exposed through a small extern "C" API (orc_*) for ctypes.

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <type_traits>
#include <vector>

namespace {

// ---------------------------------------------------------------- lattices
struct Dir {
    int ex, ey, ez;
    float w;  // the reference keeps weights as float literals (1/3, 1/18, ...)
    double wd;  // exact-as-double weight for the FP64 oracle
    int opp;
};

// LatticeModel.h:20-43 (D3Q19) and :45-64 (D3Q15): same order, same pairing.
const Dir kD3Q19[19] = {
    {0, 0, 0, 1.0f / 3.0f, 1.0 / 3.0, 0},    {1, 0, 0, 1.0f / 18.0f, 1.0 / 18.0, 2},
    {-1, 0, 0, 1.0f / 18.0f, 1.0 / 18.0, 1}, {0, 1, 0, 1.0f / 18.0f, 1.0 / 18.0, 4},
    {0, -1, 0, 1.0f / 18.0f, 1.0 / 18.0, 3}, {0, 0, 1, 1.0f / 18.0f, 1.0 / 18.0, 6},
    {0, 0, -1, 1.0f / 18.0f, 1.0 / 18.0, 5}, {1, 1, 0, 1.0f / 36.0f, 1.0 / 36.0, 8},
    {-1, -1, 0, 1.0f / 36.0f, 1.0 / 36.0, 7}, {1, -1, 0, 1.0f / 36.0f, 1.0 / 36.0, 10},
    {-1, 1, 0, 1.0f / 36.0f, 1.0 / 36.0, 9}, {1, 0, 1, 1.0f / 36.0f, 1.0 / 36.0, 12},
    {-1, 0, -1, 1.0f / 36.0f, 1.0 / 36.0, 11}, {1, 0, -1, 1.0f / 36.0f, 1.0 / 36.0, 14},
    {-1, 0, 1, 1.0f / 36.0f, 1.0 / 36.0, 13}, {0, 1, 1, 1.0f / 36.0f, 1.0 / 36.0, 16},
    {0, -1, -1, 1.0f / 36.0f, 1.0 / 36.0, 15}, {0, 1, -1, 1.0f / 36.0f, 1.0 / 36.0, 18},
    {0, -1, 1, 1.0f / 36.0f, 1.0 / 36.0, 17},
};
const Dir kD3Q15[15] = {
    {0, 0, 0, 2.0f / 9.0f, 2.0 / 9.0, 0},     {1, 0, 0, 1.0f / 9.0f, 1.0 / 9.0, 2},
    {-1, 0, 0, 1.0f / 9.0f, 1.0 / 9.0, 1},    {0, 1, 0, 1.0f / 9.0f, 1.0 / 9.0, 4},
    {0, -1, 0, 1.0f / 9.0f, 1.0 / 9.0, 3},    {0, 0, 1, 1.0f / 9.0f, 1.0 / 9.0, 6},
    {0, 0, -1, 1.0f / 9.0f, 1.0 / 9.0, 5},    {1, 1, 1, 1.0f / 72.0f, 1.0 / 72.0, 8},
    {-1, -1, -1, 1.0f / 72.0f, 1.0 / 72.0, 7}, {1, 1, -1, 1.0f / 72.0f, 1.0 / 72.0, 10},
    {-1, -1, 1, 1.0f / 72.0f, 1.0 / 72.0, 9}, {1, -1, 1, 1.0f / 72.0f, 1.0 / 72.0, 12},
    {-1, 1, -1, 1.0f / 72.0f, 1.0 / 72.0, 11}, {-1, 1, 1, 1.0f / 72.0f, 1.0 / 72.0, 14},
    {1, -1, -1, 1.0f / 72.0f, 1.0 / 72.0, 13},
};

// ------------------------------------------------------ storage emulation
// IEEE binary32 -> binary16, round-to-nearest-even, subnormals kept, overflow
// to infinity (the behaviour of a hardware RNE convert; written here from the
// IEEE 754 definition and checked in tests against numpy.float16).
uint16_t F32ToF16(float v) {
    uint32_t b;
    std::memcpy(&b, &v, 4);
    uint32_t sign = (b >> 16) & 0x8000u;
    uint32_t absb = b & 0x7FFFFFFFu;
    if (absb > 0x7F800000u) return static_cast<uint16_t>(sign | 0x7E00u | ((absb >> 13) & 0x3FFu));  // NaN
    if (absb >= 0x477FF000u) return static_cast<uint16_t>(sign | 0x7C00u);  // rounds to >= 65520 -> inf
    int e = static_cast<int>(absb >> 23) - 127;
    uint32_t m = (absb & 0x7FFFFFu) | 0x800000u;
    if (e >= -14) {
        // normal half: keep 10 mantissa bits
        uint32_t hm = (absb & 0x7FFFFFu) >> 13;
        uint32_t rest = absb & 0x1FFFu;
        uint32_t h = (static_cast<uint32_t>(e + 15) << 10) | hm;
        if (rest > 0x1000u || (rest == 0x1000u && (h & 1u))) h += 1;  // carry may bump the exponent: still correct
        return static_cast<uint16_t>(sign | h);
    }
    if (e < -25) return static_cast<uint16_t>(sign);  // below half the smallest subnormal -> 0
    // subnormal half: value = m * 2^(e-23); half subnormal unit = 2^-24
    int shift = -e - 1;  // number of bits to drop from the 24-bit significand to land on 2^-24 units
    uint32_t hm = m >> shift;
    uint32_t rest = m & ((1u << shift) - 1u);
    uint32_t halfway = 1u << (shift - 1);
    if (rest > halfway || (rest == halfway && (hm & 1u))) hm += 1;
    return static_cast<uint16_t>(sign | hm);
}

// The engine's FP16 storage rule: RNE conversion, except that values tiny in the IEEE
// "tininess after rounding" sense — below 2^-14 once rounded to 11 significant bits with
// an unbounded exponent — become a zero of the same sign (the device's conversion with
// half denormals flushed).
uint16_t F32ToF16Flush(float v) {
    uint32_t b;
    std::memcpy(&b, &v, 4);
    const uint32_t a = b & 0x7FFFFFFFu;
    if (a < 0x38800000u) {
        uint32_t hb = a >> 13;
        const uint32_t rest = a & 0x1FFFu;
        if (rest > 0x1000u || (rest == 0x1000u && (hb & 1u))) ++hb;
        if ((hb << 13) < 0x38800000u) return static_cast<uint16_t>((b >> 16) & 0x8000u);
    }
    return F32ToF16(v);
}

float F16ToF32(uint16_t h) {
    uint32_t sign = static_cast<uint32_t>(h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1Fu;
    uint32_t m = h & 0x3FFu;
    uint32_t b;
    if (e == 0) {
        if (m == 0) {
            b = sign;
        } else {  // subnormal: m * 2^-24, exact in binary32
            float f = static_cast<float>(m) * 5.9604644775390625e-08f;
            std::memcpy(&b, &f, 4);
            b |= sign;
        }
    } else if (e == 31) {
        b = sign | 0x7F800000u | (m << 13);
    } else {
        b = sign | ((e - 15 + 127) << 23) | (m << 13);
    }
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

// binary32 -> bfloat16, round-to-nearest-even (NaN kept quiet).
uint16_t F32ToBF16(float v) {
    uint32_t b;
    std::memcpy(&b, &v, 4);
    if ((b & 0x7FFFFFFFu) > 0x7F800000u) return static_cast<uint16_t>((b >> 16) | 0x40u);
    uint32_t lsb = (b >> 16) & 1u;
    b += 0x7FFFu + lsb;
    return static_cast<uint16_t>(b >> 16);
}

float BF16ToF32(uint16_t h) {
    uint32_t b = static_cast<uint32_t>(h) << 16;
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

// Stochastic rounding of the engine's *_SR formats (flowcraft_amd/csrc/lbm_kernels.hpp:
// SrHash, StoreF16SR, StoreBF16SR), restated from their definition.  The random bits of a
// population are a function of (cell, step, direction): the cell's hash ("lowbias32"
// finaliser of C. Wellons' hash-prospector on cell*0x9E3779B1 ^ step*0x85EBCA77) plus d
// golden-ratio increments.
uint32_t SrHash(uint32_t cell, uint32_t step) {
    uint32_t h = (cell * 0x9E3779B1u) ^ (step * 0x85EBCA77u);
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return h;
}
constexpr uint32_t kSrWeyl = 0x9E3779B9u;

// FP16: the top 13 random bits added to the binary32 pattern, then truncation to binary16
// (round toward zero: overflow gives the largest finite half) with half denormals
// flushed — a magnitude below 2^-14 after the addition becomes a zero of its sign.
uint16_t F32ToF16SR(float v, uint32_t r) {
    uint32_t b;
    std::memcpy(&b, &v, 4);
    b += r >> 19;
    const uint32_t sign = (b >> 16) & 0x8000u;
    const uint32_t a = b & 0x7FFFFFFFu;
    if (a > 0x7F800000u) return static_cast<uint16_t>(sign | 0x7E00u | ((a >> 13) & 0x3FFu));  // NaN
    if (a == 0x7F800000u) return static_cast<uint16_t>(sign | 0x7C00u);
    if (a >= 0x47800000u) return static_cast<uint16_t>(sign | 0x7BFFu);
    if (a < 0x38800000u) return static_cast<uint16_t>(sign);
    return static_cast<uint16_t>(sign | ((a - 0x38000000u) >> 13));
}

// BF16: the top 16 random bits added to the binary32 pattern, high half kept.
uint16_t F32ToBF16SR(float v, uint32_t r) {
    uint32_t b;
    std::memcpy(&b, &v, 4);
    return static_cast<uint16_t>((b + (r >> 16)) >> 16);
}

// Storage formats of the engine: 0 = native Real, 1 = FP16 offset, 2 = BF16 offset,
// 3 = FP16 offset and 4 = BF16 offset with stochastically rounded step stores.
// Offset formats hold g = (f - w) * S and decode f = w + g * (1/S).
constexpr float kF16Scale = 4096.0f;
constexpr float kF16InvScale = 1.0f / 4096.0f;

// A decimal constant of the reference as the float literal (Real=float, so the
// value is the one the reference's `...f` literal rounds to) or as a double.
template <typename R>
constexpr R K(float f, double d) {
    return std::is_same<R, float>::value ? static_cast<R>(f) : static_cast<R>(d);
}

template <typename Real>
struct Oracle {
    int nx = 0, ny = 0, nz = 0, Q = 0;
    size_t N = 0;
    int storage = 0;
    int nthreads = 1;
    bool periodicY = false, periodicZ = false;
    // Params (LbmReference.h:41-52)
    float tau = 0.8f, inletVelocity = 0.05f, cs = 0.12f;
    bool reverseFlow = false, freeSlipY = true, freeSlipZ = true;

    std::vector<int> ex, ey, ez, opp, mirrorY, mirrorZ;
    std::vector<Real> w;
    std::vector<float> wf;
    std::vector<uint8_t> solid;
    std::vector<float> sdf;  // empty = no SDF
    std::vector<Real> f[2];
    int cur = 0;
    uint32_t stepIndex = 0;  // steps since the last reset (the SR formats' random stream)
    std::atomic<bool> diverged{false};

    size_t Cell(int x, int y, int z) const { return (static_cast<size_t>(z) * ny + y) * nx + x; }
    size_t Pop(int d, size_t c) const { return static_cast<size_t>(d) * N + c; }

    // Bounds test with the periodic extension: y/z wrap in place when periodic.
    bool Resolve(int& x, int& y, int& z) const {
        if (x < 0 || x >= nx) return false;
        if (y < 0 || y >= ny) {
            if (!periodicY) return false;
            y = (y + ny) % ny;
        }
        if (z < 0 || z >= nz) {
            if (!periodicZ) return false;
            z = (z + nz) % nz;
        }
        return true;
    }

    void Init(int nx_, int ny_, int nz_, int q) {
        nx = nx_;
        ny = ny_;
        nz = nz_;
        N = static_cast<size_t>(nx) * ny * nz;
        const Dir* t = q == 15 ? kD3Q15 : kD3Q19;
        Q = q == 15 ? 15 : 19;
        ex.resize(Q);
        ey.resize(Q);
        ez.resize(Q);
        opp.resize(Q);
        w.resize(Q);
        wf.resize(Q);
        for (int i = 0; i < Q; ++i) {
            ex[i] = t[i].ex;
            ey[i] = t[i].ey;
            ez[i] = t[i].ez;
            opp[i] = t[i].opp;
            wf[i] = t[i].w;
            w[i] = sizeof(Real) == 4 ? static_cast<Real>(t[i].w) : static_cast<Real>(t[i].wd);
        }
        // Mirror tables found by search (LbmReference.cpp:172-183).
        mirrorY.assign(Q, 0);
        mirrorZ.assign(Q, 0);
        for (int i = 0; i < Q; ++i)
            for (int j = 0; j < Q; ++j) {
                if (ex[j] == ex[i] && ey[j] == -ey[i] && ez[j] == ez[i]) mirrorY[i] = j;
                if (ex[j] == ex[i] && ey[j] == ey[i] && ez[j] == -ez[i]) mirrorZ[i] = j;
            }
        solid.assign(N, 0);
        sdf.clear();
        f[0].assign(static_cast<size_t>(Q) * N, Real(0));
        f[1].assign(static_cast<size_t>(Q) * N, Real(0));
        cur = 0;
    }

    template <typename Body>
    void ParallelForZ(Body body) const {
        int nt = std::max(1, std::min(nthreads, nz));
        int per = (nz + nt - 1) / nt;
        std::vector<std::thread> th;
        for (int t = 0; t < nt; ++t) {
            int z0 = t * per, z1 = std::min(z0 + per, nz);
            if (z0 >= z1) break;
            th.emplace_back(body, z0, z1, t);
        }
        for (auto& x : th) x.join();
    }

    // Offset formats: the stored element decoded to g (exact), and f = w + g / S.
    bool IsF16() const { return storage == 1 || storage == 3; }
    bool IsSR() const { return storage >= 3; }
    float Scale() const { return IsF16() ? kF16Scale : 1.0f; }
    // Round to nearest (seeding, restore, and the step stores of formats 1 and 2).  FP16:
    // offsets below the smallest normal half (2^-14) are stored as 0.
    float EncDec(float g) const {
        if (IsF16()) return F16ToF32(F32ToF16Flush(g));
        return BF16ToF32(F32ToBF16(g));
    }
    // The step stores of the stochastically rounded formats, with random bits r.
    float EncDecSR(float g, uint32_t r) const {
        if (IsF16()) return F16ToF32(F32ToF16SR(g, r));
        return BF16ToF32(F32ToBF16SR(g, r));
    }
    Real ToF(Real v, int d) const {
        if (storage == 0) return v;
        return IsF16() ? static_cast<Real>(wf[d] + static_cast<float>(v) * kF16InvScale)
                       : static_cast<Real>(wf[d] + static_cast<float>(v));
    }
    // f -> stored state value (the engine's Store: g = (f - w) * S, encoded).
    Real FromF(Real v, int d) const {
        if (storage == 0) return v;
        float vf = static_cast<float>(v);
        return static_cast<Real>(IsF16() ? EncDec((vf - wf[d]) * kF16Scale) : EncDec(vf - wf[d]));
    }
    // The state as populations f (decoded for the offset formats).
    const std::vector<Real>& FState(std::vector<Real>& tmp) const {
        if (storage == 0) return f[cur];
        tmp.resize(f[cur].size());
        for (int d = 0; d < Q; ++d)
            for (size_t c = 0; c < N; ++c) tmp[Pop(d, c)] = ToF(f[cur][Pop(d, c)], d);
        return tmp;
    }

    // ---- offset-space collision (engine: lbm_kernels.hpp ReconstructOffset/CollideOffset)
    // Per opposite pair (d, d+1), e_{d+1} = -e_d: go_d = A + B, go_{d+1} = A - B with
    // A = W drho - 1.5 W tr T + 4.5 W sum_a e_a^2 T_aa + 9 W sum_{a<b} e_a e_b T_ab and
    // B = 3 W e.j, W = w_d S; the constant factors are rounded to float once.
    void ReconstructOffset(float drho, float jx, float jy, float jz, float txx, float tyy, float tzz, float txy,
                           float txz, float tyz, float* go) const {
        const float tr = (txx + tyy) + tzz;
        {
            const float W = wf[0] * Scale();
            const float W15 = 1.5f * W;
            go[0] = W * drho - W15 * tr;
        }
        const float tdiag[3] = {txx, tyy, tzz};
        const float j[3] = {jx, jy, jz};
        for (int d = 1; d < Q; d += 2) {
            const int c[3] = {ex[d], ey[d], ez[d]};
            const float W = wf[d] * Scale();
            const float W15 = 1.5f * W, W3 = 3.0f * W, W45 = 4.5f * W, W9 = 9.0f * W;
            float a = W * drho - W15 * tr;
            for (int k = 0; k < 3; ++k)
                if (c[k] != 0) a = a + W45 * tdiag[k];
            const int sxy = c[0] * c[1], sxz = c[0] * c[2], syz = c[1] * c[2];
            if (sxy != 0) a = sxy > 0 ? a + W9 * txy : a - W9 * txy;
            if (sxz != 0) a = sxz > 0 ? a + W9 * txz : a - W9 * txz;
            if (syz != 0) a = syz > 0 ? a + W9 * tyz : a - W9 * tyz;
            bool have = false;
            float b = 0.0f;
            for (int k = 0; k < 3; ++k)
                if (c[k] != 0) {
                    const float term = c[k] > 0 ? W3 * j[k] : -(W3 * j[k]);
                    b = have ? b + term : term;
                    have = true;
                }
            go[d] = a + b;
            go[d + 1] = a - b;
        }
    }

    void EqOffset(float drho, float ux, float uy, float uz, float* go) const {
        const float rho = 1.0f + drho;
        const float jx = rho * ux, jy = rho * uy, jz = rho * uz;
        ReconstructOffset(drho, jx, jy, jz, jx * ux, jy * uy, jz * uz, jx * uy, jx * uz, jy * uz, go);
    }

    // Sums of stored values in index order, zero-coefficient terms skipped.
    void RawSums(const float* g, float& sg, float& Jx, float& Jy, float& Jz) const {
        sg = Jx = Jy = Jz = 0.0f;
        for (int d = 0; d < Q; ++d) {
            sg += g[d];
            if (ex[d] == 1) Jx += g[d];
            if (ex[d] == -1) Jx -= g[d];
            if (ey[d] == 1) Jy += g[d];
            if (ey[d] == -1) Jy -= g[d];
            if (ez[d] == 1) Jz += g[d];
            if (ez[d] == -1) Jz -= g[d];
        }
    }

    bool CollideOffset(const float* g, float* go) const {
        const float invS = 1.0f / Scale();
        // Moments from the pair sums ps = g_d + g_{d+1} and differences pd = g_d - g_{d+1},
        // accumulated in pair order from -0 (the additive identity).
        float sg = g[0], Jx = -0.0f, Jy = -0.0f, Jz = -0.0f;
        float gxx = -0.0f, gyy = -0.0f, gzz = -0.0f, gxy = -0.0f, gxz = -0.0f, gyz = -0.0f;
        for (int d = 1; d < Q; d += 2) {
            const float ps = g[d] + g[d + 1];
            const float pd = g[d] - g[d + 1];
            sg += ps;
            if (ex[d] == 1) Jx += pd;
            if (ex[d] == -1) Jx -= pd;
            if (ey[d] == 1) Jy += pd;
            if (ey[d] == -1) Jy -= pd;
            if (ez[d] == 1) Jz += pd;
            if (ez[d] == -1) Jz -= pd;
            if (ex[d] != 0) gxx += ps;
            if (ey[d] != 0) gyy += ps;
            if (ez[d] != 0) gzz += ps;
            const int sxy = ex[d] * ey[d], sxz = ex[d] * ez[d], syz = ey[d] * ez[d];
            if (sxy == 1) gxy += ps;
            if (sxy == -1) gxy -= ps;
            if (sxz == 1) gxz += ps;
            if (sxz == -1) gxz -= ps;
            if (syz == 1) gyz += ps;
            if (syz == -1) gyz -= ps;
        }
        const float drho = sg * invS;
        const float jx = Jx * invS, jy = Jy * invS, jz = Jz * invS;
        const float rho = 1.0f + drho;
        const float r = rho > 1e-6f ? 1.0f / rho : 1.0f;
        const float ux = jx * r, uy = jy * r, uz = jz * r;
        const bool bad = !std::isfinite(rho) || !std::isfinite(ux) || !std::isfinite(uy) || !std::isfinite(uz) ||
                         rho <= 0.0f || (ux * ux + uy * uy + uz * uz) > 4.0f;
        const float c3 = 1.0f / 3.0f;
        const float pxx = (gxx * invS - jx * ux) - drho * c3;
        const float pyy = (gyy * invS - jy * uy) - drho * c3;
        const float pzz = (gzz * invS - jz * uz) - drho * c3;
        const float pxy = gxy * invS - jx * uy;
        const float pxz = gxz * invS - jx * uz;
        const float pyz = gyz * invS - jy * uz;
        const float piMag = std::sqrt(pxx * pxx + pyy * pyy + pzz * pzz + 2.0f * (pxy * pxy + pxz * pxz + pyz * pyz));
        const float rmax = rho < 1e-6f ? 1e-6f : rho;
        const float kSmag = 18.0f * 1.41421356f;
        const float ftau = tau, fcs = cs;
        const float tauEff = 0.5f * (ftau + std::sqrt(ftau * ftau + kSmag * fcs * fcs * piMag / rmax));
        const float keep = 1.0f - 1.0f / tauEff;
        ReconstructOffset(drho, jx, jy, jz, jx * ux + keep * pxx, jy * uy + keep * pyy, jz * uz + keep * pzz,
                          jx * uy + keep * pxy, jx * uz + keep * pxz, jy * uz + keep * pyz, go);
        return bad;
    }

    // LbmReference.cpp:9-13, same association: w*rho*(((1+3eu)+(4.5eu)eu)-1.5u2)
    Real Feq(int i, Real rho, Real ux, Real uy, Real uz) const {
        Real eu = static_cast<Real>(ex[i]) * ux + static_cast<Real>(ey[i]) * uy + static_cast<Real>(ez[i]) * uz;
        Real u2 = ux * ux + uy * uy + uz * uz;
        return w[i] * rho * (Real(1) + Real(3) * eu + Real(4.5) * eu * eu - Real(1.5) * u2);
    }

    // LbmReference.cpp:43-61
    void Moments(const std::vector<Real>& fv, size_t c, Real& rho, Real& ux, Real& uy, Real& uz) const {
        rho = ux = uy = uz = Real(0);
        for (int i = 0; i < Q; ++i) {
            Real fi = fv[Pop(i, c)];
            rho += fi;
            ux += fi * static_cast<Real>(ex[i]);
            uy += fi * static_cast<Real>(ey[i]);
            uz += fi * static_cast<Real>(ez[i]);
        }
        if (rho > K<Real>(1e-6f, 1e-6)) {
            ux /= rho;
            uy /= rho;
            uz /= rho;
        }
    }

    static Real Clamp01(Real v) { return v < Real(0) ? Real(0) : (Real(1) < v ? Real(1) : v); }

    // LbmReference.cpp:100-122.  (x,y,z) fluid, (sx,sy,sz) its solid direction-i source.
    Real Bouzidi(const std::vector<Real>& fv, int x, int y, int z, size_t c, int i, int sx, int sy, int sz) const {
        int o = opp[i];
        Real plain = fv[Pop(o, c)];
        if (sdf.empty()) return plain;
        float phiF = sdf[c];
        float phiS = sdf[Cell(sx, sy, sz)];
        if (!std::isfinite(phiF) || !std::isfinite(phiS)) return plain;
        Real denom = static_cast<Real>(phiF) - static_cast<Real>(phiS);
        if (!(denom > K<Real>(1e-4f, 1e-4))) return plain;
        Real q = Clamp01(static_cast<Real>(phiF) / denom);
        if (q < Real(0.5)) {
            int fx = x + ex[i], fy = y + ey[i], fz = z + ez[i];
            if (!Resolve(fx, fy, fz) || solid[Cell(fx, fy, fz)]) return plain;
            return Real(2) * q * plain + (Real(1) - Real(2) * q) * fv[Pop(o, Cell(fx, fy, fz))];
        }
        Real inv2q = Real(1) / (Real(2) * q);
        return inv2q * plain + (Real(1) - inv2q) * fv[Pop(i, c)];
    }

    // LbmReference.cpp:127-148 (force of the CURRENT state).
    void SolidForce(const std::vector<Real>& fv, int x, int y, int z, Real& stress, Real& fx, Real& fy,
                    Real& fz) const {
        fx = fy = fz = Real(0);
        Real wsum = Real(0);
        for (int i = 1; i < Q; ++i) {
            int nxf = x + ex[i], nyf = y + ey[i], nzf = z + ez[i];
            if (!Resolve(nxf, nyf, nzf) || solid[Cell(nxf, nyf, nzf)]) continue;
            size_t fc = Cell(nxf, nyf, nzf);
            Real plain = fv[Pop(opp[i], fc)];
            Real fb = Bouzidi(fv, nxf, nyf, nzf, fc, i, x, y, z);
            Real dp = (plain + fb) - Real(2) * w[i];
            fx -= static_cast<Real>(ex[i]) * dp;
            fy -= static_cast<Real>(ey[i]) * dp;
            fz -= static_cast<Real>(ez[i]) * dp;
            wsum += w[i];
        }
        Real mag = std::sqrt(fx * fx + fy * fy + fz * fz);
        stress = wsum > Real(0) ? mag / (Real(6) * wsum) : Real(0);
    }

    void SeedValues(Real rho, Real ux, Real uy, Real uz, Real* out) const {
        if (storage == 0) {
            for (int i = 0; i < Q; ++i) out[i] = Feq(i, rho, ux, uy, uz);
            return;
        }
        float go[19];
        EqOffset(static_cast<float>(rho) - 1.0f, static_cast<float>(ux), static_cast<float>(uy),
                 static_cast<float>(uz), go);
        for (int i = 0; i < Q; ++i) out[i] = static_cast<Real>(EncDec(go[i]));
    }

    void FillUniform(Real rho, Real ux, Real uy, Real uz) {
        std::vector<Real> fe(Q);
        SeedValues(rho, ux, uy, uz, fe.data());
        for (int i = 0; i < Q; ++i) std::fill(f[0].begin() + Pop(i, 0), f[0].begin() + Pop(i, 0) + N, fe[i]);
        cur = 0;
        stepIndex = 0;
        diverged = false;
    }

    // LbmReference.cpp:192-217
    void Reset(bool mid) {
        Real rho = Real(1), ux = Real(0);
        if (mid) {
            Real u = static_cast<Real>(inletVelocity);
            Real midSpeed = Real(0.5) * u;
            ux = reverseFlow ? -midSpeed : midSpeed;
            Real dev = Real(0.5) * (u * u - midSpeed * midSpeed);
            rho = Real(1) + Real(3) * dev;
        }
        FillUniform(rho, ux, Real(0), Real(0));
    }

    // LbmReference.cpp:219-239
    void ResetFromField(const float* field) {
        diverged = false;
        ParallelForZ([&](int z0, int z1, int) {
            for (size_t c = static_cast<size_t>(z0) * nx * ny; c < static_cast<size_t>(z1) * nx * ny; ++c) {
                Real rho = solid[c] ? Real(1) : static_cast<Real>(field[c * 4 + 0]);
                Real ux = solid[c] ? Real(0) : static_cast<Real>(field[c * 4 + 1]);
                Real uy = solid[c] ? Real(0) : static_cast<Real>(field[c * 4 + 2]);
                Real uz = solid[c] ? Real(0) : static_cast<Real>(field[c * 4 + 3]);
                Real v[19];
                SeedValues(rho, ux, uy, uz, v);
                for (int i = 0; i < Q; ++i) f[0][Pop(i, c)] = v[i];
            }
        });
        cur = 0;
        stepIndex = 0;
    }

    // LbmReference.cpp:241-469
    void Step() {
        if (storage != 0) {
            StepOffset();
            ++stepIndex;
            return;
        }
        ++stepIndex;
        const std::vector<Real>& fin = f[cur];
        std::vector<Real>& fout = f[1 - cur];
        const int inletX = reverseFlow ? nx - 1 : 0;
        const int outletX = reverseFlow ? 0 : nx - 1;
        const Real inletUx = static_cast<Real>(reverseFlow ? -inletVelocity : inletVelocity);
        const Real rtau = static_cast<Real>(tau);
        const Real rcs = static_cast<Real>(cs);
        const Real kCs2 = Real(1) / Real(3);
        const Real kSmag = Real(18) * K<Real>(1.41421356f, 1.41421356);
        ParallelForZ([&](int z0, int z1, int) {
            bool localDiv = false;
            Real fpre[19];
            for (int z = z0; z < z1; ++z)
                for (int y = 0; y < ny; ++y)
                    for (int x = 0; x < nx; ++x) {
                        size_t c = Cell(x, y, z);
                        if (solid[c]) {
                            for (int i = 0; i < Q; ++i) fout[Pop(i, c)] = w[i];
                            continue;
                        }
                        if (x == inletX) {
                            for (int i = 0; i < Q; ++i)
                                fout[Pop(i, c)] = Feq(i, Real(1), inletUx, Real(0), Real(0));
                            continue;
                        }
                        if (x == outletX) {
                            int nbx = reverseFlow ? x + 1 : x - 1;
                            Real nr, nu, nv, nw;
                            Moments(fin, Cell(nbx, y, z), nr, nu, nv, nw);
                            Real rho = nr > K<Real>(1e-6f, 1e-6) ? Real(1) + K<Real>(0.999f, 0.999) * (nr - Real(1))
                                                                 : Real(1);
                            for (int i = 0; i < Q; ++i) fout[Pop(i, c)] = Feq(i, rho, nu, nv, nw);
                            continue;
                        }
                        for (int i = 0; i < Q; ++i) {
                            int sx = x - ex[i], sy = y - ey[i], sz = z - ez[i];
                            int rx = sx, ry = sy, rz = sz;
                            bool inb = Resolve(rx, ry, rz);
                            if (inb && !solid[Cell(rx, ry, rz)]) {
                                fpre[i] = fin[Pop(i, Cell(rx, ry, rz))];
                            } else if (inb) {
                                fpre[i] = Bouzidi(fin, x, y, z, c, i, rx, ry, rz);
                            } else {
                                // Only a non-periodic axis is a wall (with the periodic extension
                                // a crossing of a periodic axis has already been wrapped).
                                bool outY = !periodicY && (sy < 0 || sy >= ny);
                                bool outZ = !periodicZ && (sz < 0 || sz >= nz);
                                bool ok = (!outY || freeSlipY) && (!outZ || freeSlipZ);
                                if ((outY || outZ) && ok) {
                                    int d = i;
                                    if (outY) d = mirrorY[d];
                                    if (outZ) d = mirrorZ[d];
                                    fpre[i] = fin[Pop(d, c)];
                                } else {
                                    fpre[i] = fin[Pop(opp[i], c)];
                                }
                            }
                        }
                        Real rho = 0, ux = 0, uy = 0, uz = 0;
                        for (int i = 0; i < Q; ++i) {
                            rho += fpre[i];
                            ux += fpre[i] * static_cast<Real>(ex[i]);
                            uy += fpre[i] * static_cast<Real>(ey[i]);
                            uz += fpre[i] * static_cast<Real>(ez[i]);
                        }
                        if (rho > K<Real>(1e-6f, 1e-6)) {
                            ux /= rho;
                            uy /= rho;
                            uz /= rho;
                        }
                        if (!std::isfinite(rho) || !std::isfinite(ux) || !std::isfinite(uy) || !std::isfinite(uz) ||
                            rho <= Real(0) || (ux * ux + uy * uy + uz * uz) > Real(4))
                            localDiv = true;
                        Real pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0;
                        for (int i = 0; i < Q; ++i) {
                            Real feq = Feq(i, rho, ux, uy, uz);
                            Real fneq = fpre[i] - feq;
                            fpre[i] = feq;
                            Real fx = static_cast<Real>(ex[i]), fy = static_cast<Real>(ey[i]),
                                 fz = static_cast<Real>(ez[i]);
                            pxx += fx * fx * fneq;
                            pyy += fy * fy * fneq;
                            pzz += fz * fz * fneq;
                            pxy += fx * fy * fneq;
                            pxz += fx * fz * fneq;
                            pyz += fy * fz * fneq;
                        }
                        Real piMag =
                            std::sqrt(pxx * pxx + pyy * pyy + pzz * pzz + Real(2) * (pxy * pxy + pxz * pxz + pyz * pyz));
                        Real rmax = rho < K<Real>(1e-6f, 1e-6) ? K<Real>(1e-6f, 1e-6) : rho;  // std::max(rho, 1e-6f)
                        Real tauEff = Real(0.5) * (rtau + std::sqrt(rtau * rtau + kSmag * rcs * rcs * piMag / rmax));
                        Real omega = Real(1) / tauEff;
                        for (int i = 0; i < Q; ++i) {
                            Real fx = static_cast<Real>(ex[i]), fy = static_cast<Real>(ey[i]),
                                 fz = static_cast<Real>(ez[i]);
                            Real qxx = fx * fx - kCs2, qyy = fy * fy - kCs2, qzz = fz * fz - kCs2;
                            Real qxy = fx * fy, qxz = fx * fz, qyz = fy * fz;
                            Real reg = (w[i] * Real(4.5)) * (qxx * pxx + qyy * pyy + qzz * pzz + Real(2) * qxy * pxy +
                                                             Real(2) * qxz * pxz + Real(2) * qyz * pyz);
                            fout[Pop(i, c)] = fpre[i] + (Real(1) - omega) * reg;
                        }
                    }
            if (localDiv) diverged = true;
        });
        cur = 1 - cur;
    }

    // The offset formats' step: the same gather and boundary rules as Step() applied to
    // the stored offsets g, then CollideOffset / EqOffset, then encode-decode.
    void StepOffset() {
        const std::vector<Real>& fin = f[cur];
        std::vector<Real>& fout = f[1 - cur];
        const int inletX = reverseFlow ? nx - 1 : 0;
        const int outletX = reverseFlow ? 0 : nx - 1;
        const float inletUx = reverseFlow ? -inletVelocity : inletVelocity;
        const float invS = 1.0f / Scale();
        ParallelForZ([&](int z0, int z1, int) {
            bool localDiv = false;
            float g[19], go[19];
            for (int z = z0; z < z1; ++z)
                for (int y = 0; y < ny; ++y)
                    for (int x = 0; x < nx; ++x) {
                        size_t c = Cell(x, y, z);
                        if (solid[c]) {
                            for (int i = 0; i < Q; ++i) fout[Pop(i, c)] = Real(0);
                            continue;
                        }
                        if (x == inletX) {
                            EqOffset(0.0f, inletUx, 0.0f, 0.0f, go);
                        } else if (x == outletX) {
                            const int nbx = reverseFlow ? x + 1 : x - 1;
                            const size_t nb = Cell(nbx, y, z);
                            for (int i = 0; i < Q; ++i) g[i] = static_cast<float>(fin[Pop(i, nb)]);
                            float sg, Jx, Jy, Jz;
                            RawSums(g, sg, Jx, Jy, Jz);
                            const float dn = sg * invS;
                            const float rn = 1.0f + dn;
                            float ux = Jx * invS, uy = Jy * invS, uz = Jz * invS, drho = 0.0f;
                            if (rn > 1e-6f) {
                                const float r = 1.0f / rn;
                                ux *= r;
                                uy *= r;
                                uz *= r;
                                drho = 0.999f * dn;
                            }
                            EqOffset(drho, ux, uy, uz, go);
                        } else {
                            for (int i = 0; i < Q; ++i) {
                                int sx = x - ex[i], sy = y - ey[i], sz = z - ez[i];
                                int rx = sx, ry = sy, rz = sz;
                                bool inb = Resolve(rx, ry, rz);
                                if (inb && !solid[Cell(rx, ry, rz)]) {
                                    g[i] = static_cast<float>(fin[Pop(i, Cell(rx, ry, rz))]);
                                } else if (inb) {
                                    g[i] = static_cast<float>(Bouzidi(fin, x, y, z, c, i, rx, ry, rz));
                                } else {
                                    bool outY = !periodicY && (sy < 0 || sy >= ny);
                                    bool outZ = !periodicZ && (sz < 0 || sz >= nz);
                                    bool ok = (!outY || freeSlipY) && (!outZ || freeSlipZ);
                                    if ((outY || outZ) && ok) {
                                        int d = i;
                                        if (outY) d = mirrorY[d];
                                        if (outZ) d = mirrorZ[d];
                                        g[i] = static_cast<float>(fin[Pop(d, c)]);
                                    } else {
                                        g[i] = static_cast<float>(fin[Pop(opp[i], c)]);
                                    }
                                }
                            }
                            if (CollideOffset(g, go)) localDiv = true;
                        }
                        if (IsSR()) {
                            const uint32_t h = SrHash(static_cast<uint32_t>(c), stepIndex);
                            for (int i = 0; i < Q; ++i)
                                fout[Pop(i, c)] =
                                    static_cast<Real>(EncDecSR(go[i], h + static_cast<uint32_t>(i) * kSrWeyl));
                        } else {
                            for (int i = 0; i < Q; ++i) fout[Pop(i, c)] = static_cast<Real>(EncDec(go[i]));
                        }
                    }
            if (localDiv) diverged = true;
        });
        cur = 1 - cur;
    }

    // LbmReference.cpp:475-500
    template <typename Out>
    void ReadbackField(Out* out) const {
        std::vector<Real> tmp;
        const std::vector<Real>& fc = FState(tmp);
        ParallelForZ([&](int z0, int z1, int) {
            for (int z = z0; z < z1; ++z)
                for (int y = 0; y < ny; ++y)
                    for (int x = 0; x < nx; ++x) {
                        size_t c = Cell(x, y, z);
                        Real a, b, d, e;
                        if (solid[c])
                            SolidForce(fc, x, y, z, a, b, d, e);
                        else
                            Moments(fc, c, a, b, d, e);
                        out[c * 4 + 0] = static_cast<Out>(a);
                        out[c * 4 + 1] = static_cast<Out>(b);
                        out[c * 4 + 2] = static_cast<Out>(d);
                        out[c * 4 + 3] = static_cast<Out>(e);
                    }
        });
    }

    // LbmReference.cpp:502-585.  out[11]: avgDensity, minDensity, maxDensity, avgSpeed,
    // maxSpeed, forceX, forceY, forceZ, maxStress, totalMass, hasNaN.
    // Partial sums are kept per z-layer and merged in layer order, so the result does
    // not depend on the thread count.
    void Stats(double* out) const {
        struct Part {
            Real minD = Real(1e30), maxD = Real(-1e30), maxS = 0, maxStress = 0;
            double dsum = 0, ssum = 0, mass = 0, fx = 0, fy = 0, fz = 0;
            size_t n = 0;
            bool nan = false;
        };
        std::vector<Part> parts(nz);
        std::vector<Real> tmp;
        const std::vector<Real>& fc = FState(tmp);
        ParallelForZ([&](int z0, int z1, int) {
            for (int z = z0; z < z1; ++z) {
                Part& p = parts[z];
                for (int y = 0; y < ny; ++y)
                    for (int x = 0; x < nx; ++x) {
                        size_t c = Cell(x, y, z);
                        if (solid[c]) {
                            Real st, fx, fy, fz;
                            SolidForce(fc, x, y, z, st, fx, fy, fz);
                            if (std::isfinite(st) && std::isfinite(fx) && std::isfinite(fy) && std::isfinite(fz)) {
                                p.fx += static_cast<double>(fx);
                                p.fy += static_cast<double>(fy);
                                p.fz += static_cast<double>(fz);
                                p.maxStress = std::max(p.maxStress, st);
                            }
                            continue;
                        }
                        Real rho, ux, uy, uz;
                        Moments(fc, c, rho, ux, uy, uz);
                        if (!std::isfinite(rho) || !std::isfinite(ux) || !std::isfinite(uy) || !std::isfinite(uz)) {
                            p.nan = true;
                            continue;
                        }
                        Real sp = std::sqrt(ux * ux + uy * uy + uz * uz);
                        p.minD = std::min(p.minD, rho);
                        p.maxD = std::max(p.maxD, rho);
                        p.maxS = std::max(p.maxS, sp);
                        p.dsum += static_cast<double>(rho);
                        p.ssum += static_cast<double>(sp);
                        p.mass += static_cast<double>(rho);
                        ++p.n;
                    }
            }
        });
        double minD = 1e30, maxD = -1e30, maxS = 0, maxStress = 0, dsum = 0, ssum = 0, mass = 0, fx = 0, fy = 0,
               fz = 0;
        size_t n = 0;
        bool nan = false;
        for (const Part& p : parts) {
            nan = nan || p.nan;
            maxStress = std::max(maxStress, static_cast<double>(p.maxStress));
            fx += p.fx;
            fy += p.fy;
            fz += p.fz;
            if (p.n == 0) continue;
            minD = std::min(minD, static_cast<double>(p.minD));
            maxD = std::max(maxD, static_cast<double>(p.maxD));
            maxS = std::max(maxS, static_cast<double>(p.maxS));
            dsum += p.dsum;
            ssum += p.ssum;
            mass += p.mass;
            n += p.n;
        }
        out[0] = n ? dsum / static_cast<double>(n) : 0.0;
        out[1] = n ? minD : 0.0;
        out[2] = n ? maxD : 0.0;
        out[3] = n ? ssum / static_cast<double>(n) : 0.0;
        out[4] = maxS;
        out[5] = fx;
        out[6] = fy;
        out[7] = fz;
        out[8] = maxStress;
        out[9] = mass;
        out[10] = nan ? 1.0 : 0.0;
    }
};

struct Handle {
    int bits = 32;
    Oracle<float>* f32 = nullptr;
    Oracle<double>* f64 = nullptr;
};

template <typename Fn>
void Dispatch(Handle* h, Fn fn) {
    if (h->bits == 64)
        fn(*h->f64);
    else
        fn(*h->f32);
}

}  // namespace

// FieldOverlay::ComputeAerodynamicForces on a 4-floats/cell field (solid cells hold
// (stress, Fx, Fy, Fz)), in one thread, cells in index order.  out[11] = Fx, Fy, Fz,
// minStress, maxStress, centre of pressure (3), lift-arrow base (3).
static void AeroForces(int nx, int ny, int nz, float cs, const float* origin, const uint8_t* solid, const float* field,
                float* out, int* has_nan) {
    for (int i = 0; i < 11; ++i) out[i] = 0.0f;
    *has_nan = 0;
    const size_t n = static_cast<size_t>(nx) * ny * nz;
    if (n == 0) return;
    double F[3] = {0, 0, 0}, lw = 0, lwp[3] = {0, 0, 0}, fw = 0, fwp[3] = {0, 0, 0}, gp[3] = {0, 0, 0};
    size_t count = 0;
    int top_any = -1;
    float lo = 1e30f, hi = -1e30f;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const size_t c = (static_cast<size_t>(z) * ny + y) * nx + x;
                if (!solid[c]) continue;
                // world position: grid-y along X, grid-x along Y, grid-z along Z
                const float p[3] = {origin[0] + (y + 0.5f) * cs, origin[1] + (x + 0.5f) * cs,
                                    origin[2] + (z + 0.5f) * cs};
                for (int k = 0; k < 3; ++k) gp[k] += p[k];
                ++count;
                top_any = std::max(top_any, z);
                const float* v = field + c * 4;
                if (!std::isfinite(v[0]) || !std::isfinite(v[1]) || !std::isfinite(v[2]) || !std::isfinite(v[3])) {
                    *has_nan = 1;
                    continue;
                }
                for (int k = 0; k < 3; ++k) F[k] += v[1 + k];
                const double wl = std::fabs(static_cast<double>(v[3]));
                lw += wl;
                for (int k = 0; k < 3; ++k) lwp[k] += wl * p[k];
                const double wf = std::sqrt(static_cast<double>(v[1]) * v[1] + static_cast<double>(v[2]) * v[2] +
                                            static_cast<double>(v[3]) * v[3]);
                fw += wf;
                for (int k = 0; k < 3; ++k) fwp[k] += wf * p[k];
                if (v[0] > 0.0f) {
                    lo = std::min(lo, v[0]);
                    hi = std::max(hi, v[0]);
                }
            }
    // stress colour range anchored at zero
    if (lo > hi) {
        lo = 0.0f;
        hi = 1e-6f;
    } else {
        lo = 0.0f;
        if (!(hi > lo)) hi = lo + 1e-6f;
    }
    float cp[3] = {0.0f, 0.0f, 0.0f};
    if (lw > 1e-9) {
        for (int k = 0; k < 3; ++k) cp[k] = static_cast<float>(lwp[k] / lw);
    } else if (fw > 1e-9) {
        for (int k = 0; k < 3; ++k) cp[k] = static_cast<float>(fwp[k] / fw);
    } else if (count > 0) {
        const double inv = 1.0 / static_cast<double>(count);
        for (int k = 0; k < 3; ++k) cp[k] = static_cast<float>(gp[k] * inv);
    }
    // top solid cell of the 5x5 grid columns under the centre of pressure
    const int cx = std::clamp(static_cast<int>(std::floor((cp[1] - origin[1]) / cs)), 0, std::max(nx - 1, 0));
    const int cy = std::clamp(static_cast<int>(std::floor((cp[0] - origin[0]) / cs)), 0, std::max(ny - 1, 0));
    int top = -1;
    for (int dy = -2; dy <= 2; ++dy)
        for (int dx = -2; dx <= 2; ++dx) {
            const int gx = std::clamp(cx + dx, 0, std::max(nx - 1, 0));
            const int gy = std::clamp(cy + dy, 0, std::max(ny - 1, 0));
            for (int z = nz - 1; z >= 0; --z)
                if (solid[(static_cast<size_t>(z) * ny + gy) * nx + gx]) {
                    top = std::max(top, z);
                    break;
                }
        }
    if (top < 0) top = std::max(top_any, 0);
    for (int k = 0; k < 3; ++k) out[k] = static_cast<float>(F[k]);
    out[3] = lo;
    out[4] = hi;
    for (int k = 0; k < 3; ++k) out[5 + k] = cp[k];
    out[8] = cp[0];
    out[9] = cp[1];
    out[10] = origin[2] + (static_cast<float>(top) + 1.0f) * cs;
}

extern "C" {

void orc_aero_forces(int nx, int ny, int nz, float cs, const float* origin, const uint8_t* solid, const float* field,
                     float* out, int* has_nan) {
    AeroForces(nx, ny, nz, cs, origin, solid, field, out, has_nan);
}

// real_bits: 32 or 64.  storage: 0 native, 1 FP16 offset, 2 BF16 offset, 3/4 the same with
// stochastically rounded step stores (offset formats: 32-bit only).
void* orc_create(int nx, int ny, int nz, int q, int real_bits, int storage, int periodic_y, int periodic_z,
                 int nthreads) {
    if (nx < 3 || ny < 1 || nz < 1 || (q != 15 && q != 19)) return nullptr;
    if (storage < 0 || storage > 4 || (real_bits == 64 && storage != 0)) return nullptr;
    Handle* h = new Handle();
    h->bits = real_bits == 64 ? 64 : 32;
    if (h->bits == 64)
        h->f64 = new Oracle<double>();
    else
        h->f32 = new Oracle<float>();
    int nt = nthreads > 0 ? nthreads : static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
    Dispatch(h, [&](auto& o) {
        o.Init(nx, ny, nz, q);
        o.storage = storage;
        o.periodicY = periodic_y != 0;
        o.periodicZ = periodic_z != 0;
        o.nthreads = nt;
    });
    return h;
}

void orc_destroy(void* p) {
    Handle* h = static_cast<Handle*>(p);
    if (!h) return;
    delete h->f32;
    delete h->f64;
    delete h;
}

void orc_set_geometry(void* p, const uint8_t* solid, const float* sdf) {
    Dispatch(static_cast<Handle*>(p), [&](auto& o) {
        o.solid.assign(solid, solid + o.N);
        if (sdf)
            o.sdf.assign(sdf, sdf + o.N);
        else
            o.sdf.clear();
    });
}

void orc_set_params(void* p, float tau, float u, int reverse, int fsy, int fsz, float cs) {
    Dispatch(static_cast<Handle*>(p), [&](auto& o) {
        o.tau = tau;
        o.inletVelocity = u;
        o.reverseFlow = reverse != 0;
        o.freeSlipY = fsy != 0;
        o.freeSlipZ = fsz != 0;
        o.cs = cs;
    });
}

void orc_reset(void* p, int mid) {
    Dispatch(static_cast<Handle*>(p), [&](auto& o) { o.Reset(mid != 0); });
}

void orc_reset_from_field(void* p, const float* field) {
    Dispatch(static_cast<Handle*>(p), [&](auto& o) { o.ResetFromField(field); });
}

void orc_step(void* p, int n) {
    Dispatch(static_cast<Handle*>(p), [&](auto& o) {
        for (int i = 0; i < n; ++i) o.Step();
    });
}

void orc_readback_field_f32(void* p, float* out) {
    Dispatch(static_cast<Handle*>(p), [&](auto& o) { o.ReadbackField(out); });
}

void orc_readback_field_f64(void* p, double* out) {
    Dispatch(static_cast<Handle*>(p), [&](auto& o) { o.ReadbackField(out); });
}

void orc_stats(void* p, double* out) {
    Dispatch(static_cast<Handle*>(p), [&](auto& o) { o.Stats(out); });
}

int orc_has_diverged(void* p) {
    int r = 0;
    Dispatch(static_cast<Handle*>(p), [&](auto& o) { r = o.diverged ? 1 : 0; });
    return r;
}

// Current post-collision populations f[d*N + cell] as double.
void orc_get_pops(void* p, double* out) {
    Dispatch(static_cast<Handle*>(p), [&](auto& o) {
        using R = typename std::remove_reference<decltype(o.f[0][0])>::type;
        std::vector<R> tmp;
        const auto& fc = o.FState(tmp);
        for (size_t k = 0; k < fc.size(); ++k) out[k] = static_cast<double>(fc[k]);
    });
}

// Overwrite the current populations (f[d*N + cell]); values pass through the storage format.
void orc_set_pops(void* p, const double* in) {
    Dispatch(static_cast<Handle*>(p), [&](auto& o) {
        auto& fc = o.f[o.cur];
        for (int d = 0; d < o.Q; ++d)
            for (size_t c = 0; c < o.N; ++c) {
                using R = typename std::remove_reference<decltype(fc[0])>::type;
                fc[o.Pop(d, c)] = o.FromF(static_cast<R>(in[o.Pop(d, c)]), d);
            }
    });
}

// Storage-format helpers, exported so tests can pin them against numpy.
uint16_t orc_f32_to_f16(float v) { return F32ToF16(v); }
uint16_t orc_f32_to_f16_flush(float v) { return F32ToF16Flush(v); }
float orc_f16_to_f32(uint16_t h) { return F16ToF32(h); }
uint16_t orc_f32_to_bf16(float v) { return F32ToBF16(v); }
float orc_bf16_to_f32(uint16_t h) { return BF16ToF32(h); }
void orc_f32_to_f16_array(const float* in, uint16_t* out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = F32ToF16(in[i]);
}
void orc_f32_to_bf16_array(const float* in, uint16_t* out, size_t n) {
    for (size_t i = 0; i < n; ++i) out[i] = F32ToBF16(in[i]);
}
float orc_f16_scale() { return kF16Scale; }
uint32_t orc_sr_hash(uint32_t cell, uint32_t step) { return SrHash(cell, step); }
uint16_t orc_f32_to_f16_sr(float v, uint32_t r) { return F32ToF16SR(v, r); }
uint16_t orc_f32_to_bf16_sr(float v, uint32_t r) { return F32ToBF16SR(v, r); }

// Times `steps` oracle steps after `warmup` untimed ones; returns seconds.
double orc_time_steps(void* p, int warmup, int steps) {
    orc_step(p, warmup);
    auto t0 = std::chrono::steady_clock::now();
    orc_step(p, steps);
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
