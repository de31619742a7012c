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
// HBM layout.  Populations are SoA planes f[d*P + n] with a one-cell ghost
// shell in y and z only:  n = ((z+1)*(ny+2) + (y+1))*nx + x,
// P = nx*(ny+2)*(nz+2).  Rows stay nx long, so a 64-lane wave reading 64
// consecutive x of one plane is one contiguous 256 B (FP32) / 128 B (16-bit)
// segment.  Ghost cells are never updated; in the in-place layout they hold
// the populations that leave the domain through a y/z wall, which is what
// makes wall reads race-free (see InPlace below).

#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lattice.hpp"

namespace fc {

// Population stores stream through the cache hierarchy: a step writes every
// population once and nothing reads it again within the step.
#ifndef FC_LBM_NT_STORES
#define FC_LBM_NT_STORES 0
#endif
template <class T>
__device__ __forceinline__ void PutPop(T* p, T v) {
#if FC_LBM_NT_STORES
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
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

// -------------------------------------------------------- storage formats
// FP32: the reference's float populations.
template <int Q>
struct StoreF32 {
    using T = float;
    using R = float;
    static __device__ __forceinline__ R Load(T v, int) { return v; }
    static __device__ __forceinline__ T Store(R f, int) { return f; }
};

// FP16 offset: g = (f - w) * 4096 as IEEE binary16 (RNE); collision in FP32.
template <int Q>
struct StoreF16 {
    using T = _Float16;
    using R = float;
    static __device__ __forceinline__ R Load(T v, int d) { return Lat<Q>::w[d] + static_cast<float>(v) * (1.0f / 4096.0f); }
    static __device__ __forceinline__ T Store(R f, int d) { return static_cast<T>((f - Lat<Q>::w[d]) * 4096.0f); }
};

// BF16 offset: g = f - w as bfloat16 (RNE, NaN kept quiet); collision in FP32.
template <int Q>
struct StoreBF16 {
    using T = uint16_t;
    using R = float;
    static __device__ __forceinline__ R Load(T v, int d) {
        return Lat<Q>::w[d] + __uint_as_float(static_cast<uint32_t>(v) << 16);
    }
    static __device__ __forceinline__ T Store(R f, int d) {
        uint32_t b = __float_as_uint(f - Lat<Q>::w[d]);
        if ((b & 0x7FFFFFFFu) > 0x7F800000u) return static_cast<T>((b >> 16) | 0x40u);
        b += 0x7FFFu + ((b >> 16) & 1u);
        return static_cast<T>(b >> 16);
    }
};

// FP64: the whole step in double (GPU twin of the FP64 oracle).
template <int Q>
struct StoreF64 {
    using T = double;
    using R = double;
    static __device__ __forceinline__ R Load(T v, int) { return v; }
    static __device__ __forceinline__ T Store(R f, int) { return f; }
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
    long long P;     // plane stride (cells)
    int nx, ny, nz;
    int z0, zc;
    int parity;      // step index & 1
    int reverse, fsy, fsz, pery, perz;
    float tau, u_in, cs;
    double tau_d, u_in_d, cs_d;
};

struct Geo {
    int nx, ny, nz, pery, perz;
    int sy, sz;  // padded strides: sy = nx, sz = nx*(ny+2)
    __device__ __forceinline__ int Idx(int x, int y, int z) const { return (z + 1) * sz + (y + 1) * sy + x; }
    // Neighbour index with the layout's wrap rules: x always wraps (only the
    // inlet/outlet columns ever reach across it, and they never read a slot);
    // y/z wrap when periodic, otherwise land in the ghost shell.
    __device__ __forceinline__ int Nbr(int x, int y, int z, int dx, int dy, int dz) const {
        int ax = x + dx, ay = y + dy, az = z + dz;
        ax = ax >= nx ? ax - nx : (ax < 0 ? ax + nx : ax);
        if (pery) ay = ay >= ny ? ay - ny : (ay < 0 ? ay + ny : ay);
        if (perz) az = az >= nz ? az - nz : (az < 0 ? az + nz : az);
        return Idx(ax, ay, az);
    }
};

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

// Slot (plane, cell) of population d read (load) / written (store) at a parity.
// `self` is the cell, `anc` the anchor of d's pair.
__device__ __forceinline__ void LoadSlot(int d, int par, int self, int anc, int& plane, int& cell) {
    if (d == 0) {
        plane = 0;
        cell = self;
    } else if (d & 1) {
        plane = par == 0 ? d + 1 : d;
        cell = self;
    } else {
        plane = par == 0 ? d - 1 : d;
        cell = anc;
    }
}

__device__ __forceinline__ void StoreSlot(int d, int par, int self, int anc, int& plane, int& cell) {
    if (d == 0) {
        plane = 0;
        cell = self;
    } else if (d & 1) {
        plane = par == 0 ? d : d + 1;
        cell = anc;
    } else {
        plane = par == 0 ? d : d - 1;
        cell = self;
    }
}

template <int Q>
__device__ __forceinline__ int PairAnchor(int d) {
    return d == 0 ? 0 : ((d & 1) ? d : d - 1);
}

// Current post-collision population d of cell (x,y,z) after `t` steps.
template <int Q, class S, class Lay>
__device__ __forceinline__ typename S::R CurrentPop(const void* pops0, const void* pops1, long long P, long long t,
                                                   const Geo& g, int x, int y, int z, int d) {
    using T = typename S::T;
    const int self = g.Idx(x, y, z);
    if constexpr (!Lay::kInPlace) {
        const T* f = static_cast<const T*>((t & 1) ? pops1 : pops0);
        return S::Load(f[d * P + self], d);
    } else {
        const T* f = static_cast<const T*>(pops0);
        const int a = PairAnchor<Q>(d);
        const int anc = d == 0 ? self : g.Nbr(x, y, z, Lat<Q>::ex[a], Lat<Q>::ey[a], Lat<Q>::ez[a]);
        int pl, c;
        StoreSlot(d, static_cast<int>((t - 1) & 1), self, anc, pl, c);
        return S::Load(f[pl * P + c], d);
    }
}

template <int Q, class R>
__device__ __forceinline__ R Feq(int d, R rho, R ux, R uy, R uz) {
    R eu = static_cast<R>(Lat<Q>::ex[d]) * ux + static_cast<R>(Lat<Q>::ey[d]) * uy + static_cast<R>(Lat<Q>::ez[d]) * uz;
    R u2 = ux * ux + uy * uy + uz * uz;
    return W<Q, R>(d) * rho * (R(1) + R(3) * eu + R(4.5) * eu * eu - R(1.5) * u2);
}

// ----------------------------------------------------------- step kernel
template <int Q, class S, class Lay>
__global__ __launch_bounds__(256) void KStep(const StepArgs a) {
    using T = typename S::T;
    using R = typename S::R;
    constexpr bool IP = Lay::kInPlace;

    const unsigned plane = static_cast<unsigned>(a.nx) * static_cast<unsigned>(a.ny);
    const unsigned lin = blockIdx.x * 256u + threadIdx.x;
    if (lin >= static_cast<unsigned>(a.zc) * plane) return;
    const unsigned zr = lin / plane;
    const unsigned rem = lin - zr * plane;
    const int y = static_cast<int>(rem / static_cast<unsigned>(a.nx));
    const int x = static_cast<int>(rem) - y * a.nx;
    const int z = a.z0 + static_cast<int>(zr);

    const Geo g{a.nx, a.ny, a.nz, a.pery, a.perz, a.nx, a.nx * (a.ny + 2)};
    const int self = g.Idx(x, y, z);
    const long long P = a.P;
    const int par = a.parity;

    T* __restrict__ dst = static_cast<T*>(a.pop_dst);
    const T* src = static_cast<const T*>(a.pop_src);
    R* side_mom = static_cast<R*>(a.side_mom);
    const int plane_yz = a.ny * a.nz;
    const int mom_col = x == 1 ? 0 : 1;
    const bool mom_writer = IP && (x == 1 || x == a.nx - 2);

    const uint8_t fl = a.flags[self];
    R fo[Q];

    if (fl & kSolid) {
        // Solid cell: populations w_i (LbmReference.cpp:264-267).  In place, solid
        // cells are inactive; only the outlet-column moments they feed are written.
#pragma unroll
        for (int d = 0; d < Q; ++d) fo[d] = W<Q, R>(d);
        if constexpr (!IP) {
#pragma unroll
            for (int d = 0; d < Q; ++d) dst[d * P + self] = S::Store(fo[d], d);
        }
        if (mom_writer) {
            R m0 = 0, m1 = 0, m2 = 0, m3 = 0;
#pragma unroll
            for (int d = 0; d < Q; ++d) {
                R v = S::Load(S::Store(fo[d], d), d);
                m0 += v;
                m1 += v * static_cast<R>(Lat<Q>::ex[d]);
                m2 += v * static_cast<R>(Lat<Q>::ey[d]);
                m3 += v * static_cast<R>(Lat<Q>::ez[d]);
            }
            R* m = side_mom + ((static_cast<long long>(par) * 2 + mom_col) * plane_yz + z * a.ny + y) * 4;
            m[0] = m0;
            m[1] = m1;
            m[2] = m2;
            m[3] = m3;
        }
        return;
    }

    const R tau = sizeof(R) == 4 ? static_cast<R>(a.tau) : static_cast<R>(a.tau_d);
    const R uin = sizeof(R) == 4 ? static_cast<R>(a.u_in) : static_cast<R>(a.u_in_d);
    const R cs = sizeof(R) == 4 ? static_cast<R>(a.cs) : static_cast<R>(a.cs_d);
    const int inletX = a.reverse ? a.nx - 1 : 0;
    const int outletX = a.reverse ? 0 : a.nx - 1;
    const bool near = (fl & kNearSolid) != 0;
    BRec<R>* rec = nullptr;

    if (x == inletX) {
        // Equilibrium inlet (LbmReference.cpp:269-273).
        const R ux = a.reverse ? -uin : uin;
#pragma unroll
        for (int d = 0; d < Q; ++d) fo[d] = Feq<Q, R>(d, R(1), ux, R(0), R(0));
    } else if (x == outletX) {
        // Zero-gradient outlet from the neighbour's previous post-collision moments,
        // density relaxed toward 1 (LbmReference.cpp:275-327).
        const int nbx = a.reverse ? x + 1 : x - 1;
        R nr = 0, nu = 0, nv = 0, nw = 0;
        if constexpr (!IP) {
            const int nb = self + (nbx - x);
#pragma unroll
            for (int d = 0; d < Q; ++d) {
                R v = S::Load(src[d * P + nb], d);
                nr += v;
                nu += v * static_cast<R>(Lat<Q>::ex[d]);
                nv += v * static_cast<R>(Lat<Q>::ey[d]);
                nw += v * static_cast<R>(Lat<Q>::ez[d]);
            }
        } else {
            const int col = nbx == 1 ? 0 : 1;
            const R* m = side_mom + ((static_cast<long long>(par ^ 1) * 2 + col) * plane_yz + z * a.ny + y) * 4;
            nr = m[0];
            nu = m[1];
            nv = m[2];
            nw = m[3];
        }
        if (nr > K<R>(1e-6f, 1e-6)) {
            nu /= nr;
            nv /= nr;
            nw /= nr;
        }
        const R rho = nr > K<R>(1e-6f, 1e-6) ? R(1) + K<R>(0.999f, 0.999) * (nr - R(1)) : R(1);
#pragma unroll
        for (int d = 0; d < Q; ++d) fo[d] = Feq<Q, R>(d, rho, nu, nv, nw);
    } else {
        // ---------------------------------------------------------- gather
        R f[Q];
        const bool edge = (y == 0) | (y == a.ny - 1) | (z == 0) | (z == a.nz - 1);
        if (!(edge | near)) {
            // Bulk fast path: every source is an in-domain fluid cell, no wrap.
#pragma unroll
            for (int d = 0; d < Q; ++d) {
                if constexpr (!IP) {
                    const int off = Lat<Q>::ex[d] + Lat<Q>::ey[d] * g.sy + Lat<Q>::ez[d] * g.sz;
                    f[d] = S::Load(src[d * P + (self - off)], d);
                } else {
                    const int pa = PairAnchor<Q>(d);
                    const int anc = self + Lat<Q>::ex[pa] + Lat<Q>::ey[pa] * g.sy + Lat<Q>::ez[pa] * g.sz;
                    int pl, c;
                    LoadSlot(d, par, self, anc, pl, c);
                    f[d] = S::Load(dst[pl * P + c], d);
                }
            }
        } else {
            uint32_t kinds[3] = {0u, 0u, 0u};
            if (near) {
                rec = static_cast<BRec<R>*>(a.recs) + a.bidx[self];
                kinds[0] = rec->kind[0];
                kinds[1] = rec->kind[1];
                kinds[2] = rec->kind[2];
            }
            // Phase 1: ordinary streaming for every direction whose source is an
            // in-domain fluid cell (periodic wrap included).
#pragma unroll
            for (int d = 0; d < Q; ++d) {
                const int sy_ = y - Lat<Q>::ey[d], sz_ = z - Lat<Q>::ez[d];
                const bool outY = !a.pery && (sy_ < 0 || sy_ >= a.ny);
                const bool outZ = !a.perz && (sz_ < 0 || sz_ >= a.nz);
                const unsigned kd = KindOf(kinds, d);
                f[d] = R(0);
                if (d == 0 || (!outY && !outZ && kd == kPull)) {
                    if constexpr (!IP) {
                        const int s = g.Nbr(x, y, z, -Lat<Q>::ex[d], -Lat<Q>::ey[d], -Lat<Q>::ez[d]);
                        f[d] = S::Load(src[d * P + s], d);
                    } else {
                        const int pa = PairAnchor<Q>(d);
                        const int anc = d == 0 ? self : g.Nbr(x, y, z, Lat<Q>::ex[pa], Lat<Q>::ey[pa], Lat<Q>::ez[pa]);
                        int pl, c;
                        LoadSlot(d, par, self, anc, pl, c);
                        f[d] = S::Load(dst[pl * P + c], d);
                    }
                }
            }
            // Phase 2: walls and solid sources, from the cell's own previous
            // outgoing populations (LbmReference.cpp:100-122, 336-381).
            auto own_prev = [&](int d) -> R {
                if constexpr (!IP) {
                    return S::Load(src[d * P + self], d);
                } else {
                    const int pa = PairAnchor<Q>(d);
                    const int anc = d == 0 ? self : g.Nbr(x, y, z, Lat<Q>::ex[pa], Lat<Q>::ey[pa], Lat<Q>::ez[pa]);
                    int pl, c;
                    StoreSlot(d, par ^ 1, self, anc, pl, c);
                    return S::Load(dst[pl * P + c], d);
                }
            };
#pragma unroll
            for (int d = 1; d < Q; ++d) {
                const int sy_ = y - Lat<Q>::ey[d], sz_ = z - Lat<Q>::ez[d];
                const bool outY = !a.pery && (sy_ < 0 || sy_ >= a.ny);
                const bool outZ = !a.perz && (sz_ < 0 || sz_ >= a.nz);
                if (outY || outZ) {
                    const bool ok = (!outY || a.fsy) && (!outZ || a.fsz);
                    if (ok) {
                        if (outY && outZ)
                            f[d] = own_prev(MirrorZ<Q>(MirrorY<Q>(d)));
                        else if (outY)
                            f[d] = own_prev(MirrorY<Q>(d));
                        else
                            f[d] = own_prev(MirrorZ<Q>(d));
                    } else {
                        f[d] = own_prev(Opp(d));
                    }
                    continue;
                }
                const unsigned kd = KindOf(kinds, d);
                if (kd == kPull) continue;
                const R plain = own_prev(Opp(d));
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
                        fi = S::Load(src[d * P + self], d);
                    else
                        fi = rec->side[d];
                    f[d] = inv2q * plain + (R(1) - inv2q) * fi;
                }
            }
        }

        // ------------------------------------------------------- collide
        R rho = 0, ux = 0, uy = 0, uz = 0;
#pragma unroll
        for (int d = 0; d < Q; ++d) {
            rho += f[d];
            ux += f[d] * static_cast<R>(Lat<Q>::ex[d]);
            uy += f[d] * static_cast<R>(Lat<Q>::ey[d]);
            uz += f[d] * static_cast<R>(Lat<Q>::ez[d]);
        }
        if (rho > K<R>(1e-6f, 1e-6)) {
            ux /= rho;
            uy /= rho;
            uz /= rho;
        }
        if (!__builtin_isfinite(rho) || !__builtin_isfinite(ux) || !__builtin_isfinite(uy) ||
            !__builtin_isfinite(uz) || rho <= R(0) || (ux * ux + uy * uy + uz * uz) > R(4))
            atomicOr(a.diverged, 1);

        const R kCs2 = R(1) / R(3);
        R pxx = 0, pyy = 0, pzz = 0, pxy = 0, pxz = 0, pyz = 0;
#pragma unroll
        for (int d = 0; d < Q; ++d) {
            const R fe = Feq<Q, R>(d, rho, ux, uy, uz);
            const R fneq = f[d] - fe;
            f[d] = fe;
            const R cx = static_cast<R>(Lat<Q>::ex[d]), cy = static_cast<R>(Lat<Q>::ey[d]),
                    cz = static_cast<R>(Lat<Q>::ez[d]);
            pxx += cx * cx * fneq;
            pyy += cy * cy * fneq;
            pzz += cz * cz * fneq;
            pxy += cx * cy * fneq;
            pxz += cx * cz * fneq;
            pyz += cy * cz * fneq;
        }
        const R piMag = sqrt(pxx * pxx + pyy * pyy + pzz * pzz + R(2) * (pxy * pxy + pxz * pxz + pyz * pyz));
        const R rmax = rho < K<R>(1e-6f, 1e-6) ? K<R>(1e-6f, 1e-6) : rho;
        const R kSmag = R(18) * K<R>(1.41421356f, 1.41421356);
        const R tauEff = R(0.5) * (tau + sqrt(tau * tau + kSmag * cs * cs * piMag / rmax));
        const R omega = R(1) / tauEff;
#pragma unroll
        for (int d = 0; d < Q; ++d) {
            const R cx = static_cast<R>(Lat<Q>::ex[d]), cy = static_cast<R>(Lat<Q>::ey[d]),
                    cz = static_cast<R>(Lat<Q>::ez[d]);
            const R qxx = cx * cx - kCs2, qyy = cy * cy - kCs2, qzz = cz * cz - kCs2;
            const R qxy = cx * cy, qxz = cx * cz, qyz = cy * cz;
            const R reg = (W<Q, R>(d) * R(4.5)) *
                          (qxx * pxx + qyy * pyy + qzz * pzz + R(2) * qxy * pxy + R(2) * qxz * pxz + R(2) * qyz * pyz);
            fo[d] = f[d] + (R(1) - omega) * reg;
        }
    }

    // ---------------------------------------------------------------- store
    if constexpr (!IP) {
#pragma unroll
        for (int d = 0; d < Q; ++d) PutPop(&dst[d * P + self], S::Store(fo[d], d));
    } else {
        const bool interior = x > 0 && x < a.nx - 1 && y > 0 && y < a.ny - 1 && z > 0 && z < a.nz - 1;
#pragma unroll
        for (int d = 0; d < Q; ++d) {
            const int pa = PairAnchor<Q>(d);
            int anc = self;
            if (d != 0) {
                anc = interior ? self + Lat<Q>::ex[pa] + Lat<Q>::ey[pa] * g.sy + Lat<Q>::ez[pa] * g.sz
                               : g.Nbr(x, y, z, Lat<Q>::ex[pa], Lat<Q>::ey[pa], Lat<Q>::ez[pa]);
            }
            int pl, c;
            StoreSlot(d, par, self, anc, pl, c);
            PutPop(&dst[pl * P + c], S::Store(fo[d], d));
        }
        if (rec != nullptr) {
            uint32_t kinds[3] = {rec->kind[0], rec->kind[1], rec->kind[2]};
#pragma unroll
            for (int d = 1; d < Q; ++d)
                if (KindOf(kinds, d) == kBouzidiHi) rec->side[d] = S::Load(S::Store(fo[d], d), d);
        }
        if (mom_writer) {
            R m0 = 0, m1 = 0, m2 = 0, m3 = 0;
#pragma unroll
            for (int d = 0; d < Q; ++d) {
                R v = S::Load(S::Store(fo[d], d), d);
                m0 += v;
                m1 += v * static_cast<R>(Lat<Q>::ex[d]);
                m2 += v * static_cast<R>(Lat<Q>::ey[d]);
                m3 += v * static_cast<R>(Lat<Q>::ez[d]);
            }
            R* m = side_mom + ((static_cast<long long>(par) * 2 + mom_col) * plane_yz + z * a.ny + y) * 4;
            m[0] = m0;
            m[1] = m1;
            m[2] = m2;
            m[3] = m3;
        }
    }
}

}  // namespace fc
