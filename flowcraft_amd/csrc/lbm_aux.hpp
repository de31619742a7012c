// Device kernels around the step: geometry preprocessing, seeding, the macro
// field, momentum-exchange forces, statistics and state dump/restore.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lbm_kernels.hpp"

namespace fc {

// Decodes an unpadded global cell index (z*ny + y)*nx + x.
__device__ __forceinline__ void Coords(long long c, int nx, int ny, int& x, int& y, int& z) {
    const long long plane = static_cast<long long>(nx) * ny;
    z = static_cast<int>(c / plane);
    const long long r = c - static_cast<long long>(z) * plane;
    y = static_cast<int>(r / nx);
    x = static_cast<int>(r - static_cast<long long>(y) * nx);
}

// Decodes a padded cell index ((z+1)*(ny+2) + (y+1))*nx + x, the index of the per-cell
// arrays (flags, boundary-record index) and of the cell lists.  64-bit: a lattice may hold
// more than 2^31 padded cells.
__device__ __forceinline__ void PaddedCoords(long long n, int nx, int ny, int& x, int& y, int& z) {
    const long long layer = static_cast<long long>(nx) * (ny + 2);
    const long long zp = n / layer;
    const long long r = n - zp * layer;
    const long long yp = r / nx;
    x = static_cast<int>(r - yp * nx);
    y = static_cast<int>(yp) - 1;
    z = static_cast<int>(zp) - 1;
}

__device__ __forceinline__ long long PaddedIndex(int x, int y, int z, int nx, int ny) {
    return (static_cast<long long>(z + 1) * (ny + 2) + (y + 1)) * nx + x;
}

// In-domain test with the periodic wrap (applied in place).
__device__ __forceinline__ bool Resolve(int& x, int& y, int& z, int nx, int ny, int nz, int pery, int perz) {
    if (x < 0 || x >= nx) return false;
    if (y < 0 || y >= ny) {
        if (!pery) return false;
        y = y < 0 ? y + ny : y - ny;
    }
    if (z < 0 || z >= nz) {
        if (!perz) return false;
        z = z < 0 ? z + nz : z - nz;
    }
    return true;
}

// ------------------------------------------------------------ geometry
// Cell flags: solid / near-solid fluid / surface solid.  Ghost cells stay 0.
template <int Q>
__global__ void KClassify(const uint8_t* __restrict__ solid, uint8_t* __restrict__ flags, int nx, int ny, int nz,
                          int pery, int perz) {
    const long long N = static_cast<long long>(nx) * ny * nz;
    const long long c = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= N) return;
    int x, y, z;
    Coords(c, nx, ny, x, y, z);
    const long long n = PaddedIndex(x, y, z, nx, ny);
    uint8_t fl = 0;
    if (solid[c]) {
        fl = kSolid;
        for (int d = 1; d < Q; ++d) {
            int a = x + Lat<Q>::ex[d], b = y + Lat<Q>::ey[d], e = z + Lat<Q>::ez[d];
            if (Resolve(a, b, e, nx, ny, nz, pery, perz) &&
                !solid[(static_cast<long long>(e) * ny + b) * nx + a]) {
                fl |= kSurface;
                break;
            }
        }
    } else {
        for (int d = 1; d < Q; ++d) {
            int a = x - Lat<Q>::ex[d], b = y - Lat<Q>::ey[d], e = z - Lat<Q>::ez[d];
            if (Resolve(a, b, e, nx, ny, nz, pery, perz) && solid[(static_cast<long long>(e) * ny + b) * nx + a]) {
                fl = kNearSolid;
                break;
            }
        }
    }
    flags[n] = fl;
}

// Deterministic stream compaction of the cells whose flag has `mask`, in
// increasing global cell order: per-chunk counts, a host-side exclusive scan,
// then an in-order scatter.  Output entries are padded indices.
constexpr int kChunkIters = 16;
constexpr int kChunk = 256 * kChunkIters;

__device__ __forceinline__ long long PaddedOf(long long c, int nx, int ny) {
    int x, y, z;
    Coords(c, nx, ny, x, y, z);
    return PaddedIndex(x, y, z, nx, ny);
}

__global__ void KCountFlags(const uint8_t* __restrict__ flags, uint8_t mask, int nx, int ny, int nz,
                            int* __restrict__ counts) {
    const long long N = static_cast<long long>(nx) * ny * nz;
    const long long base = static_cast<long long>(blockIdx.x) * kChunk;
    int cnt = 0;
    for (int it = 0; it < kChunkIters; ++it) {
        long long c = base + it * 256 + threadIdx.x;
        cnt += (c < N && (flags[PaddedOf(c, nx, ny)] & mask)) ? 1 : 0;
    }
    __shared__ int red[256];
    red[threadIdx.x] = cnt;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) counts[blockIdx.x] = red[0];
}

__global__ void KScatterFlags(const uint8_t* __restrict__ flags, uint8_t mask, int nx, int ny, int nz,
                              const long long* __restrict__ offsets, long long* __restrict__ out) {
    const long long N = static_cast<long long>(nx) * ny * nz;
    const long long base = static_cast<long long>(blockIdx.x) * kChunk;
    __shared__ int wave_tot[4];
    long long running = offsets[blockIdx.x];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int it = 0; it < kChunkIters; ++it) {
        long long c = base + it * 256 + threadIdx.x;
        long long n = 0;
        bool p = false;
        if (c < N) {
            n = PaddedOf(c, nx, ny);
            p = (flags[n] & mask) != 0;
        }
        const unsigned long long bal = __ballot(p);
        const int before = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wave_tot[wave] = __popcll(bal);
        __syncthreads();
        int woff = 0;
        for (int w = 0; w < wave; ++w) woff += wave_tot[w];
        if (p) out[running + woff + before] = n;
        running += wave_tot[0] + wave_tot[1] + wave_tot[2] + wave_tot[3];
        __syncthreads();
    }
}

// Boundary records for the near-solid cells: link kind and Bouzidi q per
// direction, computed with the reference's float operations
// (LbmReference.cpp:104-118): valid iff both phi finite and phi_f - phi_s > 1e-4,
// q = clamp(phi_f / (phi_f - phi_s), 0, 1); q < 1/2 needs x + e_i in-domain fluid.
template <int Q, class R>
__global__ void KBuildRecords(const uint8_t* __restrict__ solid, const float* __restrict__ sdf,
                              const long long* __restrict__ list, int count, int* __restrict__ bidx,
                              BRec<R>* __restrict__ recs, int nx, int ny, int nz, int pery, int perz) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const long long n = list[k];
    int x, yy, zz;
    PaddedCoords(n, nx, ny, x, yy, zz);
    const long long c = (static_cast<long long>(zz) * ny + yy) * nx + x;
    bidx[n] = k;
    BRec<R> r;
    r.kind[0] = r.kind[1] = r.kind[2] = 0u;
    r.pad = 0u;
    for (int d = 0; d < 19; ++d) {
        r.q[d] = R(0);
        r.side[d] = R(0);
    }
    for (int d = 1; d < Q; ++d) {
        int sx = x - Lat<Q>::ex[d], sy = yy - Lat<Q>::ey[d], sz = zz - Lat<Q>::ez[d];
        if (!Resolve(sx, sy, sz, nx, ny, nz, pery, perz)) continue;
        const long long s = (static_cast<long long>(sz) * ny + sy) * nx + sx;
        if (!solid[s]) continue;
        unsigned kind = kPlainBB;
        R q = R(0);
        if (sdf != nullptr) {
            const float phiF = sdf[c], phiS = sdf[s];
            if (__builtin_isfinite(phiF) && __builtin_isfinite(phiS)) {
                const R denom = static_cast<R>(phiF) - static_cast<R>(phiS);
                if (denom > K<R>(1e-4f, 1e-4)) {
                    R v = static_cast<R>(phiF) / denom;
                    q = v < R(0) ? R(0) : (R(1) < v ? R(1) : v);
                    if (q < R(0.5)) {
                        int fx = x + Lat<Q>::ex[d], fy = yy + Lat<Q>::ey[d], fz = zz + Lat<Q>::ez[d];
                        if (Resolve(fx, fy, fz, nx, ny, nz, pery, perz) &&
                            !solid[(static_cast<long long>(fz) * ny + fy) * nx + fx])
                            kind = kBouzidiLo;
                    } else {
                        kind = kBouzidiHi;
                    }
                }
            }
        }
        r.kind[d >> 3] |= kind << ((d & 7) * 4);
        r.q[d] = q;
    }
    recs[k] = r;
}

// ------------------------------------------------------------- seeding
// Writes the equilibrium of (rho,u) per cell as the "current" state after t
// steps: ping-pong into copy t&1, in place into the store slots of parity (t-1)&1.
// field4 == nullptr -> the uniform (rho0,u0); solid cells take (1,0,0,0) when
// seeding from a field (LbmReference.cpp:219-239).
template <int Q, class S, class Lay>
__global__ void KSeed(void* pops0, void* pops1, long long P, long long t, int nx, int ny, int nz, int pery, int perz,
                      const uint8_t* __restrict__ solid, const float* __restrict__ field4, double rho0, double u0) {
    using T = typename S::T;
    using R = typename S::R;
    const long long N = static_cast<long long>(nx) * ny * nz;
    const long long c = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= N) return;
    int x, y, z;
    Coords(c, nx, ny, x, y, z);
    const Geo g{nx, ny, nz, pery, perz, nx, nx * (ny + 2), Q, PopRowStride(nx, Q)};
    R rho, ux, uy, uz;
    if (field4 == nullptr) {
        rho = static_cast<R>(rho0);
        ux = static_cast<R>(u0);
        uy = R(0);
        uz = R(0);
    } else if (solid[c]) {
        rho = R(1);
        ux = uy = uz = R(0);
    } else {
        rho = static_cast<R>(field4[c * 4 + 0]);
        ux = static_cast<R>(field4[c * 4 + 1]);
        uy = static_cast<R>(field4[c * 4 + 2]);
        uz = static_cast<R>(field4[c * 4 + 3]);
    }
    R feq[Q];
    if constexpr (S::kOffset)
        EqOffset<Q, S>(rho - 1.0f, ux, uy, uz, feq);
    else
        FeqAll<Q, R>(rho, ux, uy, uz, feq);
    T* f = static_cast<T*>(const_cast<void*>(CurrentCopy<Lay>(pops0, pops1, t)));
#pragma unroll
    for (int d = 0; d < Q; ++d) f[CurrentIndex<Q, Lay>(g, t, x, y, z, d)] = S::Enc(feq[d]);
}

// Exact state restore: values f[d*N + cell] in the storage's compute type.
template <int Q, class S, class Lay>
__global__ void KSetPops(void* pops0, void* pops1, long long P, long long t, int nx, int ny, int nz, int pery,
                         int perz, const typename S::R* __restrict__ in) {
    using T = typename S::T;
    const long long N = static_cast<long long>(nx) * ny * nz;
    const long long c = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= N) return;
    int x, y, z;
    Coords(c, nx, ny, x, y, z);
    const Geo g{nx, ny, nz, pery, perz, nx, nx * (ny + 2), Q, PopRowStride(nx, Q)};
    T* f = static_cast<T*>(const_cast<void*>(CurrentCopy<Lay>(pops0, pops1, t)));
#pragma unroll
    for (int d = 0; d < Q; ++d) f[CurrentIndex<Q, Lay>(g, t, x, y, z, d)] = S::Store(in[d * N + c], d);
}

// State dump.  In place, solid cells are never updated after the first step;
// their reference value (w_i) is reported instead of the stale slots.
template <int Q, class S, class Lay>
__global__ void KGetPops(const void* pops0, const void* pops1, long long P, long long t, int nx, int ny, int nz,
                         int pery, int perz, const uint8_t* __restrict__ solid, typename S::R* __restrict__ out) {
    using R = typename S::R;
    const long long N = static_cast<long long>(nx) * ny * nz;
    const long long c = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= N) return;
    int x, y, z;
    Coords(c, nx, ny, x, y, z);
    const Geo g{nx, ny, nz, pery, perz, nx, nx * (ny + 2), Q, PopRowStride(nx, Q)};
    const bool stale = Lay::kInPlace && t > 0 && solid[c];
#pragma unroll
    for (int d = 0; d < Q; ++d)
        out[d * N + c] = stale ? W<Q, R>(d) : CurrentPop<Q, S, Lay>(pops0, pops1, t, g, x, y, z, d);
}

// In place: rebuild the side values (q >= 1/2 links, outlet-column moments)
// from the current state after a seed or a restore.
template <int Q, class S, class Lay>
__global__ void KRefreshSide(const void* pops0, const void* pops1, long long P, long long t, int nx, int ny, int nz,
                             int pery, int perz, const long long* __restrict__ list, int count, BRec<typename S::R>* recs,
                             typename S::R* side_mom) {
    using R = typename S::R;
    const Geo g{nx, ny, nz, pery, perz, nx, nx * (ny + 2), Q, PopRowStride(nx, Q)};
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int ncol = 2 * ny * nz;
    if (k < count) {
        int x, y, z;
        PaddedCoords(list[k], nx, ny, x, y, z);
        BRec<R>& r = recs[k];
        for (int d = 1; d < Q; ++d)
            if (KindOf(r.kind, d) == kBouzidiHi) r.side[d] = CurrentRaw<Q, S, Lay>(pops0, pops1, t, g, x, y, z, d);
    } else if (k < count + ncol) {
        const int j = k - count;
        const int col = j / (ny * nz);
        const int yz = j - col * ny * nz;
        const int z = yz / ny, y = yz - z * ny;
        const int x = col == 0 ? 1 : nx - 2;
        R m0 = 0, m1 = 0, m2 = 0, m3 = 0;
        for (int d = 0; d < Q; ++d) {
            R v = CurrentRaw<Q, S, Lay>(pops0, pops1, t, g, x, y, z, d);
            m0 += v;
            m1 += v * static_cast<R>(Lat<Q>::ex[d]);
            m2 += v * static_cast<R>(Lat<Q>::ey[d]);
            m3 += v * static_cast<R>(Lat<Q>::ez[d]);
        }
        R* m = side_mom + ((static_cast<long long>((t - 1) & 1) * 2 + col) * ny * nz + z * ny + y) * 4;
        m[0] = m0;
        m[1] = m1;
        m[2] = m2;
        m[3] = m3;
    }
}

// ------------------------------------------------------------ observables
// Momentum-exchange force of the CURRENT populations on each surface solid cell
// (LbmReference.cpp:127-148): out[k] = (stress, Fx, Fy, Fz) in the compute type.
template <int Q, class S, class Lay>
__global__ void KSurfaceForce(const void* pops0, const void* pops1, long long P, long long t, int nx, int ny, int nz,
                              int pery, int perz, const uint8_t* __restrict__ flags, const int* __restrict__ bidx,
                              const BRec<typename S::R>* __restrict__ recs, const long long* __restrict__ surf, int nsurf,
                              typename S::R* __restrict__ out) {
    using R = typename S::R;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nsurf) return;
    const Geo g{nx, ny, nz, pery, perz, nx, nx * (ny + 2), Q, PopRowStride(nx, Q)};
    int x, y, z;
    PaddedCoords(surf[k], nx, ny, x, y, z);
    R fx = 0, fy = 0, fz = 0, wsum = 0;
    for (int d = 1; d < Q; ++d) {
        int a = x + Lat<Q>::ex[d], b = y + Lat<Q>::ey[d], e = z + Lat<Q>::ez[d];
        if (!Resolve(a, b, e, nx, ny, nz, pery, perz)) continue;
        const long long fn = g.Idx(a, b, e);
        if (flags[fn] & kSolid) continue;
        const int o = Opp(d);
        const R plain = CurrentPop<Q, S, Lay>(pops0, pops1, t, g, a, b, e, o);
        R fb = plain;
        const BRec<R>& r = recs[bidx[fn]];
        const unsigned kd = KindOf(r.kind, d);
        if (kd == kBouzidiLo) {
            int ffx = a + Lat<Q>::ex[d], ffy = b + Lat<Q>::ey[d], ffz = e + Lat<Q>::ez[d];
            Resolve(ffx, ffy, ffz, nx, ny, nz, pery, perz);
            const R q = r.q[d];
            fb = R(2) * q * plain + (R(1) - R(2) * q) * CurrentPop<Q, S, Lay>(pops0, pops1, t, g, ffx, ffy, ffz, o);
        } else if (kd == kBouzidiHi) {
            const R q = r.q[d];
            const R inv2q = R(1) / (R(2) * q);
            fb = inv2q * plain + (R(1) - inv2q) * CurrentPop<Q, S, Lay>(pops0, pops1, t, g, a, b, e, d);
        }
        const R dp = (plain + fb) - R(2) * W<Q, R>(d);
        fx -= static_cast<R>(Lat<Q>::ex[d]) * dp;
        fy -= static_cast<R>(Lat<Q>::ey[d]) * dp;
        fz -= static_cast<R>(Lat<Q>::ez[d]) * dp;
        wsum += W<Q, R>(d);
    }
    const R mag = sqrt(fx * fx + fy * fy + fz * fz);
    out[k * 4 + 0] = wsum > R(0) ? mag / (R(6) * wsum) : R(0);
    out[k * 4 + 1] = fx;
    out[k * 4 + 2] = fy;
    out[k * 4 + 3] = fz;
}

// Macro field of the layers [z0, z0+zc) into out (4 floats per cell, chunk-local
// indexing).  Fluid: moments of the current populations (LbmReference.cpp:43-61);
// solid: zero here, surface cells are filled by KScatterSurface.
template <int Q, class S, class Lay>
__device__ __forceinline__ float4 FluidField(const void* pops0, const void* pops1, long long t, const Geo& g, int x,
                                             int y, int z) {
    using R = typename S::R;
    R rho = 0, ux = 0, uy = 0, uz = 0;
#pragma unroll
    for (int d = 0; d < Q; ++d) {
        const R fi = CurrentPop<Q, S, Lay>(pops0, pops1, t, g, x, y, z, d);
        rho += fi;
        ux += fi * static_cast<R>(Lat<Q>::ex[d]);
        uy += fi * static_cast<R>(Lat<Q>::ey[d]);
        uz += fi * static_cast<R>(Lat<Q>::ez[d]);
    }
    if (rho > K<R>(1e-6f, 1e-6)) {
        ux /= rho;
        uy /= rho;
        uz /= rho;
    }
    return make_float4(static_cast<float>(rho), static_cast<float>(ux), static_cast<float>(uy), static_cast<float>(uz));
}

template <int Q, class S, class Lay>
__global__ void KField(const void* pops0, const void* pops1, long long P, long long t, int nx, int ny, int nz,
                       int pery, int perz, const uint8_t* __restrict__ flags, int z0, int zc,
                       float* __restrict__ out) {
    const long long cnt = static_cast<long long>(nx) * ny * zc;
    const long long j = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= cnt) return;
    int x, y, zr;
    Coords(j, nx, ny, x, y, zr);
    const int z = z0 + zr;
    const Geo g{nx, ny, nz, pery, perz, nx, nx * (ny + 2), Q, PopRowStride(nx, Q)};
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (!(flags[g.Idx(x, y, z)] & kSolid)) v = FluidField<Q, S, Lay>(pops0, pops1, t, g, x, y, z);
    reinterpret_cast<float4*>(out)[j] = v;
}

// The fluid part of the field at a list of cells (global index (z*ny + y)*nx + x): the
// same values as KField there, zero for solid cells and for indices outside
// [0, nx*ny*nz), which read nothing.
template <int Q, class S, class Lay>
__global__ void KFieldAt(const void* pops0, const void* pops1, long long t, int nx, int ny, int nz, int pery,
                         int perz, const uint8_t* __restrict__ flags, const int* __restrict__ cells, long long n,
                         float* __restrict__ out) {
    const long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int c = cells[k];
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (c < 0 || static_cast<long long>(c) >= static_cast<long long>(nx) * ny * nz) {
        reinterpret_cast<float4*>(out)[k] = v;
        return;
    }
    int x, y, z;
    Coords(c, nx, ny, x, y, z);
    const Geo g{nx, ny, nz, pery, perz, nx, nx * (ny + 2), Q, PopRowStride(nx, Q)};
    if (!(flags[g.Idx(x, y, z)] & kSolid)) v = FluidField<Q, S, Lay>(pops0, pops1, t, g, x, y, z);
    reinterpret_cast<float4*>(out)[k] = v;
}

template <class R>
__global__ void KScatterSurface(const long long* __restrict__ surf, int nsurf, const R* __restrict__ sforce, int nx,
                                int ny, int z0, int zc, float* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nsurf) return;
    int x, y, z;
    PaddedCoords(surf[k], nx, ny, x, y, z);
    if (z < z0 || z >= z0 + zc) return;
    const long long j = (static_cast<long long>(z - z0) * ny + y) * nx + x;
    out[j * 4 + 0] = static_cast<float>(sforce[k * 4 + 0]);
    out[j * 4 + 1] = static_cast<float>(sforce[k * 4 + 1]);
    out[j * 4 + 2] = static_cast<float>(sforce[k * 4 + 2]);
    out[j * 4 + 3] = static_cast<float>(sforce[k * 4 + 3]);
}

// Fixed-order force reduction (single block): sums of finite (Fx,Fy,Fz) in FP64
// and the max stress (LbmReference.cpp:523-531).  out: Fx, Fy, Fz, maxStress.
template <class R>
__global__ __launch_bounds__(1024) void KReduceForces(const R* __restrict__ sforce, int nsurf, double* __restrict__ out) {
    __shared__ double sx[1024], sy[1024], sz[1024], sm[1024];
    double ax = 0, ay = 0, az = 0, am = 0;
    for (int k = threadIdx.x; k < nsurf; k += 1024) {
        const R st = sforce[k * 4 + 0], fx = sforce[k * 4 + 1], fy = sforce[k * 4 + 2], fz = sforce[k * 4 + 3];
        if (__builtin_isfinite(st) && __builtin_isfinite(fx) && __builtin_isfinite(fy) && __builtin_isfinite(fz)) {
            ax += static_cast<double>(fx);
            ay += static_cast<double>(fy);
            az += static_cast<double>(fz);
            am = am < static_cast<double>(st) ? static_cast<double>(st) : am;
        }
    }
    sx[threadIdx.x] = ax;
    sy[threadIdx.x] = ay;
    sz[threadIdx.x] = az;
    sm[threadIdx.x] = am;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sx[threadIdx.x] += sx[threadIdx.x + s];
            sy[threadIdx.x] += sy[threadIdx.x + s];
            sz[threadIdx.x] += sz[threadIdx.x + s];
            sm[threadIdx.x] = sm[threadIdx.x] < sm[threadIdx.x + s] ? sm[threadIdx.x + s] : sm[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sx[0];
        out[1] = sy[0];
        out[2] = sz[0];
        out[3] = sm[0];
    }
}

// ------------------------------------------------------- aerodynamic forces
// FieldOverlay::ComputeAerodynamicForces (FieldOverlay.cpp:469-621) over the solid cells
// of the current state, without the field readback.  The field value of a solid cell is
// (stress, Fx, Fy, Fz) of KSurfaceForce rounded to float for surface cells and exactly
// zero for the others, so the force and weighted-position sums run over the surface
// list; the geometric sums (solid count, centroid, top layer) depend only on the
// geometry and the world frame and are computed once per frame.  Cell centres follow
// VoxelGrid::CellCenter (VoxelGrid.h:49-51): grid-y along world X, grid-x along world
// Y, in float as there; every sum is FP64, reduced in a fixed order.
constexpr int kAeroBlocks = 256;   // KAeroSurface grid
constexpr int kAeroSVals = 13;     // its per-block partial
constexpr int kGeomBlocks = 512;   // KAeroGeom grid
constexpr int kAeroGVals = 5;      // its per-block partial
constexpr int kAeroPartDoubles = kAeroBlocks * kAeroSVals + kGeomBlocks * kAeroGVals;

__device__ __forceinline__ void CellCenterWorld(int x, int y, int z, float cs, float ox, float oy, float oz, double& px,
                                                double& py, double& pz) {
    px = static_cast<double>(ox + (static_cast<float>(y) + 0.5f) * cs);
    py = static_cast<double>(oy + (static_cast<float>(x) + 0.5f) * cs);
    pz = static_cast<double>(oz + (static_cast<float>(z) + 0.5f) * cs);
}

// Fixed-order tree reduction of V values over a 256-thread block; `is_max` selects
// max instead of sum per value.  Returns the block result in v[i][0].
template <int V>
__device__ __forceinline__ void BlockReduce256(double (&v)[V][256], const bool (&is_max)[V]) {
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            const int o = threadIdx.x + s;
#pragma unroll
            for (int i = 0; i < V; ++i)
                v[i][threadIdx.x] = is_max[i] ? (v[i][threadIdx.x] > v[i][o] ? v[i][threadIdx.x] : v[i][o])
                                              : v[i][threadIdx.x] + v[i][o];
        }
        __syncthreads();
    }
}

// Surface pass: part[b*kAeroSVals + i] = 0-2 sum F, 3 sum |Fz|, 4-6 sum |Fz| pos,
// 7 sum |F|, 8-10 sum |F| pos, 11 max positive stress (-1 if none), 12 non-finite seen.
template <class R>
__global__ __launch_bounds__(256) void KAeroSurface(const long long* __restrict__ surf, int nsurf,
                                                    const R* __restrict__ sforce, int nx, int ny, float cs, float ox,
                                                    float oy, float oz, double* __restrict__ part) {
    __shared__ double v[kAeroSVals][256];
    double a[kAeroSVals];
#pragma unroll
    for (int i = 0; i < kAeroSVals; ++i) a[i] = 0.0;
    a[11] = -1.0;
    for (int k = blockIdx.x * 256 + threadIdx.x; k < nsurf; k += kAeroBlocks * 256) {
        int x, y, z;
        PaddedCoords(surf[k], nx, ny, x, y, z);
        const float st = static_cast<float>(sforce[k * 4 + 0]);
        const float fx = static_cast<float>(sforce[k * 4 + 1]);
        const float fy = static_cast<float>(sforce[k * 4 + 2]);
        const float fz = static_cast<float>(sforce[k * 4 + 3]);
        if (!__builtin_isfinite(st) || !__builtin_isfinite(fx) || !__builtin_isfinite(fy) || !__builtin_isfinite(fz)) {
            a[12] = 1.0;
            continue;
        }
        double px, py, pz;
        CellCenterWorld(x, y, z, cs, ox, oy, oz, px, py, pz);
        const double dx = fx, dy = fy, dz = fz;
        a[0] += dx;
        a[1] += dy;
        a[2] += dz;
        const double wl = fabs(dz);
        a[3] += wl;
        a[4] += wl * px;
        a[5] += wl * py;
        a[6] += wl * pz;
        const double wf = sqrt(dx * dx + dy * dy + dz * dz);
        a[7] += wf;
        a[8] += wf * px;
        a[9] += wf * py;
        a[10] += wf * pz;
        if (st > 0.0f && static_cast<double>(st) > a[11]) a[11] = st;
    }
#pragma unroll
    for (int i = 0; i < kAeroSVals; ++i) v[i][threadIdx.x] = a[i];
    constexpr bool mx[kAeroSVals] = {false, false, false, false, false, false, false,
                                     false, false, false, false, true,  true};
    BlockReduce256<kAeroSVals>(v, mx);
    if (threadIdx.x < kAeroSVals) part[blockIdx.x * kAeroSVals + threadIdx.x] = v[threadIdx.x][0];
}

// Geometry pass over whole rows of the flag array (no per-cell division):
// part[b*kAeroGVals + i] = 0 solid cells, 1-3 sum pos, 4 top solid layer (-1 if none).
__global__ __launch_bounds__(256) void KAeroGeom(const uint8_t* __restrict__ flags, int nx, int ny, int nz, float cs,
                                                 float ox, float oy, float oz, double* __restrict__ part) {
    __shared__ double v[kAeroGVals][256];
    double cnt = 0.0, sx = 0.0, sy = 0.0, sz = 0.0, top = -1.0;
    const int rows = ny * nz;
    for (int r = blockIdx.x; r < rows; r += kGeomBlocks) {
        const int z = r / ny, y = r - (r / ny) * ny;
        const uint8_t* row = flags + (static_cast<long long>(z + 1) * (ny + 2) + (y + 1)) * nx;
        const double py = static_cast<double>(ox + (static_cast<float>(y) + 0.5f) * cs);  // world X
        const double pz = static_cast<double>(oz + (static_cast<float>(z) + 0.5f) * cs);
        double c = 0.0, s = 0.0;
        for (int x = threadIdx.x; x < nx; x += 256) {
            if (!(row[x] & kSolid)) continue;
            c += 1.0;
            s += static_cast<double>(oy + (static_cast<float>(x) + 0.5f) * cs);  // world Y
        }
        cnt += c;
        sx += c * py;
        sy += s;
        sz += c * pz;
        if (c > 0.0 && z > top) top = z;
    }
    v[0][threadIdx.x] = cnt;
    v[1][threadIdx.x] = sx;
    v[2][threadIdx.x] = sy;
    v[3][threadIdx.x] = sz;
    v[4][threadIdx.x] = top;
    constexpr bool mx[kAeroGVals] = {false, false, false, false, true};
    BlockReduce256<kAeroGVals>(v, mx);
    if (threadIdx.x < kAeroGVals) part[blockIdx.x * kAeroGVals + threadIdx.x] = v[threadIdx.x][0];
}

// Folds `blocks` partials of V values (fixed order: each thread its run of consecutive
// blocks, then a tree) into out[0..V).  One block of 256 threads.
template <int V>
__global__ __launch_bounds__(256) void KAeroFold(const double* __restrict__ part, int blocks, bool max_last_two,
                                                 bool max_last, double* __restrict__ out) {
    __shared__ double v[V][256];
    bool mx[V];
#pragma unroll
    for (int i = 0; i < V; ++i) mx[i] = (max_last_two && i >= V - 2) || (max_last && i == V - 1);
    const int per = (blocks + 255) / 256;
    double a[V];
#pragma unroll
    for (int i = 0; i < V; ++i) a[i] = mx[i] ? -1.0 : 0.0;
    for (int j = 0; j < per; ++j) {
        const int b = threadIdx.x * per + j;
        if (b >= blocks) break;
#pragma unroll
        for (int i = 0; i < V; ++i) {
            const double p = part[b * V + i];
            a[i] = mx[i] ? (a[i] > p ? a[i] : p) : a[i] + p;
        }
    }
#pragma unroll
    for (int i = 0; i < V; ++i) v[i][threadIdx.x] = a[i];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            const int o = threadIdx.x + s;
#pragma unroll
            for (int i = 0; i < V; ++i)
                v[i][threadIdx.x] = mx[i] ? (v[i][threadIdx.x] > v[i][o] ? v[i][threadIdx.x] : v[i][o])
                                          : v[i][threadIdx.x] + v[i][o];
        }
        __syncthreads();
    }
    if (threadIdx.x < V) out[threadIdx.x] = v[threadIdx.x][0];
}

// Fluid statistics, pass 1: a fixed grid of kStatBlocks blocks, grid-stride over
// cells, fixed-order block reduction.  part[b*8 + ...] = dsum, ssum, count, nan,
// minD, maxD, maxS, unused.
constexpr int kStatBlocks = 1024;

template <int Q, class S, class Lay>
__global__ __launch_bounds__(256) void KStatsPartial(const void* pops0, const void* pops1, long long P, long long t,
                                                     int nx, int ny, int nz, int pery, int perz,
                                                     const uint8_t* __restrict__ flags, double* __restrict__ part) {
    using R = typename S::R;
    const long long N = static_cast<long long>(nx) * ny * nz;
    const Geo g{nx, ny, nz, pery, perz, nx, nx * (ny + 2), Q, PopRowStride(nx, Q)};
    double dsum = 0, ssum = 0, cnt = 0, nan = 0;
    double mn = 1e30, mx = -1e30, ms = 0;
    for (long long c = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x; c < N;
         c += static_cast<long long>(kStatBlocks) * 256) {
        int x, y, z;
        Coords(c, nx, ny, x, y, z);
        if (flags[g.Idx(x, y, z)] & kSolid) continue;
        R rho = 0, ux = 0, uy = 0, uz = 0;
#pragma unroll
        for (int d = 0; d < Q; ++d) {
            const R fi = CurrentPop<Q, S, Lay>(pops0, pops1, t, g, x, y, z, d);
            rho += fi;
            ux += fi * static_cast<R>(Lat<Q>::ex[d]);
            uy += fi * static_cast<R>(Lat<Q>::ey[d]);
            uz += fi * static_cast<R>(Lat<Q>::ez[d]);
        }
        if (rho > K<R>(1e-6f, 1e-6)) {
            ux /= rho;
            uy /= rho;
            uz /= rho;
        }
        if (!__builtin_isfinite(rho) || !__builtin_isfinite(ux) || !__builtin_isfinite(uy) || !__builtin_isfinite(uz)) {
            nan = 1;
            continue;
        }
        const R sp = sqrt(ux * ux + uy * uy + uz * uz);
        const double r = static_cast<double>(rho);
        mn = r < mn ? r : mn;
        mx = r > mx ? r : mx;
        ms = static_cast<double>(sp) > ms ? static_cast<double>(sp) : ms;
        dsum += r;
        ssum += static_cast<double>(sp);
        cnt += 1;
    }
    __shared__ double v[7][256];
    v[0][threadIdx.x] = dsum;
    v[1][threadIdx.x] = ssum;
    v[2][threadIdx.x] = cnt;
    v[3][threadIdx.x] = nan;
    v[4][threadIdx.x] = mn;
    v[5][threadIdx.x] = mx;
    v[6][threadIdx.x] = ms;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            const int o = threadIdx.x + s;
            v[0][threadIdx.x] += v[0][o];
            v[1][threadIdx.x] += v[1][o];
            v[2][threadIdx.x] += v[2][o];
            v[3][threadIdx.x] = v[3][threadIdx.x] > v[3][o] ? v[3][threadIdx.x] : v[3][o];
            v[4][threadIdx.x] = v[4][threadIdx.x] < v[4][o] ? v[4][threadIdx.x] : v[4][o];
            v[5][threadIdx.x] = v[5][threadIdx.x] > v[5][o] ? v[5][threadIdx.x] : v[5][o];
            v[6][threadIdx.x] = v[6][threadIdx.x] > v[6][o] ? v[6][threadIdx.x] : v[6][o];
        }
        __syncthreads();
    }
    if (threadIdx.x < 7) part[blockIdx.x * 8 + threadIdx.x] = v[threadIdx.x][0];
}

// Pass 2: one thread folds the kStatBlocks partials in block order.
__global__ void KStatsFinal(const double* __restrict__ part, double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double dsum = 0, ssum = 0, cnt = 0, nan = 0, mn = 1e30, mx = -1e30, ms = 0;
    for (int b = 0; b < kStatBlocks; ++b) {
        const double* p = part + b * 8;
        dsum += p[0];
        ssum += p[1];
        cnt += p[2];
        nan = nan > p[3] ? nan : p[3];
        mn = mn < p[4] ? mn : p[4];
        mx = mx > p[5] ? mx : p[5];
        ms = ms > p[6] ? ms : p[6];
    }
    out[0] = dsum;
    out[1] = ssum;
    out[2] = cnt;
    out[3] = nan;
    out[4] = mn;
    out[5] = mx;
    out[6] = ms;
}

// Health probe of one layer (LbmSolver.cpp:653-720): non-finite anywhere, or a
// fluid cell with rho <= 0 or |u|^2 > 4.  Surface solid cells are checked from
// the force buffer by KProbeSurface.
template <int Q, class S, class Lay>
__global__ void KProbe(const void* pops0, const void* pops1, long long P, long long t, int nx, int ny, int nz, int pery,
                       int perz, const uint8_t* __restrict__ flags, int z, int* __restrict__ bad) {
    using R = typename S::R;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nx * ny) return;
    const int y = j / nx, x = j - y * nx;
    const Geo g{nx, ny, nz, pery, perz, nx, nx * (ny + 2), Q, PopRowStride(nx, Q)};
    if (flags[g.Idx(x, y, z)] & kSolid) return;
    R rho = 0, ux = 0, uy = 0, uz = 0;
#pragma unroll
    for (int d = 0; d < Q; ++d) {
        const R fi = CurrentPop<Q, S, Lay>(pops0, pops1, t, g, x, y, z, d);
        rho += fi;
        ux += fi * static_cast<R>(Lat<Q>::ex[d]);
        uy += fi * static_cast<R>(Lat<Q>::ey[d]);
        uz += fi * static_cast<R>(Lat<Q>::ez[d]);
    }
    if (rho > K<R>(1e-6f, 1e-6)) {
        ux /= rho;
        uy /= rho;
        uz /= rho;
    }
    const float r = static_cast<float>(rho), u = static_cast<float>(ux), v = static_cast<float>(uy),
                w = static_cast<float>(uz);
    if (!__builtin_isfinite(r) || !__builtin_isfinite(u) || !__builtin_isfinite(v) || !__builtin_isfinite(w) ||
        r <= 0.0f || (u * u + v * v + w * w) > 4.0f)
        atomicOr(bad, 1);
}

template <class R>
__global__ void KProbeSurface(const long long* __restrict__ surf, int nsurf, const R* __restrict__ sforce, int nx,
                              int ny, int z, int* __restrict__ bad) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nsurf) return;
    int x, y, zc;
    PaddedCoords(surf[k], nx, ny, x, y, zc);
    if (zc != z) return;
    for (int i = 0; i < 4; ++i)
        if (!__builtin_isfinite(static_cast<float>(sforce[k * 4 + i]))) atomicOr(bad, 1);
}

}  // namespace fc
