// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// Clean-room CPU restatement of FlowCraft's SmokeTrailSystem (FlowCraft/SmokeTrail.cpp),
// checking the engine's device smoke trails (flowcraft_amd/csrc/smoke.hip):
//   cell lookup of a world point ........... SmokeTrail.cpp:36-50
//   trilinear velocity sampling ............. SmokeTrail.cpp:52-99
//   random cell pick + jittered centre ...... SmokeTrail.cpp:26-34, 148-166
//   respawn at the inlet face / plane ....... SmokeTrail.cpp:168-200
//   scatter, plane shift, reset ............. SmokeTrail.cpp:202-275
//   step ................................... SmokeTrail.cpp:277-345
// Random draws use std::mt19937 (seed 12345) with the standard library's uniform
// distributions, in the reference's order.  Float operations in the reference's order
// (built with -ffp-contract=off).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

namespace {

struct P3 {
    float x, y, z;
};

struct Smoke {
    int nx, ny, nz;
    float cs, ox, oy, oz;
    std::vector<uint8_t> solid;
    std::vector<std::vector<P3>> trails;
    int L = 10;
    std::mt19937 rng{12345};

    bool In(int x, int y, int z) const { return x >= 0 && y >= 0 && z >= 0 && x < nx && y < ny && z < nz; }
    bool SolidAt(int x, int y, int z) const {
        return In(x, y, z) && solid[(static_cast<size_t>(z) * ny + y) * nx + x] != 0;
    }
    P3 Centre(int x, int y, int z) const { return {ox + (y + 0.5f) * cs, oy + (x + 0.5f) * cs, oz + (z + 0.5f) * cs}; }

    void Pick(int fx, int fy, int fz, int& x, int& y, int& z) {
        std::uniform_int_distribution<int> dx(0, std::max(nx - 1, 0)), dy(0, std::max(ny - 1, 0)),
            dz(0, std::max(nz - 1, 0));
        x = fx >= 0 ? fx : nx / 2;
        y = fy >= 0 ? fy : ny / 2;
        z = fz >= 0 ? fz : nz / 2;
        for (int a = 0; a < 8; ++a) {
            const int tx = fx >= 0 ? fx : dx(rng);
            const int ty = fy >= 0 ? fy : dy(rng);
            const int tz = fz >= 0 ? fz : dz(rng);
            if (!SolidAt(tx, ty, tz)) {
                x = tx;
                y = ty;
                z = tz;
                return;
            }
        }
    }

    P3 Jitter(int x, int y, int z, bool xf, bool yf, bool zf) {
        std::uniform_real_distribution<float> u(-0.5f, 0.5f);
        const float jx = xf ? 0.0f : u(rng) * cs;
        const float jy = yf ? 0.0f : u(rng) * cs;
        const float jz = zf ? 0.0f : u(rng) * cs;
        const P3 c = Centre(x, y, z);
        return {c.x + jy, c.y + jx, c.z + jz};
    }

    static int Pin(int axis, int want, int index, int n) { return axis == want ? std::clamp(index, 0, n - 1) : -1; }

    void Respawn(int reverse, int axis, int index, std::vector<P3>& t) {
        const int fx = axis == 0 ? std::clamp(index, 0, nx - 1) : (reverse ? nx - 1 : 0);
        const int fy = Pin(axis, 1, index, ny), fz = Pin(axis, 2, index, nz);
        int x, y, z;
        Pick(fx, fy, fz, x, y, z);
        t.clear();
        // the reference passes "not pinned" as the y/z fixed flags here (SmokeTrail.cpp:199)
        t.push_back(Jitter(x, y, z, true, fy < 0, fz < 0));
    }

    void Scatter(int axis, int index) {
        const int fx = Pin(axis, 0, index, nx), fy = Pin(axis, 1, index, ny), fz = Pin(axis, 2, index, nz);
        for (auto& t : trails) {
            int x, y, z;
            Pick(fx, fy, fz, x, y, z);
            t.clear();
            t.push_back(Jitter(x, y, z, fx >= 0, fy >= 0, fz >= 0));
        }
    }

    P3 Velocity(const float* field, P3 p) const {
        const float gx = (p.y - oy) / cs - 0.5f, gy = (p.x - ox) / cs - 0.5f, gz = (p.z - oz) / cs - 0.5f;
        const int x0 = static_cast<int>(std::floor(gx)), y0 = static_cast<int>(std::floor(gy)),
                  z0 = static_cast<int>(std::floor(gz));
        const float tx = gx - static_cast<float>(x0), ty = gy - static_cast<float>(y0), tz = gz - static_cast<float>(z0);
        auto at = [&](int x, int y, int z) -> P3 {
            const size_t c = (static_cast<size_t>(std::clamp(z, 0, nz - 1)) * ny + std::clamp(y, 0, ny - 1)) * nx +
                             std::clamp(x, 0, nx - 1);
            if (solid[c]) return {0.0f, 0.0f, 0.0f};
            return {field[c * 4 + 2], field[c * 4 + 1], field[c * 4 + 3]};  // grid (ux, uy, uz) -> world
        };
        auto lerp = [](P3 a, P3 b, float t) { return P3{a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t}; };
        const P3 c00 = lerp(at(x0, y0, z0), at(x0 + 1, y0, z0), tx);
        const P3 c10 = lerp(at(x0, y0 + 1, z0), at(x0 + 1, y0 + 1, z0), tx);
        const P3 c01 = lerp(at(x0, y0, z0 + 1), at(x0 + 1, y0, z0 + 1), tx);
        const P3 c11 = lerp(at(x0, y0 + 1, z0 + 1), at(x0 + 1, y0 + 1, z0 + 1), tx);
        return lerp(lerp(c00, c10, ty), lerp(c01, c11, ty), tz);
    }

    void Step(const float* field, int reverse, float mmps, float dt, int suspend, int axis, int index) {
        if (dt <= 0.0f) return;
        dt = std::min(dt, 0.1f);
        for (auto& t : trails) {
            if (t.empty()) {
                if (!suspend) Respawn(reverse, axis, index, t);
                continue;
            }
            const P3 cur = t.back();
            const int cx = static_cast<int>(std::floor((cur.y - oy) / cs));
            const int cy = static_cast<int>(std::floor((cur.x - ox) / cs));
            const int cz = static_cast<int>(std::floor((cur.z - oz) / cs));
            if (!In(cx, cy, cz) || SolidAt(cx, cy, cz)) {
                if (suspend)
                    t.clear();
                else
                    Respawn(reverse, axis, index, t);
                continue;
            }
            P3 v = Velocity(field, cur);
            if (axis == 0) v.y = 0.0f;
            if (axis == 1) v.x = 0.0f;
            if (axis == 2) v.z = 0.0f;
            const float s = mmps * dt;
            t.push_back({cur.x + v.x * s, cur.y + v.y * s, cur.z + v.z * s});
            if (static_cast<int>(t.size()) > L) t.erase(t.begin());
        }
    }
};

}  // namespace

extern "C" {

void* orc_smoke_create(int nx, int ny, int nz, float cs, const float* origin, const uint8_t* solid) {
    Smoke* s = new Smoke{nx, ny, nz, cs, origin[0], origin[1], origin[2], {}, {}, 10, std::mt19937(12345)};
    s->solid.assign(solid, solid + static_cast<size_t>(nx) * ny * nz);
    return s;
}
void orc_smoke_destroy(void* h) { delete static_cast<Smoke*>(h); }
void orc_smoke_reset(void* h, int /*reverse*/, int count, int trail, int axis, int index) {
    Smoke* s = static_cast<Smoke*>(h);
    s->L = std::max(trail, 2);
    s->trails.assign(static_cast<size_t>(std::max(count, 0)), {});
    s->Scatter(axis, index);
}
void orc_smoke_scatter(void* h, int axis, int index) { static_cast<Smoke*>(h)->Scatter(axis, index); }
void orc_smoke_shift_plane(void* h, int axis, int index) {
    Smoke* s = static_cast<Smoke*>(h);
    for (auto& t : s->trails)
        for (P3& p : t) {
            if (axis == 0) p.y = s->Centre(std::clamp(index, 0, s->nx - 1), 0, 0).y;
            if (axis == 1) p.x = s->Centre(0, std::clamp(index, 0, s->ny - 1), 0).x;
            if (axis == 2) p.z = s->Centre(0, 0, std::clamp(index, 0, s->nz - 1)).z;
        }
}
void orc_smoke_step(void* h, const float* field, int reverse, float mmps, float dt, int suspend, int axis, int index) {
    static_cast<Smoke*>(h)->Step(field, reverse, mmps, dt, suspend, axis, index);
}
// out[(p*L + k)*3 + c], lengths[p]; L = the trail length given to reset (at least 2).
void orc_smoke_trails(void* h, float* out, int* lengths) {
    Smoke* s = static_cast<Smoke*>(h);
    for (size_t p = 0; p < s->trails.size(); ++p) {
        lengths[p] = static_cast<int>(s->trails[p].size());
        for (size_t k = 0; k < s->trails[p].size(); ++k) {
            out[(p * s->L + k) * 3] = s->trails[p][k].x;
            out[(p * s->L + k) * 3 + 1] = s->trails[p][k].y;
            out[(p * s->L + k) * 3 + 2] = s->trails[p][k].z;
        }
    }
}

}  // extern "C"
