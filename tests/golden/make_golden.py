"""Writes tests/golden/lbm_reference_golden.npz from FlowCraft's own LbmReference,
compiled from the reference sources into oracle/_ref/libref_lbm.so
(make -C oracle ref).  Run in the container that holds the reference:

    python tests/golden/make_golden.py

Each case stores its inputs — parameters, seed, step count and the SHA-256 of
the solid mask + SDF bytes, which the tests regenerate with tests/_cases.py and
check against that hash — and the reference's outputs after the steps: the SHA-256 of the full 4-floats/cell
readback field (pins every bit), a strided subsample of the field, per-z-layer
float64 sums of each channel, and ComputeStats.  The reference test's own
sphere case (tests/gpu_vs_cpu_test.cpp:27-80) is included for D3Q19 and D3Q15.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import RefLbm  # noqa: E402
from tests._cases import case_a_grid, small_sphere_grid, wing_grid  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lbm_reference_golden.npz")
STRIDE = 31

GEOMETRIES = {
    "caseA": case_a_grid,
    "small": small_sphere_grid,
    "small_nosdf": lambda: small_sphere_grid(with_sdf=False),
    "wing": wing_grid,
}


def geometry_sha(g) -> str:
    h = hashlib.sha256(np.ascontiguousarray(g.solid, np.uint8).tobytes())
    if g.signedDistance is not None:
        h.update(np.ascontiguousarray(g.signedDistance, np.float32).tobytes())
    return h.hexdigest()


def cases():
    base = dict(tau=0.6, u=0.05, reverse=False, fsy=True, fsz=True, cs=0.12)
    out = []
    for q in (19, 15):
        out.append((f"caseA_q{q}_60", "caseA", q, base, True, 60))
        for steps in (1, 2, 10):
            out.append((f"small_cs0_q{q}_{steps}", "small", q, dict(base, tau=0.56, cs=0.0), True, steps))
            out.append((f"small_reverse_q{q}_{steps}", "small", q, dict(base, reverse=True), True, steps))
            out.append((f"small_noslip_q{q}_{steps}", "small", q, dict(base, fsy=False, fsz=False), True, steps))
            out.append((f"small_mixed_q{q}_{steps}", "small", q, dict(base, fsz=False), True, steps))
            out.append((f"small_nosdf_q{q}_{steps}", "small_nosdf", q, base, True, steps))
            out.append((f"small_quiescent_q{q}_{steps}", "small", q, dict(base, tau=0.7, u=0.08, cs=0.15), False,
                        steps))
        out.append((f"wing_q{q}_80", "wing", q, base, True, 80))
    return out


def main():
    data = {}
    meta = {}
    grids = {k: f() for k, f in GEOMETRIES.items()}
    for name, gname, q, p, mid, steps in cases():
        g = grids[gname]
        r = RefLbm(g.solid, g.signedDistance, q=q, tau=p["tau"], u=p["u"], reverse=p["reverse"],
                   free_slip_y=p["fsy"], free_slip_z=p["fsz"], cs=p["cs"], seed_mid=mid)
        r.step(steps)
        f = r.field()
        st = r.stats()
        f4 = f.reshape(g.nz, g.ny, g.nx, 4)
        data[f"{name}/sha256"] = np.frombuffer(hashlib.sha256(f.tobytes()).digest(), np.uint8)
        data[f"{name}/sample"] = f.reshape(-1, 4)[::STRIDE].copy()
        data[f"{name}/layer_sums"] = f4.astype(np.float64).sum(axis=(1, 2))
        data[f"{name}/stats"] = np.array([st[k] for k in ("avgDensity", "minDensity", "maxDensity", "avgSpeed",
                                                          "maxSpeed", "forceX", "forceY", "forceZ", "maxStress",
                                                          "totalMass", "hasNaN")], np.float64)
        data[f"{name}/diverged"] = np.array([r.diverged()], np.uint8)
        meta[name] = dict(q=q, params=p, seed_mid=mid, steps=steps, shape=[g.nz, g.ny, g.nx], stride=STRIDE,
                          geometry=gname, geometry_sha256=geometry_sha(g))
        print(name, "forceX", st["forceX"], flush=True)
    data["meta"] = np.frombuffer(json.dumps(meta).encode(), np.uint8)
    np.savez_compressed(OUT, **data)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
