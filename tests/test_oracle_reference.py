"""Pins the FP32 oracle to FlowCraft's own LbmReference.

* Always: against the golden fixtures written by tests/golden/make_golden.py from
  the compiled reference (oracle/_ref): the oracle's readback field must hash to
  the reference's SHA-256 (every bit), and ComputeStats must agree.
* Where oracle/_ref is built (the container with the reference sources): live
  bit-for-bit comparisons on scenarios the fixtures do not cover.
"""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np
import pytest

import oracle
from tests._cases import (RAGGED_CASES, case_a_grid, make_oracle, oracle_params, ragged_params,
                          small_sphere_grid, wing_grid)
from flowcraft_amd.lbm import LatticeModel, Params

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "lbm_reference_golden.npz")
STATS_KEYS = ("avgDensity", "minDensity", "maxDensity", "avgSpeed", "maxSpeed", "forceX", "forceY", "forceZ",
              "maxStress", "totalMass", "hasNaN")
GEOMETRIES = {
    "caseA": case_a_grid,
    "small": small_sphere_grid,
    "small_nosdf": lambda: small_sphere_grid(with_sdf=False),
    "wing": wing_grid,
}


def load_golden():
    z = np.load(GOLDEN, allow_pickle=False)
    meta = json.loads(bytes(z["meta"]).decode())
    return z, meta


def geometry_sha(g) -> str:
    h = hashlib.sha256(np.ascontiguousarray(g.solid, np.uint8).tobytes())
    if g.signedDistance is not None:
        h.update(np.ascontiguousarray(g.signedDistance, np.float32).tobytes())
    return h.hexdigest()


_z, _meta = load_golden()
CASES = sorted(_meta)


def golden_case(name):
    m = _meta[name]
    g = GEOMETRIES[m["geometry"]]()
    assert geometry_sha(g) == m["geometry_sha256"], "geometry generator no longer reproduces the fixture inputs"
    p = m["params"]
    params = Params(tau=p["tau"], inletVelocity=p["u"], reverseFlow=p["reverse"], freeSlipY=p["fsy"],
                    freeSlipZ=p["fsz"], smagorinskyCs=p["cs"])
    model = LatticeModel.D3Q15 if m["q"] == 15 else LatticeModel.D3Q19
    return g, params, model, m


@pytest.mark.parametrize("name", CASES)
def test_oracle_reproduces_reference_golden(name):
    g, params, model, m = golden_case(name)
    o = make_oracle(g, model)
    oracle_params(o, params)
    o.reset(m["seed_mid"])
    o.step(m["steps"])
    f = o.field()
    sample = _z[f"{name}/sample"]
    assert np.array_equal(f.reshape(-1, 4)[::m["stride"]].view(np.uint32), sample.view(np.uint32))
    assert hashlib.sha256(f.tobytes()).digest() == bytes(_z[f"{name}/sha256"])
    sums = f.reshape(*m["shape"], 4).astype(np.float64).sum(axis=(1, 2))
    assert np.array_equal(sums, _z[f"{name}/layer_sums"])
    st = o.stats()
    ref = dict(zip(STATS_KEYS, _z[f"{name}/stats"]))
    for k in STATS_KEYS:
        # the reference narrows the averages/force sums to float (LbmReference.h:112-119)
        assert np.float32(st[k]) == pytest.approx(np.float32(ref[k]), rel=1e-6, abs=1e-9), k
    assert o.diverged() == bool(_z[f"{name}/diverged"][0])


needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built (no reference sources)")


@needs_ref
@pytest.mark.parametrize("q", [19, 15])
def test_live_reverse_switch_and_reset_from_field_bitexact(q):
    from oracle import RefLbm
    g = wing_grid()
    model = LatticeModel.D3Q15 if q == 15 else LatticeModel.D3Q19
    r = RefLbm(g.solid, g.signedDistance, q=q, tau=0.58, u=0.06, cs=0.1, seed_mid=True)
    o = make_oracle(g, model)
    o.set_params(0.58, 0.06, False, True, True, 0.1)
    o.reset(True)
    r.step(12)
    o.step(12)
    assert np.array_equal(r.field().view(np.uint32), o.field().view(np.uint32))
    r.set_params(0.58, 0.06, True, True, False, 0.1)
    o.set_params(0.58, 0.06, True, True, False, 0.1)
    r.step(9)
    o.step(9)
    assert np.array_equal(r.field().view(np.uint32), o.field().view(np.uint32))
    seed = r.field()
    seed = seed.copy()
    seed.reshape(-1, 4)[:, 2] += np.float32(0.001)
    r.reset_from_field(seed)
    o.reset_from_field(seed)
    r.step(7)
    o.step(7)
    assert np.array_equal(r.field().view(np.uint32), o.field().view(np.uint32))
    assert r.diverged() == o.diverged()


@needs_ref
def test_random_sdf_and_divergence_bitexact():
    """Perturbed SDF values (+-inf sentinels, degenerate pairs) and an unstable run that
    diverges: field bits and the divergence flag match the reference."""
    from oracle import RefLbm
    g = small_sphere_grid()
    rng = np.random.default_rng(11)
    sdf = g.signedDistance.copy()
    sdf += rng.normal(0, 0.3, sdf.shape).astype(np.float32)
    sdf[rng.random(sdf.shape) < 0.05] = np.inf
    sdf[rng.random(sdf.shape) < 0.02] = -np.inf
    sdf = np.where(g.solid == 1, -np.abs(sdf), np.abs(sdf)).astype(np.float32)
    r = RefLbm(g.solid, sdf, q=19, tau=0.503, u=0.3, cs=0.0, seed_mid=False)
    from oracle import OracleLbm
    o = OracleLbm(g.solid, sdf, q=19, bits=32)
    o.set_params(0.503, 0.3, False, True, True, 0.0)
    o.reset(False)
    for _ in range(6):
        r.step(25)
        o.step(25)
        a, b = r.field(), o.field()
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert r.diverged() == o.diverged()


@needs_ref
@pytest.mark.parametrize("q", [19, 15])
@pytest.mark.parametrize("name", sorted(RAGGED_CASES))
def test_ragged_lattices_bitexact(name, q):
    """Lattices at the layout's edges (partial 64-cell blocks, nx = 4, one- and two-cell
    y/z extents, obstacles on a wall): field bits and stats match the reference."""
    from oracle import RefLbm
    build, fsy, fsz = RAGGED_CASES[name]
    g = build()
    p = ragged_params(fsy, fsz)
    model = LatticeModel.D3Q15 if q == 15 else LatticeModel.D3Q19
    r = RefLbm(g.solid, g.signedDistance, q=q, tau=p.tau, u=p.inletVelocity, free_slip_y=fsy, free_slip_z=fsz,
               cs=p.smagorinskyCs, seed_mid=True)
    o = make_oracle(g, model)
    oracle_params(o, p)
    o.reset(True)
    for n in (1, 14):
        r.step(n)
        o.step(n)
        assert np.array_equal(r.field().view(np.uint32), o.field().view(np.uint32))
    rs, os_ = r.stats(), o.stats()
    for k in STATS_KEYS:
        # the reference narrows the averages/force sums to float (LbmReference.h:112-119)
        assert np.float32(os_[k]) == pytest.approx(np.float32(rs[k]), rel=1e-6, abs=1e-9), k
