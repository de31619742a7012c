"""GPU parity: the HIP engine (through libfc_lbm.so) against the CPU oracle on the
same inputs.

Bar (stated per test): the FP32 engine must match the FP32 oracle — itself
bit-exact with FlowCraft's compiled LbmReference (test_oracle_reference.py) —
within 1e-5 relative to each field channel's scale; the run also records the
max float32 ULP distance, which is 0 when the kernels reproduce the reference's
rounding.  Reduced-storage runs are compared with the oracle's exact emulation of
the same storage format; FP64 runs with the FP64 oracle at 1e-12.
"""
from __future__ import annotations

import numpy as np
import pytest

from flowcraft_amd.lbm import LatticeModel, Layout, Params, Storage
from tests._cases import (RAGGED_CASES, case_a_grid, case_a_params, make_engine, make_oracle, oracle_params,
                          ragged_params, rel_err, small_sphere_grid, ulp_diff32, wing_grid)

pytestmark = pytest.mark.gpu

LAYOUTS = [Layout.PINGPONG, Layout.INPLACE]
MODELS = [LatticeModel.D3Q19, LatticeModel.D3Q15]


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu-case] {request.node.nodeid}", flush=True)
    yield


def _fields_close(g, o, tol, solid):
    """Per-channel relative error: max|g-o| / max|o| over fluid cells (rho, ux, uy, uz)
    and over solid cells (stress, Fx, Fy, Fz)."""
    g4, o4 = g.reshape(-1, 4), o.reshape(-1, 4)
    fl = solid.reshape(-1) == 0
    errs = []
    for mask in (fl, ~fl):
        if not mask.any():
            continue
        for k in range(4):
            a, b = g4[mask, k], o4[mask, k]
            scale = max(float(np.abs(b).max()), 1e-6)
            errs.append(float(np.abs(a.astype(np.float64) - b).max()) / scale)
    worst = max(errs)
    assert worst <= tol, f"relative field error {worst:.3e} > {tol:.1e}"
    return worst


def _run_pair(grid, params, model=LatticeModel.D3Q19, storage=Storage.F32, layout=Layout.PINGPONG, steps=60,
              seed_mid=True, periodic=(False, False)):
    eng = make_engine(grid, model, storage, layout, periodic)
    eng.Reset(params, seed_mid)
    orc = make_oracle(grid, model, storage, periodic)
    oracle_params(orc, params)
    orc.reset(seed_mid)
    eng.Step(None, steps)
    orc.step(steps)
    return eng, orc


@pytest.mark.parametrize("layout", LAYOUTS, ids=lambda l: l.name)
@pytest.mark.parametrize("model", MODELS, ids=lambda m: m.name)
def test_case_a_fp32(model, layout):
    """The reference's own test case (tests/gpu_vs_cpu_test.cpp:27-80), 60 steps."""
    grid = case_a_grid()
    eng, orc = _run_pair(grid, case_a_params(), model, Storage.F32, layout, steps=60)
    g, o = eng.ReadbackField(), orc.field()
    _fields_close(g, o, 1e-5, grid.solid)
    print(f"max ulp {ulp_diff32(g, o)}")
    s, so = eng.ReadbackStats(), orc.stats()
    for k_e, k_o in (("forceX", "forceX"), ("avgDensity", "avgDensity"), ("avgSpeed", "avgSpeed"),
                     ("totalMass", "totalMass")):
        assert abs(getattr(s, k_e) - so[k_o]) <= 1e-5 * max(abs(so[k_o]), 1e-6), (k_e, getattr(s, k_e), so[k_o])
    assert abs(s.forceY - so["forceY"]) <= 1e-5 * abs(so["forceX"])
    assert abs(s.forceZ - so["forceZ"]) <= 1e-5 * abs(so["forceX"])
    assert abs(s.maxStress - so["maxStress"]) <= 1e-5 * so["maxStress"]
    assert s.minDensity == pytest.approx(so["minDensity"], rel=1e-6)
    assert s.maxDensity == pytest.approx(so["maxDensity"], rel=1e-6)
    assert not s.hasNaN
    # and its error against the FP64 oracle, recorded
    o64 = make_oracle(grid, model, Storage.F64)
    oracle_params(o64, case_a_params())
    o64.reset(True)
    o64.step(60)
    print(f"vs FP64 oracle: rel field err {rel_err(g, o64.field()):.3e}")


VARIANTS = {
    "cs0": dict(params=Params(tau=0.56, inletVelocity=0.05, smagorinskyCs=0.0)),
    "reverse": dict(params=Params(tau=0.6, inletVelocity=0.05, reverseFlow=True)),
    "noslip": dict(params=Params(tau=0.6, inletVelocity=0.05, freeSlipY=False, freeSlipZ=False)),
    "mixed_walls": dict(params=Params(tau=0.6, inletVelocity=0.05, freeSlipY=True, freeSlipZ=False)),
    "nosdf": dict(params=Params(tau=0.6, inletVelocity=0.05), sdf=False),
    "quiescent": dict(params=Params(tau=0.7, inletVelocity=0.08, smagorinskyCs=0.15), seed_mid=False),
}


@pytest.mark.parametrize("steps", [1, 2, 10])
@pytest.mark.parametrize("layout", LAYOUTS, ids=lambda l: l.name)
@pytest.mark.parametrize("variant", sorted(VARIANTS))
def test_variants_fp32(variant, layout, steps):
    v = VARIANTS[variant]
    grid = small_sphere_grid(with_sdf=v.get("sdf", True))
    for model in MODELS:
        eng, orc = _run_pair(grid, v["params"], model, Storage.F32, layout, steps, v.get("seed_mid", True))
        _fields_close(eng.ReadbackField(), orc.field(), 1e-5, grid.solid)


@pytest.mark.parametrize("layout", LAYOUTS, ids=lambda l: l.name)
@pytest.mark.parametrize("name", sorted(RAGGED_CASES))
def test_ragged_lattices_bitexact(name, layout):
    """Lattices at the layout's edges (partial 64-cell blocks, nx = 4, one- and two-cell
    y/z extents, an obstacle on a wall; the oracle equals FlowCraft's LbmReference on
    them bit for bit): FP32 fields identical to the oracle's, FP16 identical to its
    emulation of the format, after 1 and 15 steps."""
    build, fsy, fsz = RAGGED_CASES[name]
    grid = build()
    p = ragged_params(fsy, fsz)
    for model in MODELS:
        for storage in (Storage.F32, Storage.F16, Storage.F16_SR, Storage.BF16_SR):
            eng = make_engine(grid, model, storage, layout)
            eng.Reset(p, True)
            orc = make_oracle(grid, model, storage)
            oracle_params(orc, p)
            orc.reset(True)
            for n in (1, 14):
                eng.Step(None, n)
                orc.step(n)
                g, o = eng.ReadbackField(), orc.field()
                assert np.array_equal(g.view(np.uint32), o.view(np.uint32)), \
                    f"{model.name} {storage.name}: max ulp {ulp_diff32(g, o)}"
            eng.Shutdown()


@pytest.mark.parametrize("layout", LAYOUTS, ids=lambda l: l.name)
@pytest.mark.parametrize("storage", [Storage.F16, Storage.BF16, Storage.F16_SR, Storage.BF16_SR],
                         ids=lambda s: s.name)
@pytest.mark.parametrize("model", MODELS, ids=lambda m: m.name)
def test_reduced_storage_matches_emulation(model, storage, layout):
    """FP16/BF16 offset storage, rounded to nearest or stochastically, vs the oracle's
    exact emulation of the same format: identical fields."""
    grid = case_a_grid()
    eng, orc = _run_pair(grid, case_a_params(), model, storage, layout, steps=40)
    g, o = eng.ReadbackField(), orc.field()
    print(f"max ulp vs emulation {ulp_diff32(g, o)}")
    _fields_close(g, o, 1e-5, grid.solid)
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32))


@pytest.mark.parametrize("layout", LAYOUTS, ids=lambda l: l.name)
def test_fp64_matches_fp64_oracle(layout):
    """FP64 engine vs FP64 oracle on the double populations (the float readback field
    would cap the comparison at float32 rounding): 1e-12 relative."""
    grid = case_a_grid()
    eng, orc = _run_pair(grid, case_a_params(), LatticeModel.D3Q19, Storage.F64, layout, steps=30)
    fluid = grid.solid == 0
    pe, po = eng.Populations(), orc.pops()
    assert pe.dtype == np.float64
    assert rel_err(pe[:, fluid], po[:, fluid]) <= 1e-12
    f = eng.Forces()
    so = orc.stats()
    assert abs(f[0] - so["forceX"]) <= 1e-12 * abs(so["forceX"])


@pytest.mark.parametrize("layout", LAYOUTS, ids=lambda l: l.name)
def test_periodic_yz(layout):
    """Periodic y/z extension (no reference counterpart) vs the oracle's periodic wrap."""
    grid = small_sphere_grid()
    p = Params(tau=0.56, inletVelocity=0.05, smagorinskyCs=0.0)
    for periodic in [(True, True), (True, False), (False, True)]:
        print("periodic", periodic, flush=True)
        eng, orc = _run_pair(grid, p, LatticeModel.D3Q19, Storage.F32, layout, steps=12, periodic=periodic)
        _fields_close(eng.ReadbackField(), orc.field(), 1e-5, grid.solid)


@pytest.mark.parametrize("storage", [Storage.F32, Storage.F16, Storage.F16_SR, Storage.BF16_SR],
                         ids=lambda s: s.name)
@pytest.mark.parametrize("model", MODELS, ids=lambda m: m.name)
def test_inplace_equals_pingpong_bitwise(model, storage):
    grid = wing_grid()
    p = Params(tau=0.6, inletVelocity=0.05, smagorinskyCs=0.12)
    a = make_engine(grid, model, storage, Layout.PINGPONG)
    b = make_engine(grid, model, storage, Layout.INPLACE)
    a.Reset(p, True)
    b.Reset(p, True)
    a.Step(None, 25)
    b.Step(None, 25)
    fa, fb = a.ReadbackField(), b.ReadbackField()
    assert np.array_equal(fa.view(np.uint32), fb.view(np.uint32))
    assert np.array_equal(a.Populations()[:, grid.solid == 0], b.Populations()[:, grid.solid == 0])


@pytest.mark.parametrize("layout", LAYOUTS, ids=lambda l: l.name)
def test_step_range_chunks_equal_full_step(layout):
    """StepRange chunks covering [0,nz) + AdvanceSubstep == Step (LbmSolver.h:145-170)."""
    grid = case_a_grid()
    p = case_a_params()
    a = make_engine(grid, layout=layout)
    b = make_engine(grid, layout=layout)
    a.Reset(p, True)
    b.Reset(p, True)
    for _ in range(7):
        a.Step(None, 1)
        for z0, zc in ((0, 5), (5, 1), (6, 11), (17, 7)):
            b.StepRange(None, z0, zc)
        b.AdvanceSubstep()
    assert np.array_equal(a.ReadbackField().view(np.uint32), b.ReadbackField().view(np.uint32))


@pytest.mark.parametrize("layout", LAYOUTS, ids=lambda l: l.name)
def test_reset_from_field_and_populations_roundtrip(layout):
    grid = case_a_grid()
    p = case_a_params()
    orc = make_oracle(grid)
    oracle_params(orc, p)
    orc.reset(True)
    orc.step(20)
    seed = orc.field()
    eng = make_engine(grid, layout=layout)
    eng.SetParams(p)
    eng.ResetFromField(seed)
    orc.reset_from_field(seed)
    _fields_close(eng.ReadbackField(), orc.field(), 1e-5, grid.solid)
    eng.Step(None, 15)
    orc.step(15)
    _fields_close(eng.ReadbackField(), orc.field(), 1e-5, grid.solid)
    # exact dump/restore: continuing from a restored state equals continuing uninterrupted
    pops = eng.Populations()
    other = make_engine(grid, layout=layout)
    other.SetParams(p)
    other.ResetFromField(seed)  # any state, then overwritten
    other.SetPopulations(pops)
    eng.Step(None, 9)
    other.Step(None, 9)
    assert np.array_equal(eng.ReadbackField().view(np.uint32), other.ReadbackField().view(np.uint32))


@pytest.mark.parametrize("layout", LAYOUTS, ids=lambda l: l.name)
def test_live_reverse_flow(layout):
    """SetParams(reverseFlow) mid-run without reseeding, as the reference test does."""
    grid = wing_grid()
    p = Params(tau=0.6, inletVelocity=0.05)
    eng, orc = _run_pair(grid, p, LatticeModel.D3Q19, Storage.F32, layout, steps=10)
    p2 = Params(tau=0.6, inletVelocity=0.05, reverseFlow=True)
    eng.SetParams(p2)
    oracle_params(orc, p2)
    eng.Step(None, 10)
    orc.step(10)
    _fields_close(eng.ReadbackField(), orc.field(), 1e-5, grid.solid)


def test_probe_health_and_divergence_flag():
    """ProbeSliceHealth (LbmSolver.cpp:653-720) and the sticky divergence flag
    (LbmReference.cpp:397-400): a NaN seeded through ResetFromField at layer 9."""
    grid = case_a_grid()
    eng = make_engine(grid)
    eng.SetParams(case_a_params())
    field = np.tile(np.array([1.0, 0.02, 0.0, 0.0], np.float32), grid.CellCount())
    eng.ResetFromField(field)
    assert not eng.ProbeSliceHealth(9)
    bad = field.copy().reshape(24, 24, 48, 4)
    bad[9, 3, 30, 1] = np.nan
    eng.ResetFromField(bad.reshape(-1))
    assert eng.ProbeSliceHealth(9)
    assert not eng.ProbeSliceHealth(2)
    assert not eng.HasDiverged()
    eng.Step(None, 1)
    assert eng.HasDiverged()
    eng.Reset(case_a_params(), True)
    assert not eng.HasDiverged()
    # over-density / over-speed rule: |u|^2 > 4 on a fluid cell
    fast = field.copy().reshape(24, 24, 48, 4)
    fast[4, 5, 6, 1] = 2.5
    eng.ResetFromField(fast.reshape(-1))
    assert eng.ProbeSliceHealth(4)


def test_boundary_links_bitexact():
    """Mask and q-fractions are bit-exact with the reference formula
    q = clamp(phi_f/(phi_f - phi_s), 0, 1) (LbmReference.cpp:104-112) in float32."""
    grid = case_a_grid()
    eng = make_engine(grid)
    cell, d, kind, q = eng.BoundaryLinks()
    from tests.test_oracle_invariants import expected_links
    exp = expected_links(grid, 19)
    got = {(int(c), int(i)): (int(k), np.float32(v).view(np.uint32)) for c, i, k, v in zip(cell, d, kind, q)}
    assert set(got) == set(exp)
    for key, (k, qbits) in exp.items():
        assert got[key][0] == k, key
        if k in (3, 4):
            assert got[key][1] == qbits, key
    assert eng.SurfaceCells() > 0


def test_wing_cross_check_q19_vs_q15():
    """The reference's D3Q19-vs-D3Q15 cross-check (tests/gpu_vs_cpu_test.cpp:149-172).
    At its 80 steps the lift is still in the start-up transient: FlowCraft's own
    LbmReference gives Fz = -0.0203 (D3Q19) there (golden fixture wing_q19_80), which
    the engine reproduces; the invariants drag > 0, lift > 0 and
    |avgSpeed19 - avgSpeed15| < 0.005 are checked once the lift has settled (400 steps)."""
    from tests.test_oracle_reference import _z
    grid = wing_grid()
    p = Params(tau=0.6, inletVelocity=0.05, smagorinskyCs=0.12)
    s = {}
    for model in MODELS:
        e = make_engine(grid, model)
        e.Reset(p, True)
        e.Step(None, 80)
        st = e.ReadbackStats()
        ref = _z[f"wing_q{19 if model == LatticeModel.D3Q19 else 15}_80/stats"]
        assert abs(st.forceX - ref[5]) <= 1e-5 * abs(ref[5])
        assert abs(st.forceZ - ref[7]) <= 1e-5 * abs(ref[5])
        assert abs(st.avgSpeed - ref[3]) <= 1e-5 * ref[3]
        e.Step(None, 320)
        s[model] = e.ReadbackStats()
    for m in MODELS:
        assert not s[m].hasNaN and s[m].forceX > 0 and s[m].forceZ > 0
    assert abs(s[LatticeModel.D3Q19].avgSpeed - s[LatticeModel.D3Q15].avgSpeed) < 0.005


@pytest.mark.parametrize("layout", LAYOUTS, ids=lambda l: l.name)
@pytest.mark.parametrize("storage", [Storage.F32, Storage.F16_SR, Storage.F64], ids=lambda s: s.name)
def test_store_cache_policies_agree(storage, layout, monkeypatch):
    """Cached population stores (lattices that fit in the last-level cache, the default at
    this size) and streamed ones (larger lattices) give identical bits, on a periodic
    lattice (the in-place seam cells store through their own path), and the streamed
    build matches the oracle."""
    grid = small_sphere_grid()
    p = Params(tau=0.6, inletVelocity=0.05, smagorinskyCs=0.12)
    out = {}
    for res in ("1", "0"):
        monkeypatch.setenv("FC_LBM_RESIDENT", res)
        eng = make_engine(grid, LatticeModel.D3Q19, storage, layout, periodic=(True, True))
        eng.Reset(p, True)
        eng.Step(None, 12)
        out[res] = (eng.Populations(), eng.ReadbackField())
        eng.Shutdown()
    assert np.array_equal(out["1"][0], out["0"][0], equal_nan=True)
    orc = make_oracle(grid, LatticeModel.D3Q19, storage, periodic=(True, True))
    oracle_params(orc, p)
    orc.reset(True)
    orc.step(12)
    if storage == Storage.F64:
        fluid = grid.solid == 0
        assert rel_err(out["0"][0][:, fluid], orc.pops()[:, fluid]) <= 1e-12
    else:
        _fields_close(out["0"][1], orc.field(), 1e-5, grid.solid)


@pytest.mark.parametrize("layout", LAYOUTS, ids=lambda l: l.name)
@pytest.mark.parametrize("storage", [Storage.F32, Storage.F16, Storage.F64], ids=lambda s: s.name)
def test_cells_per_thread_variants_agree(storage, layout, monkeypatch):
    """The step kernel's one- and two-cells-per-thread builds give identical bits, on a
    lattice whose y extent (12) is not a multiple of either tile height, and the
    two-cell build matches the oracle."""
    grid = small_sphere_grid()
    p = Params(tau=0.6, inletVelocity=0.05, smagorinskyCs=0.12)
    out = {}
    for cpt in ("1", "2"):
        monkeypatch.setenv("FC_LBM_CELLS_PER_THREAD", cpt)
        eng = make_engine(grid, LatticeModel.D3Q19, storage, layout)
        eng.Reset(p, True)
        eng.Step(None, 12)
        out[cpt] = (eng.Populations(), eng.ReadbackField())
        eng.Shutdown()
    assert np.array_equal(out["1"][0], out["2"][0], equal_nan=True)
    orc = make_oracle(grid, LatticeModel.D3Q19, storage)
    oracle_params(orc, p)
    orc.reset(True)
    orc.step(12)
    if storage == Storage.F64:
        # the float readback field would cap this at float32 rounding
        fluid = grid.solid == 0
        assert rel_err(out["2"][0][:, fluid], orc.pops()[:, fluid]) <= 1e-12
    else:
        _fields_close(out["2"][1], orc.field(), 1e-5, grid.solid)


def _sphere_c1_grid():
    """128x64x64 lattice, sphere of radius 8 cells centred at (32, 32, 32), SDF as in the
    reference test (tests/gpu_vs_cpu_test.cpp:29-41)."""
    from flowcraft_amd import geometry
    return geometry.sphere_grid(128, 64, 64, (32, 32, 32), 8)


C1_PARAMS = Params(tau=0.56, inletVelocity=0.05, smagorinskyCs=0.0)


@pytest.mark.parametrize("layout", LAYOUTS, ids=lambda l: l.name)
def test_sphere_128x64x64_1000_steps_matches_reference(layout):
    """The plain regularized BGK sphere workload (D3Q19, Cs = 0, tau = 0.56, u = 0.05,
    rest start) run 1000 steps in FP32 against the oracle — bit-exact with FlowCraft's
    LbmReference — with the reference's free-slip y/z walls: field within 1e-5 per
    channel, force sums and total mass within 1e-5; the drag coefficient's error
    against the FP64 oracle is printed."""
    grid = _sphere_c1_grid()
    eng, orc = _run_pair(grid, C1_PARAMS, LatticeModel.D3Q19, Storage.F32, layout, steps=1000, seed_mid=False)
    g, o = eng.ReadbackField(), orc.field()
    err = _fields_close(g, o, 1e-5, grid.solid)
    s, so = eng.ReadbackStats(), orc.stats()
    assert abs(s.forceX - so["forceX"]) <= 1e-5 * abs(so["forceX"])
    assert abs(s.totalMass - so["totalMass"]) <= 1e-5 * so["totalMass"]
    o64 = make_oracle(grid, LatticeModel.D3Q19, Storage.F64)
    oracle_params(o64, C1_PARAMS)
    o64.reset(False)
    o64.step(1000)
    s64 = o64.stats()
    q = 0.5 * 0.05 ** 2 * np.pi * 8 ** 2
    print(f"field err vs FP32 oracle {err:.2e}, max ulp {ulp_diff32(g, o)}; Cd {s.forceX / q:.6f} vs FP64 oracle "
          f"{s64['forceX'] / q:.6f} (rel {abs(s.forceX - s64['forceX']) / abs(s64['forceX']):.2e})")
    assert abs(s.forceX - s64["forceX"]) <= 1e-3 * abs(s64["forceX"])


def test_sphere_128x64x64_periodic_1000_steps():
    """The same workload with periodic y/z (an extension: the reference has only walls),
    in place, against the oracle's periodic wrap in FP32 (within 1e-5) and FP64 (drag
    within 1e-3)."""
    grid = _sphere_c1_grid()
    eng, orc = _run_pair(grid, C1_PARAMS, LatticeModel.D3Q19, Storage.F32, Layout.INPLACE, steps=1000,
                         seed_mid=False, periodic=(True, True))
    _fields_close(eng.ReadbackField(), orc.field(), 1e-5, grid.solid)
    o64 = make_oracle(grid, LatticeModel.D3Q19, Storage.F64, periodic=(True, True))
    oracle_params(o64, C1_PARAMS)
    o64.reset(False)
    o64.step(1000)
    f, f64 = eng.Forces()[0], o64.stats()["forceX"]
    print(f"periodic Fx {f:.6e} vs FP64 oracle {f64:.6e}")
    assert abs(f - f64) <= 1e-3 * abs(f64)
