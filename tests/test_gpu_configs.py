"""GPU parity at the full sizes of the benchmark configurations, against the CPU oracle.

The reference's only harness compares its GPU solver with LbmReference on the same
inputs (tests/gpu_vs_cpu_test.cpp:27-80).  These tests do the same at the lattices the
benchmark configurations name, for a short window of steps that the multithreaded
oracle finishes in seconds:

* the 512x256x256 NACA 0012 wing (chord 128, AoA 4 deg, u 0.08, free-slip y/z): D3Q19
  in FP32 (tau 0.5005, Cs 0.1) against the FP32 oracle, which is bit-exact with
  FlowCraft's LbmReference; FP16 and BF16 offset storage (and FP16 with stochastically
  rounded stores) against the oracle's emulation of the same format; D3Q15 in FP32, FP16
  and stochastically rounded FP16 at tau 0.6 (FlowCraft's own LbmReference diverges with
  D3Q15 at this inlet for tau <= 0.55);
* the 768x512x576 ellipsoid + 64 spheres lattice (226M cells, u 0.06, tau 0.501) in
  FP16 storage against the oracle's emulation;
* in place == ping-pong, bit for bit, for the variants not covered elsewhere.

The runs start from FlowCraft's mid-speed seed (LbmSolver::Reset(params, true),
LbmSolver.cpp:389-427), so the flow meets the body, its Bouzidi links, the walls, the
inlet and the outlet from the first step.  Bar: every field value identical to the
oracle's (fluid (rho, u) and solid (stress, F) channels); the FP64 force sums within
1e-9 relative of the oracle's LbmReference::ComputeStats sums (same terms, another
summation order).

ReadbackFieldRange (FlowCraft/LbmSolver.cpp:584-651) is checked on sub-ranges: z0 > 0,
odd counts, ranges clamped at both ends, empty and out-of-lattice ranges; the covered
layers equal the whole-field readback bit for bit and the other layers of the caller's
buffer are left untouched.
"""
from __future__ import annotations

import numpy as np
import pytest

from flowcraft_amd import geometry
from flowcraft_amd.lbm import LatticeModel, Layout, LbmSolver, Params, Storage
from tests._cases import case_a_grid, case_a_params, make_engine, make_oracle, oracle_params, ulp_diff32

pytestmark = pytest.mark.gpu

# the oracle's threads: the host cores of one GPU's share of the box
ORACLE_THREADS = 16

C2_PARAMS = Params(tau=0.5005, inletVelocity=0.08, smagorinskyCs=0.1)
C2_D3Q15_PARAMS = Params(tau=0.6, inletVelocity=0.08, smagorinskyCs=0.1)
C3_PARAMS = Params(tau=0.501, inletVelocity=0.06, smagorinskyCs=0.1)


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu-case] {request.node.nodeid}", flush=True)
    yield


_GRIDS = {}


def _c2_wing():
    if "c2" not in _GRIDS:
        _GRIDS["c2"] = geometry.naca_wing_grid(512, 256, 256, m=0.0, p=0.0, t=0.12, chord=128.0, le_x=128.0,
                                               le_y=128.0, aoa_deg=4.0)
    return _GRIDS["c2"]


def _c3_lattice():
    return geometry.ellipsoid_and_spheres_grid(768, 512, 576, semi_axes=(96, 40, 40), n_spheres=64, r_range=(4, 12),
                                               seed=0)


def _engine_field(s):
    import torch
    out = torch.empty(s.CellCount() * 4, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    s.FieldDevice(out)
    s.Sync()
    return out.cpu().numpy()


def _compare_with_oracle(grid, model, storage, params, steps, layout=Layout.PINGPONG):
    eng = make_engine(grid, model, storage, layout)
    eng.Reset(params, True)
    eng.Step(None, steps)
    got = _engine_field(eng)
    forces = eng.Forces()
    eng.Shutdown()
    del eng
    orc = make_oracle(grid, model, storage, nthreads=ORACLE_THREADS)
    oracle_params(orc, params)
    orc.reset(True)
    orc.step(steps)
    ref = orc.field()
    st = orc.stats()
    diverged = orc.diverged()
    del orc
    assert not diverged
    assert np.isfinite(got).all()
    same = got == ref
    n_diff = int(same.size - same.sum())
    print(f"{model.name} {storage.name} {grid.nx}x{grid.ny}x{grid.nz}, {steps} steps: {n_diff} of {same.size} "
          f"field values differ, max ulp {ulp_diff32(got, ref)}; forces engine {forces[:3]} "
          f"oracle {(st['forceX'], st['forceY'], st['forceZ'])}", flush=True)
    assert n_diff == 0
    for k, name in enumerate(("forceX", "forceY", "forceZ")):
        assert abs(forces[k] - st[name]) <= 1e-9 * max(abs(st["forceX"]), 1e-12), (name, forces[k], st[name])


@pytest.mark.parametrize("storage", [Storage.F32, Storage.F16, Storage.BF16, Storage.F16_SR], ids=lambda s: s.name)
def test_c2_wing_d3q19_matches_oracle(storage):
    _compare_with_oracle(_c2_wing(), LatticeModel.D3Q19, storage, C2_PARAMS, steps=20)


@pytest.mark.parametrize("storage", [Storage.F32, Storage.F16, Storage.F16_SR], ids=lambda s: s.name)
def test_c2_wing_d3q15_matches_oracle(storage):
    _compare_with_oracle(_c2_wing(), LatticeModel.D3Q15, storage, C2_D3Q15_PARAMS, steps=20)


def test_c3_fp16_matches_oracle():
    """226M cells: 8.6 GB of FP16 populations per copy on the device; the oracle holds
    ~36 GB of host memory."""
    _compare_with_oracle(_c3_lattice(), LatticeModel.D3Q19, Storage.F16, C3_PARAMS, steps=3)


def _inplace_vs_pingpong(grid, model, storage, params, steps):
    import torch
    fields, forces = [], []
    for layout in (Layout.INPLACE, Layout.PINGPONG):
        s = make_engine(grid, model, storage, layout)
        s.Reset(params, True)
        s.Step(None, steps)
        out = torch.empty(s.CellCount() * 4, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        s.FieldDevice(out)
        s.Sync()
        fields.append(out)
        forces.append(s.Forces())
        assert not s.HasDiverged()
        s.Shutdown()
        del s
    assert torch.equal(fields[0].view(torch.int32), fields[1].view(torch.int32))
    assert np.array_equal(forces[0], forces[1])
    assert bool(torch.isfinite(fields[0]).all())


@pytest.mark.parametrize("model,storage", [(LatticeModel.D3Q15, Storage.F32), (LatticeModel.D3Q15, Storage.F16),
                                           (LatticeModel.D3Q19, Storage.BF16)],
                         ids=["D3Q15_F32", "D3Q15_F16", "D3Q19_BF16"])
def test_c2_wing_inplace_equals_pingpong(model, storage):
    params = C2_D3Q15_PARAMS if model == LatticeModel.D3Q15 else C2_PARAMS
    _inplace_vs_pingpong(_c2_wing(), model, storage, params, steps=30)


def test_c3_fp16_inplace_equals_pingpong():
    _inplace_vs_pingpong(_c3_lattice(), LatticeModel.D3Q19, Storage.F16, C3_PARAMS, steps=12)


SENTINEL = np.float32(-7777.25)
RANGES = [(0, 24), (5, 7), (23, 1), (1, 1), (-4, 9), (20, 100), (-3, 40), (24, 3), (30, 2), (3, 0), (6, -2)]


@pytest.mark.parametrize("layout", [Layout.PINGPONG, Layout.INPLACE], ids=lambda l: l.name)
@pytest.mark.parametrize("storage", [Storage.F32, Storage.F16], ids=lambda s: s.name)
def test_readback_field_range_subranges(storage, layout):
    grid = case_a_grid()
    s = make_engine(grid, LatticeModel.D3Q19, storage, layout)
    s.Reset(case_a_params(), True)
    s.Step(None, 30)
    whole = s.ReadbackField().reshape(grid.nz, -1)
    nz = grid.nz
    for z0, zc in RANGES:
        out = np.full(grid.CellCount() * 4, SENTINEL, np.float32)
        s.ReadbackFieldRange(z0, zc, out)
        out = out.reshape(nz, -1)
        lo, hi = max(z0, 0), min(z0 + zc, nz)
        covered = np.zeros(nz, bool)
        if zc > 0 and hi > lo:
            covered[lo:hi] = True
        assert np.array_equal(out[covered].view(np.uint32), whole[covered].view(np.uint32)), (z0, zc)
        assert np.all(out[~covered] == SENTINEL), (z0, zc)
    # a buffer of the wrong size is left alone (LbmSolver.h's presized-vector contract)
    short = np.full(16, SENTINEL, np.float32)
    s.ReadbackFieldRange(0, 4, short)
    assert np.all(short == SENTINEL)
    s.Shutdown()


def test_readback_field_range_ragged_lattice():
    """A lattice whose x extent is not a whole number of 64-cell blocks and whose z extent
    is odd: every single-layer range and a sweep of 16-layer chunks (FlowCraft's paced
    readback, main.cpp:2506-2555) reassemble the whole field."""
    grid = geometry.sphere_grid(130, 20, 37, (40, 10, 18), 6)
    s = make_engine(grid, LatticeModel.D3Q19, Storage.F32, Layout.INPLACE)
    s.Reset(Params(tau=0.6, inletVelocity=0.05, smagorinskyCs=0.1), True)
    s.Step(None, 17)
    whole = s.ReadbackField()
    per_layer = np.full_like(whole, SENTINEL)
    for z in range(grid.nz):
        s.ReadbackFieldRange(z, 1, per_layer)
    chunked = np.full_like(whole, SENTINEL)
    for z in range(0, grid.nz, 16):
        s.ReadbackFieldRange(z, 16, chunked)
    assert np.array_equal(per_layer.view(np.uint32), whole.view(np.uint32))
    assert np.array_equal(chunked.view(np.uint32), whole.view(np.uint32))
    s.Shutdown()
