"""Device smoke trails (fc_smoke_*, csrc/smoke.hip) against the oracle's restatement of
FlowCraft's SmokeTrailSystem (pinned trail for trail to FlowCraft's own class in
tests/test_smoke.py): the same scripted sessions, bit for bit, with the field given as
a device tensor and with the field taken from the solver's current state."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from flowcraft_amd.smoke import PlaneConfinement, SmokeTrailSystem
from tests._cases import case_a_grid, case_a_params, make_engine
from tests.test_smoke import flow_case, session, synthetic_case

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu-case] {request.node.nodeid}", flush=True)
    yield


class DeviceSmoke:
    """The oracle front end's interface over the device system."""

    def __init__(self, solid, cs, org):
        from flowcraft_amd.lbm import VoxelGrid
        nz, ny, nx = solid.shape
        self.grid = VoxelGrid(nx, ny, nz, solid, None, cs, org)
        self.sys = SmokeTrailSystem(self.grid)
        self.dfield = None

    @staticmethod
    def _c(confine):
        return None if confine is None else PlaneConfinement(*confine)

    def reset(self, rev, count, trail, confine=None):
        self.sys.Reset(self.grid, rev, count, trail, self._c(confine))

    def scatter(self, confine=None):
        self.sys.ScatterRandom(self.grid, self._c(confine))

    def shift_plane(self, confine):
        self.sys.ShiftPlane(self.grid, self._c(confine))

    def step(self, field, rev, mmps, dt, suspend=False, confine=None):
        import torch
        if self.dfield is None:
            self.dfield = torch.from_numpy(np.ascontiguousarray(field, np.float32)).cuda()
            torch.cuda.synchronize()
        self.sys.Step(self.grid, self.dfield, rev, mmps, dt, suspend, self._c(confine))

    def trails(self):
        return self.sys.Trails()


@pytest.mark.parametrize("rev", [False, True], ids=["forward", "reverse"])
@pytest.mark.parametrize("case", ["synthetic", "flow"])
def test_device_equals_restatement(case, rev):
    solid, field, cs, org = synthetic_case() if case == "synthetic" else flow_case()
    a, b = DeviceSmoke(solid, cs, org), oracle.OracleSmoke(solid, cs, org)
    for la, lb in zip(session(a, field, rev), session(b, field, rev)):
        assert la == lb
        ta, na = a.trails()
        tb, nb = b.trails()
        assert np.array_equal(na, nb), la
        for p in range(len(na)):  # slots past a trail's length are unspecified on the device
            assert np.array_equal(ta[p, :na[p]].view(np.uint32), tb[p, :nb[p]].view(np.uint32)), (la, p)


def test_device_steps_through_the_solver_state():
    g = case_a_grid()
    eng = make_engine(g)
    eng.Reset(case_a_params(), True)
    dev = SmokeTrailSystem(g)
    orc = oracle.OracleSmoke(g.solid, g.cellSize, g.origin)
    dev.Reset(g, False, 500, 12)
    orc.reset(False, 500, 12)
    for k in range(10):
        eng.Step(None, 5)
        dev.Step(g, eng, False, 300.0, 0.05)
        orc.step(eng.ReadbackField(), False, 300.0, 0.05)
    ta, na = dev.Trails()
    tb, nb = orc.trails()
    assert np.array_equal(na, nb) and na.max() == 11
    for p in range(len(na)):
        assert np.array_equal(ta[p, :na[p]].view(np.uint32), tb[p, :nb[p]].view(np.uint32))
    eng.Shutdown()


def test_field_source_is_checked():
    """A field that is not a contiguous float32 device tensor of 4 values per cell is
    refused before anything reaches the device."""
    import torch
    g = case_a_grid()
    dev = SmokeTrailSystem(g)
    dev.Reset(g, False, 50, 4)
    n = g.nx * g.ny * g.nz * 4
    for bad in (torch.zeros(n), torch.zeros(n - 4, device="cuda"), torch.zeros(n, dtype=torch.float64, device="cuda"),
                torch.zeros(2 * n, device="cuda")[::2]):
        with pytest.raises(ValueError):
            dev.Step(g, bad, False, 300.0, 0.05)
    dev.Step(g, torch.zeros(n, device="cuda"), False, 300.0, 0.05)  # a still field: accepted
    _, lengths = dev.Trails()
    assert (lengths == 2).all()


@pytest.mark.parametrize("layout,storage", [("PINGPONG", "F32"), ("INPLACE", "F32"), ("PINGPONG", "F16")])
def test_field_at_equals_whole_field(layout, storage):
    """fc_lbm_field_at (what the smoke step gathers) gives, bit for bit, the whole field's
    values at the listed cells for fluid cells and zeros for solid ones."""
    import torch
    from flowcraft_amd.lbm import Layout, Storage
    g = case_a_grid()
    eng = make_engine(g, storage=Storage[storage], layout=Layout[layout])
    eng.Reset(case_a_params(), True)
    eng.Step(None, 7)
    N = g.nx * g.ny * g.nz
    whole = torch.empty(N * 4, device="cuda")
    eng.FieldDevice(whole)
    rng = np.random.default_rng(5)
    cells = np.concatenate([rng.integers(0, N, 3000), np.flatnonzero(g.solid.reshape(-1))[:50], [0, N - 1],
                            [-1, N, N + 12345]])  # the last three are outside the lattice: zeros
    dc = torch.from_numpy(cells.astype(np.int32)).cuda()
    out = torch.empty(len(cells) * 4, device="cuda")
    eng.FieldAt(dc, out)
    eng.Sync()
    got = out.view(-1, 4).cpu().numpy()
    assert (got[-3:] == 0).all()
    got, cells = got[:-3], cells[:-3]
    ref = whole.view(-1, 4).cpu().numpy()[cells]
    solid = g.solid.reshape(-1)[cells] == 1
    assert solid.any() and (~solid).any()
    assert np.array_equal(got[~solid].view(np.uint32), ref[~solid].view(np.uint32))
    assert (got[solid] == 0).all()
    eng.Shutdown()
