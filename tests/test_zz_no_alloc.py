"""fc_lbm_readback_field_range_noalloc against fc_lbm_readback_field_range, through the C
ABI: the same layers, bit for bit, with the other layers of the output untouched, on two
calls in a row (the second on a smaller lattice and range), and the rejection of a lattice
whose single z layer does not fit the fixed 32 MiB scratch."""
from __future__ import annotations

import numpy as np
import pytest

from flowcraft_amd import _lib, geometry
from flowcraft_amd.lbm import LatticeModel, Layout, LbmSolver, Params, Storage

pytestmark = pytest.mark.gpu

SENTINEL = np.float32(-12345.5)


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu-case] {request.node.nodeid}", flush=True)
    yield


def _solver(grid, storage=Storage.F32, layout=Layout.PINGPONG):
    s = LbmSolver()
    assert s.Init(None, grid, LatticeModel.D3Q19, storage=storage, layout=layout), s.LastError()
    s.Reset(Params(tau=0.6, inletVelocity=0.05, smagorinskyCs=0.12), True)
    s.Step(None, 25)
    return s


def _range(s, fn, z0, zc):
    out = np.full(s.CellCount() * 4, SENTINEL, np.float32)
    rc = fn(s._h, int(z0), int(zc), out.ctypes.data)
    assert rc == _lib.FC_LBM_OK, s._lib.fc_lbm_last_error(s._h).decode()
    return out


def _check_pair(s, z0, zc):
    L = s._lib
    ref = _range(s, L.fc_lbm_readback_field_range, z0, zc)
    got = _range(s, L.fc_lbm_readback_field_range_noalloc, z0, zc)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    layer = s.Nx() * s.Ny() * 4
    lo, hi = max(z0, 0), min(z0 + zc, s.Nz())
    written = np.zeros(s.Nz(), bool)
    if hi > lo:
        written[lo:hi] = True
    per_layer = got.reshape(s.Nz(), layer)
    assert np.all(per_layer[~written] == SENTINEL)
    if hi > lo:
        assert not np.any(per_layer[written] == SENTINEL)
    return got


@pytest.mark.parametrize("layout", [Layout.PINGPONG, Layout.INPLACE], ids=lambda l: l.name)
def test_noalloc_readback_equals_readback_two_calls(layout):
    # first call: 512x256 layers are 2 MiB, so 16 layers per scratch chunk and a 37-layer
    # range from z0 = 5 takes three chunks
    big = _solver(geometry.naca_wing_grid(512, 256, 64, m=0.0, p=0.0, t=0.12, chord=128.0, le_x=128.0,
                                          le_y=128.0, aoa_deg=4.0), layout=layout)
    _check_pair(big, 5, 37)
    # second call, smaller input: another handle on a 48x24x24 sphere, the whole field and
    # a clamped range; the scratch holds the first call's contents at this point
    small = _solver(geometry.sphere_grid(48, 24, 24, (20, 12, 12), 5), storage=Storage.F16, layout=layout)
    whole = _check_pair(small, 0, small.Nz())
    assert np.array_equal(whole.view(np.uint32), small.ReadbackField().view(np.uint32))
    _check_pair(small, -3, 7)
    _check_pair(small, 20, 100)
    _check_pair(small, 30, 4)  # entirely outside: nothing written
    # the first handle again, after the second used the scratch
    _check_pair(big, 60, 4)
    big.Shutdown()
    small.Shutdown()


def test_noalloc_readback_rejects_a_layer_larger_than_the_scratch():
    # one layer of 2048 x 1040 cells is 34,078,720 bytes > 32 MiB
    g = geometry.empty_grid(2048, 1040, 1)
    s = LbmSolver()
    assert s.Init(None, g, LatticeModel.D3Q19, storage=Storage.F16), s.LastError()
    s.Reset(Params(), False)
    out = np.zeros(1, np.float32)
    rc = s._lib.fc_lbm_readback_field_range_noalloc(s._h, 0, 1, out.ctypes.data)
    assert rc == _lib.FC_LBM_EINVAL
    assert b"exceeds the 32 MiB scratch" in s._lib.fc_lbm_last_error(s._h)
    s.Shutdown()
