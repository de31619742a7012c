"""Host-side argument checks of the bindings, run without a GPU: each bad input is
rejected before any device pointer is used."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from flowcraft_amd import _lib, meshes
from flowcraft_amd.lbm import check_field_at_tensors, check_geometry_tensors
from flowcraft_amd.voxelizer import Params as VParams, Voxelizer, _arrays

N = 4 * 3 * 2


@pytest.mark.parametrize("solid,sdf,msg", [
    (torch.zeros(N, dtype=torch.int64), None, "dtype"),
    (torch.zeros(N, dtype=torch.float32), None, "dtype"),
    (torch.zeros(N - 1, dtype=torch.uint8), None, "must hold 24"),
    (torch.zeros(2 * N, dtype=torch.uint8)[::2], None, "contiguous"),
    (torch.zeros(N, dtype=torch.uint8), None, "GPU tensor"),
    (torch.zeros(N, dtype=torch.bool), None, "GPU tensor"),
    (np.zeros(N, np.uint8), None, "torch tensor"),
    (torch.zeros(N, dtype=torch.uint8), torch.zeros(N, dtype=torch.float64), "dtype"),
    (torch.zeros(N, dtype=torch.uint8), torch.zeros(N + 1, dtype=torch.float32), "must hold 24"),
])
def test_geometry_tensors_rejected(solid, sdf, msg):
    with pytest.raises(ValueError, match=msg):
        check_geometry_tensors(N, solid, sdf)


@pytest.mark.parametrize("cells,out,msg", [
    (torch.zeros(8, dtype=torch.int64), torch.zeros(32), "dtype"),
    (torch.zeros(8, dtype=torch.int32), torch.zeros(32, dtype=torch.float64), "dtype"),
    (torch.zeros(8, dtype=torch.int32), torch.zeros(31), "at least 32"),
    (torch.zeros(16, dtype=torch.int32)[::2], torch.zeros(32), "contiguous"),
    (torch.zeros(8, dtype=torch.int32), torch.zeros(64)[::2], "contiguous"),
    (torch.zeros(8, dtype=torch.int32), torch.zeros(32), "GPU tensor"),
])
def test_field_at_tensors_rejected(cells, out, msg):
    with pytest.raises(ValueError, match=msg):
        check_field_at_tensors(cells, out)


def test_device_voxelizer_rejects_small_output_buffers():
    """fc_voxelize_device compares the buffers' capacity with the frame on the host and
    fails before any device work, so the (never dereferenced) addresses here are safe."""
    pos, idx = _arrays(*meshes.wing(0.04, 0.4, 0.15, 100.0, 300.0, (0.0, 0.0, 0.0), "x", 8.0))
    vp = VParams(64)
    (nx, ny, nz), _, _ = Voxelizer.Frame(pos, idx, vp)
    L = _lib.load()
    for cap in (0, nx * ny * nz - 1):
        rc = L.fc_voxelize_device(pos.ctypes.data, len(pos), idx.ctypes.data, len(idx), ctypes.byref(vp._c()),
                                  ctypes.c_void_p(256), ctypes.c_void_p(512), cap)
        assert rc == _lib.FC_LBM_EINVAL
        assert f"output buffers hold {cap} cells, the frame has {nx * ny * nz}".encode() in L.fc_voxelize_last_error()


def _create(nx, ny, nz, storage=_lib.STORE_F16, layout=_lib.LAYOUT_INPLACE):
    L = _lib.load()
    desc = _lib.FcLbmDesc(nx, ny, nz, _lib.D3Q19, storage, layout, 0, 0, None)
    h = ctypes.c_void_p()
    rc = L.fc_lbm_create(ctypes.byref(desc), ctypes.byref(h))
    msg = (L.fc_lbm_last_error(h) or b"").decode()
    L.fc_lbm_destroy(h)
    return rc, msg


def test_lattice_size_limit_is_checked_before_the_device():
    """The engine's lattice limit: at most 2^32 - 2^20 cells nx*ny*nz (one work-item per
    cell in one 32-bit grid dimension; cell and padded-cell indices are 64-bit), and three
    padded z layers of populations under 2 GiB (32-bit offsets from a per-workgroup base).
    At this cross-section nz = 2047 (2^32 - 2^21 cells, well past 2^31 padded cells) is the
    last that fits; one more layer is refused with EINVAL before any device call."""
    assert 2048 * 1024 * 2047 <= (1 << 32) - (1 << 20) < 2048 * 1024 * 2048
    assert 2048 * (1024 + 2) * (2047 + 2) > (1 << 31)
    rc, msg = _create(2048, 1024, 2048)
    assert rc == _lib.FC_LBM_EINVAL and "lattice too large (at most 4293918720 cells)" in msg
    rc, msg = _create(65536, 8192, 1)
    assert rc == _lib.FC_LBM_EINVAL and "cross-section too large" in msg
    if torch.cuda.device_count() == 0:
        # the size checks pass; without a device the next step (the device query) fails
        rc, msg = _create(2048, 1024, 2047)
        assert rc == _lib.FC_LBM_EDEVICE, (rc, msg)
