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
