"""FlowCraft's mesh voxelizer, Voxelizer::Voxelize (FlowCraft/Voxelizer.cpp:128-432):
the oracle's restatement (oracle/vox_oracle.cpp) pinned to FlowCraft's own function
compiled into oracle/_ref — solid mask, signed distance and grid frame bit for bit — and
to geometry (a sphere's signed distance, an enclosed cavity, a body touching the seed
corner); plus the device library's host-side frame and error behaviour (no GPU call).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np
import pytest

import oracle
from flowcraft_amd import _lib, meshes

needs_ref = pytest.mark.skipif(not oracle.ref_voxelizer_available(),
                               reason="oracle/_ref/libref_voxelizer.so is not built (make -C oracle ref)")
STL = "/root/reference/HighWingModel.stl"


def _hollow_sphere():
    """Outer shell r 20 (outward) around an inner shell r 12 wound inward: the gap is
    solid material, the inner void an enclosed fluid pocket."""
    outer = meshes.ellipsoid((0.0, 0.0, 0.0), (20.0, 20.0, 20.0), 48, 24)
    ip, ii = meshes.ellipsoid((0.0, 0.0, 0.0), (12.0, 12.0, 12.0), 40, 20)
    ii = ii.reshape(-1, 3)[:, [0, 2, 1]].reshape(-1)  # reversed winding: normals point into the void
    return meshes.merge(outer, (ip, ii))


def _with_degenerates(m):
    pos, idx = m
    extra = np.array([[1.0, 2.0, 3.0], [1.0, 2.0, 3.0], [4.0, 5.0, 6.0]], np.float32)  # zero-area triangle
    pos2 = np.concatenate([pos, extra])
    n = len(pos)
    idx2 = np.concatenate([idx[:300], np.array([n, n + 1, n + 2, n, n, n], np.uint32), idx[300:]])
    return pos2, idx2


CASES = {
    "sphere": (lambda: meshes.ellipsoid((1.5, -2.25, 0.7), (10.0, 10.0, 10.0), 32, 16), 48, None),
    "wing_aoa8": (lambda: meshes.wing(0.04, 0.4, 0.15, 100.0, 300.0, (0.0, 0.0, 0.0), "x", 8.0), 96, None),
    "fin": (lambda: meshes.wing(0.0, 0.0, 0.1, 80.0, 120.0, (5.0, 0.0, 0.0), "z", 0.0, 48, 24), 64, None),
    "aircraft": (lambda: meshes.aircraft(), 160, None),
    "hollow": (_hollow_sphere, 64, None),
    "degenerate": (lambda: _with_degenerates(meshes.ellipsoid((0.0, 0.0, 0.0), (8.0, 5.0, 3.0), 24, 12)), 40,
                   None),
    "no_padding": (lambda: meshes.ellipsoid((0.0, 0.0, 0.0), (8.0, 8.0, 8.0), 24, 12), 24, (0.0, 0.0, 0.0)),
    "seed_in_shell": (lambda: _corner_bead(), 24, (0.0, 0.0, 0.0)),
}


def _args(name):
    mk, res, pad = CASES[name]
    return mk(), res, pad if pad is not None else oracle.VOX_DEFAULT_PADDING


def _assert_same(a, b):
    assert a[0].shape == b[0].shape
    assert a[2] == b[2] and a[3] == b[3]
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))


@needs_ref
@pytest.mark.parametrize("name", sorted(CASES))
def test_restatement_equals_flowcraft_voxelizer(name):
    (pos, idx), res, pad = _args(name)
    _assert_same(oracle.voxelize(pos, idx, res, pad), oracle.ref_voxelize(pos, idx, res, pad))


@needs_ref
@pytest.mark.skipif(not os.path.exists(STL), reason="FlowCraft's sample STL is not present")
def test_restatement_equals_flowcraft_on_its_sample_stl():
    pos, idx = meshes.read_stl(STL)
    _assert_same(oracle.voxelize(pos, idx, 96), oracle.ref_voxelize(pos, idx, 96))


@needs_ref
def test_empty_mesh_gives_empty_grid():
    for pos, idx in ((np.zeros((0, 3), np.float32), np.zeros(0, np.uint32)),
                     (np.zeros((3, 3), np.float32), np.array([0, 1], np.uint32))):
        a, b = oracle.voxelize(pos, idx, 64), oracle.ref_voxelize(pos, idx, 64)
        assert a[0].size == 0 and b[0].size == 0


def test_sphere_signed_distance_matches_geometry():
    """Cells of the surface shell (within the conservative radius sqrt(3)/2 cells, where
    the nearest triangle is always among those visited) carry the distance to the
    tessellated sphere: within the chord sagitta of the true |p - c| - r.  Cells no
    triangle visits keep +-inf with the sign of their classification."""
    r, c = 10.0, np.array([1.5, -2.25, 0.7])
    pos, idx = meshes.ellipsoid(tuple(c), (r, r, r), 96, 48)
    solid, sdf, cs, org = oracle.voxelize(pos, idx, 64)
    nz, ny, nx = solid.shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    p = np.stack([org[0] + (y + 0.5) * cs, org[1] + (x + 0.5) * cs, org[2] + (z + 0.5) * cs], -1)
    true = np.linalg.norm(p - c, axis=-1) - r
    sag = r * (1 - np.cos(np.pi / 48))
    band = np.abs(sdf) <= 0.8660254 * cs
    assert band.sum() > 1000
    assert np.abs(sdf[band] - true[band]).max() <= sag + 1e-4 * r
    assert (solid[true < -cs] == 1).all() and (solid[true > cs] == 0).all()
    far = ~np.isfinite(sdf)
    assert far.any() and (np.sign(sdf[far]) == np.where(solid[far] == 1, -1, 1)).all()


def test_enclosed_cavity_becomes_solid():
    solid, sdf, cs, org = oracle.voxelize(*_hollow_sphere(), 64)
    nz, ny, nx = solid.shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    p = np.stack([org[0] + (y + 0.5) * cs, org[1] + (x + 0.5) * cs, org[2] + (z + 0.5) * cs], -1)
    rr = np.linalg.norm(p, axis=-1)
    shell = 0.8660254 * cs  # the conservative radius of the surface shell
    assert (solid[rr < 12.0 - shell] == 1).all()  # the void is unreachable from the exterior
    assert (solid[rr > 20.0 + shell + 1e-3] == 0).all()
    assert (solid[(rr > 12.0 + shell) & (rr < 20.0 - shell)] == 1).all()  # the wall between the shells


def _corner_bead():
    """A sphere r 8 and, at the low corner of the bounds, a bead r 0.3: with no padding
    the grid starts at the bead, so cell (0,0,0) lies in its surface shell."""
    return meshes.merge(meshes.ellipsoid((0.0, 0.0, 0.0), (8.0, 8.0, 8.0), 24, 12),
                        meshes.ellipsoid((-9.0, -9.0, -9.0), (0.3, 0.3, 0.3), 8, 4))


def test_body_on_the_seed_corner_fills_the_grid():
    """Cell (0,0,0) in the surface shell: no cell is reached from the seed, so every cell
    ends solid (Voxelizer.cpp:386-395), though most lie outside both bodies."""
    solid, _, _, _ = oracle.voxelize(*_corner_bead(), 24, (0.0, 0.0, 0.0))
    assert solid.all()
    solid, _, _, _ = oracle.voxelize(*_corner_bead(), 24)  # default padding: the corner is fluid
    assert solid[0, 0, 0] == 0 and solid.mean() < 0.2


def test_device_frame_matches_restatement_without_a_gpu():
    L = _lib.load()
    for name in sorted(CASES):
        (pos, idx), res, pad = _args(name)
        p = _lib.FcVoxParams()
        p.target_resolution = res
        for k in range(3):
            p.padding[k] = pad[k]
        g = _lib.FcVoxGrid()
        pos = np.ascontiguousarray(pos, np.float32)
        idx = np.ascontiguousarray(idx, np.uint32)
        assert L.fc_voxelize_frame(pos.ctypes.data, len(pos), idx.ctypes.data, len(idx), ctypes.byref(p),
                                   ctypes.byref(g)) == _lib.FC_LBM_OK
        dims, cs, org = oracle.vox_frame(pos, idx, res, pad)
        assert (g.nx, g.ny, g.nz) == dims and g.cell_size == np.float32(cs)
        assert tuple(np.float32(v) for v in g.origin) == tuple(np.float32(v) for v in org)


def test_device_entry_points_reject_bad_input_without_a_gpu():
    L = _lib.load()
    d = _lib.FcVoxParams()
    assert L.fc_vox_params_default(ctypes.byref(d)) == _lib.FC_LBM_OK
    assert d.target_resolution == 640 and list(d.padding) == [np.float32(0.3), 1.0, 0.5]
    pos = np.zeros((3, 3), np.float32)
    bad = np.array([0, 1, 7], np.uint32)
    g = _lib.FcVoxGrid()
    assert L.fc_voxelize_frame(pos.ctypes.data, 3, bad.ctypes.data, 3, ctypes.byref(d), ctypes.byref(g)) == \
        _lib.FC_LBM_EINVAL
    assert b"out of range" in L.fc_voxelize_last_error()
    assert L.fc_voxelize_frame(pos.ctypes.data, 3, bad.ctypes.data, 3, None, ctypes.byref(g)) == _lib.FC_LBM_EINVAL
    # an empty mesh is the reference's empty grid and never reaches the device
    assert L.fc_voxelize(pos.ctypes.data, 0, bad.ctypes.data, 0, ctypes.byref(d), None, None) == _lib.FC_LBM_OK
    assert L.fc_voxelize_frame(pos.ctypes.data, 0, bad.ctypes.data, 0, ctypes.byref(d), ctypes.byref(g)) == \
        _lib.FC_LBM_OK and (g.nx, g.ny, g.nz) == (0, 0, 0)
