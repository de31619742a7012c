"""Shared parity-case helpers: the same seeded inputs driven through the HIP
engine (flowcraft_amd, via libfc_lbm.so) and through the CPU oracle."""
from __future__ import annotations

import numpy as np

from flowcraft_amd import geometry
from flowcraft_amd.lbm import LatticeModel, Layout, LbmSolver, Params, Storage, VoxelGrid

STORAGE_TO_ORACLE = {Storage.F32: (32, 0), Storage.F16: (32, 1), Storage.BF16: (32, 2), Storage.F64: (64, 0),
                     Storage.F16_SR: (32, 3), Storage.BF16_SR: (32, 4)}


def case_a_grid() -> VoxelGrid:
    """tests/gpu_vs_cpu_test.cpp:28-41: 48x24x24, sphere r=5 at (20,12,12)."""
    return geometry.sphere_grid(48, 24, 24, (20, 12, 12), 5)


def case_a_params() -> Params:
    """tests/gpu_vs_cpu_test.cpp:43: tau 0.6, u 0.05, Cs 0.12, defaults otherwise."""
    return Params(tau=0.6, inletVelocity=0.05, smagorinskyCs=0.12)


def small_sphere_grid(with_sdf=True) -> VoxelGrid:
    g = geometry.sphere_grid(16, 12, 12, (7.0, 6.0, 6.0), 2.6)
    if not with_sdf:
        g.signedDistance = None
    return g


def wing_grid() -> VoxelGrid:
    """tests/gpu_vs_cpu_test.cpp:84-105: inclined ellipsoid wing, 48x32x24, +10 deg, SDF x3."""
    nx, ny, nz = 48, 32, 24
    cx, cy, cz = np.float32(20), np.float32(16), np.float32(12)
    ca, sa = np.float32(np.cos(np.float32(0.1745))), np.float32(np.sin(np.float32(0.1745)))
    z, y, x = np.meshgrid(np.arange(nz, dtype=np.float32), np.arange(ny, dtype=np.float32),
                          np.arange(nx, dtype=np.float32), indexing="ij")
    h = np.float32(0.5)
    dx, dy, dz = x + h - cx, y + h - cy, z + h - cz
    rx = dx * ca - dz * sa
    rz = dx * sa + dz * ca
    d = (np.sqrt((rx / np.float32(7)) * (rx / np.float32(7)) + (dy / np.float32(9)) * (dy / np.float32(9)) +
                 (rz / np.float32(2.2)) * (rz / np.float32(2.2))) - np.float32(1)) * np.float32(3)
    d = d.astype(np.float32)
    return VoxelGrid(nx, ny, nz, (d < 0).astype(np.uint8), d)


def make_engine(grid: VoxelGrid, model=LatticeModel.D3Q19, storage=Storage.F32, layout=Layout.PINGPONG,
                periodic=(False, False)) -> LbmSolver:
    s = LbmSolver()
    ok = s.Init(None, grid, model, storage=storage, layout=layout, periodicY=periodic[0], periodicZ=periodic[1])
    assert ok, s.LastError()
    return s


def make_oracle(grid: VoxelGrid, model=LatticeModel.D3Q19, storage=Storage.F32, periodic=(False, False),
                bits=None, nthreads=0):
    from oracle import OracleLbm
    b, st = STORAGE_TO_ORACLE[Storage(storage)]
    if bits is not None:
        b = bits
        st = 0 if bits == 64 else st
    q = 15 if model == LatticeModel.D3Q15 else 19
    return OracleLbm(grid.solid, grid.signedDistance, q=q, bits=b, storage=st, periodic_y=periodic[0],
                     periodic_z=periodic[1], nthreads=nthreads)


def oracle_params(o, p: Params):
    o.set_params(p.tau, p.inletVelocity, p.reverseFlow, p.freeSlipY, p.freeSlipZ, p.smagorinskyCs)


def ulp_diff32(a: np.ndarray, b: np.ndarray) -> int:
    """Max distance in float32 units-in-the-last-place (monotone integer mapping)."""
    ia = a.astype(np.float32).view(np.int32).astype(np.int64)
    ib = b.astype(np.float32).view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7FFFFFFF), ia)
    ib = np.where(ib < 0, -(ib & 0x7FFFFFFF), ib)
    return int(np.abs(ia - ib).max()) if ia.size else 0


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    """max |a-b| / max(|b|, 1e-30) over the array (relative to the field scale)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(float(np.abs(b).max()), 1e-30)
    return float(np.abs(a - b).max()) / scale


# Lattices at the edges of the layout: x not a multiple of the 64-cell block (one partial
# block, a full block plus one cell), the minimum nx = 4, one-cell and two-cell y/z
# extents where a cell touches both walls, obstacles next to the walls.  Each entry:
# (grid builder, free-slip y, free-slip z).
RAGGED_CASES = {
    "100x3x5_sphere": (lambda: geometry.sphere_grid(100, 3, 5, (30, 1, 2), 1.5), True, True),
    "70x1x1_channel": (lambda: geometry.sphere_grid(70, 1, 1, (-50, 0, 0), 1.0), True, True),
    "4x2x3_noslip": (lambda: geometry.sphere_grid(4, 2, 3, (-50, 0, 0), 1.0), False, False),
    "129x5x2_mixed": (lambda: geometry.sphere_grid(129, 5, 2, (64, 2, 1), 1.8), True, False),
    "65x7x1_sphere_on_wall": (lambda: geometry.sphere_grid(65, 7, 1, (20, 0, 0), 2.5), False, True),
}


def ragged_params(fsy: bool, fsz: bool) -> Params:
    return Params(tau=0.6, inletVelocity=0.05, smagorinskyCs=0.12, freeSlipY=fsy, freeSlipZ=fsz)
