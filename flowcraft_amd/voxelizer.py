"""Python mirror of FlowCraft's mesh voxelizer (FlowCraft/Voxelizer.h:11-53), running on
the device through libfc_lbm.so (``fc_voxelize``, csrc/voxelize.hip).

``Voxelizer.Voxelize(positions, indices, params)`` takes world-space vertex positions
(float32 [n, 3]) and triangle indices (uint32, 3 per triangle, outward winding) and
returns the ``VoxelGrid`` the solver reads: solid mask and signed distance in the
reference's conventions, identical to FlowCraft's own ``Voxelizer::Voxelize``.  An
empty mesh gives an empty grid, as there; a bad index raises.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .lbm import VoxelGrid


@dataclass
class Params:
    """Voxelizer::Params (FlowCraft/Voxelizer.h:13-30), same defaults."""
    targetResolution: int = 640
    paddingFactor: tuple = (0.3, 1.0, 0.5)

    def _c(self) -> L.FcVoxParams:
        p = L.FcVoxParams()
        p.target_resolution = int(self.targetResolution)
        for k in range(3):
            p.padding[k] = float(self.paddingFactor[k])
        return p


class Voxelizer:
    @staticmethod
    def Frame(positions, indices, params: Params = Params()):
        """(nx, ny, nz), cellSize, origin of the grid Voxelize lays over the mesh."""
        lib = L.load()
        pos, idx = _arrays(positions, indices)
        g = L.FcVoxGrid()
        rc = lib.fc_voxelize_frame(pos.ctypes.data, len(pos), idx.ctypes.data, len(idx), ctypes.byref(params._c()),
                                   ctypes.byref(g))
        if rc != L.FC_LBM_OK:
            raise L.FcLbmError(rc, lib.fc_voxelize_last_error().decode())
        return (g.nx, g.ny, g.nz), float(g.cell_size), tuple(float(v) for v in g.origin)

    @staticmethod
    def Voxelize(positions, indices, params: Params = Params()) -> VoxelGrid:
        """Voxelizer::Voxelize (FlowCraft/Voxelizer.cpp:128-432) on the device."""
        lib = L.load()
        pos, idx = _arrays(positions, indices)
        (nx, ny, nz), cs, origin = Voxelizer.Frame(pos, idx, params)
        if nx * ny * nz == 0:
            return VoxelGrid(0, 0, 0, np.zeros((0, 0, 0), np.uint8), None, 1.0, (0.0, 0.0, 0.0))
        solid = np.empty((nz, ny, nx), np.uint8)
        sdf = np.empty((nz, ny, nx), np.float32)
        rc = lib.fc_voxelize(pos.ctypes.data, len(pos), idx.ctypes.data, len(idx), ctypes.byref(params._c()),
                             solid.ctypes.data, sdf.ctypes.data)
        if rc != L.FC_LBM_OK:
            raise L.FcLbmError(rc, lib.fc_voxelize_last_error().decode())
        return VoxelGrid(nx, ny, nz, solid, sdf, cs, origin)


    @staticmethod
    def VoxelizeDevice(positions, indices, params: Params = Params()):
        """Extension: the same grid left in HBM, for LbmSolver.SetGeometryDevice:
        ((nx, ny, nz), cellSize, origin, solid uint8 tensor [nz, ny, nx], signed distance
        float32 tensor [nz, ny, nx]), both on the current device."""
        import torch
        lib = L.load()
        pos, idx = _arrays(positions, indices)
        (nx, ny, nz), cs, origin = Voxelizer.Frame(pos, idx, params)
        solid = torch.empty((nz, ny, nx), dtype=torch.uint8, device="cuda")
        sdf = torch.empty((nz, ny, nx), dtype=torch.float32, device="cuda")
        if nx * ny * nz:
            torch.cuda.current_stream().synchronize()  # the buffers exist before the library writes them
            rc = lib.fc_voxelize_device(pos.ctypes.data, len(pos), idx.ctypes.data, len(idx),
                                        ctypes.byref(params._c()), solid.data_ptr(), sdf.data_ptr(), solid.numel())
            if rc != L.FC_LBM_OK:
                raise L.FcLbmError(rc, lib.fc_voxelize_last_error().decode())
        return (nx, ny, nz), cs, origin, solid, sdf


def _arrays(positions, indices):
    pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
    idx = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1)
    return pos, idx
