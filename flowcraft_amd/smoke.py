"""Python mirror of FlowCraft's SmokeTrailSystem (FlowCraft/SmokeTrail.h:10-184) with
its particles on the device (libfc_lbm.so, csrc/smoke.hip): trails advected through the
solver's field where it lives.  Same methods and argument meaning as the reference
(Reset, Step, ScatterRandom, ShiftPlane); ``confine`` is None or a
``PlaneConfinement(axis, index)`` with axis 0/1/2 = SliceAxis X/Y/Z (a grid plane).
``Step`` takes the solver (its current state) or a device field tensor in place of the
reference's host field vector.  ``Trails()`` returns the trails for drawing.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib as L
from .lbm import LbmSolver, VoxelGrid


@dataclass
class PlaneConfinement:
    axis: int   # 0, 1, 2 = SliceAxis::X, Y, Z (grid axes)
    index: int


class SmokeTrailSystem:
    def __init__(self, grid: VoxelGrid):
        self._lib = L.load()
        self._h = ctypes.c_void_p()
        solid = np.ascontiguousarray(grid.solid, dtype=np.uint8).reshape(-1)
        o = (ctypes.c_float * 3)(*[float(v) for v in grid.origin])
        self._dims = (grid.nx, grid.ny, grid.nz)
        self._check(self._lib.fc_smoke_create(grid.nx, grid.ny, grid.nz, float(grid.cellSize), o, solid.ctypes.data,
                                              ctypes.byref(self._h)))

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            self._lib.fc_smoke_destroy(self._h)
            self._h = ctypes.c_void_p()

    def _check(self, rc):
        if rc != L.FC_LBM_OK:
            raise L.FcLbmError(rc, self._lib.fc_smoke_last_error(self._h).decode())

    @staticmethod
    def _c(confine: Optional[PlaneConfinement]):
        return (-1, 0) if confine is None else (int(confine.axis), int(confine.index))

    def Reset(self, grid: VoxelGrid, reverseFlow: bool, particleCount: int, trailLength: int,
              confine: Optional[PlaneConfinement] = None) -> None:
        """SmokeTrailSystem::Reset (FlowCraft/SmokeTrail.cpp:257-275)."""
        self._check(self._lib.fc_smoke_reset(self._h, int(reverseFlow), int(particleCount), int(trailLength),
                                             *self._c(confine)))

    def ScatterRandom(self, grid: VoxelGrid, confine: Optional[PlaneConfinement] = None) -> None:
        """SmokeTrailSystem::ScatterRandom (FlowCraft/SmokeTrail.cpp:202-216)."""
        self._check(self._lib.fc_smoke_scatter(self._h, *self._c(confine)))

    def ShiftPlane(self, grid: VoxelGrid, confine: PlaneConfinement) -> None:
        """SmokeTrailSystem::ShiftPlane (FlowCraft/SmokeTrail.cpp:218-255)."""
        self._check(self._lib.fc_smoke_shift_plane(self._h, int(confine.axis), int(confine.index)))

    def Step(self, grid: VoxelGrid, source, reverseFlow: bool, mmPerSecPerLatticeUnit: float, dtSeconds: float,
             suspendRespawn: bool = False, confine: Optional[PlaneConfinement] = None) -> None:
        """SmokeTrailSystem::Step (FlowCraft/SmokeTrail.cpp:277-345).  `source`: the
        LbmSolver (its current state) or a float32 device tensor of 4 floats per cell."""
        args = (int(reverseFlow), float(mmPerSecPerLatticeUnit), float(dtSeconds), int(suspendRespawn),
                *self._c(confine))
        if isinstance(source, LbmSolver):
            self._check(self._lib.fc_smoke_step(self._h, source._h, *args))
        else:
            import torch
            n = self._dims[0] * self._dims[1] * self._dims[2] * 4
            if not (isinstance(source, torch.Tensor) and source.is_cuda and source.dtype == torch.float32
                    and source.is_contiguous() and source.numel() == n):
                raise ValueError(f"source must be an LbmSolver or a contiguous float32 device tensor of {n} values")
            # the library reads the field on its own stream: the work that wrote it must be done
            torch.cuda.current_stream(source.device).synchronize()
            self._check(self._lib.fc_smoke_step_field(self._h, ctypes.c_void_p(source.data_ptr()), *args))

    def Trails(self):
        """(points float32 [particles, trailLength, 3] oldest first, unused slots zero;
        lengths int32 [particles])."""
        p, t = ctypes.c_int(0), ctypes.c_int(0)
        self._check(self._lib.fc_smoke_info(self._h, ctypes.byref(p), ctypes.byref(t)))
        out = np.zeros((p.value, t.value, 3), np.float32)
        n = np.zeros(p.value, np.int32)
        if p.value:
            self._check(self._lib.fc_smoke_trails(self._h, out.ctypes.data, n.ctypes.data))
        return out, n
