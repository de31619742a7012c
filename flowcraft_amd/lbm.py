"""Python mirror of FlowCraft's ``LbmSolver`` (FlowCraft/LbmSolver.h:59-302).

Same method names, argument meaning and error behaviour as the reference
class, over the C ABI of ``libfc_lbm.so``:

* ``Init`` returns ``False`` and sets ``LastError()`` on failure
  (FlowCraft/LbmSolver.cpp:205-366); a missing shared library raises.
* Every other method silently does nothing on an invalid solver, and
  ``ResetFromField`` ignores a field of the wrong size
  (FlowCraft/LbmSolver.cpp:430), as the reference does.
* ``Step``/``StepRange`` only enqueue work on the solver's HIP stream (the
  reference records into the caller's command buffer); readbacks block.

The SDL handles of the reference signatures (``SDL_GPUDevice*``,
``SDL_GPUCommandBuffer*``) are accepted as ignored optional arguments.
Extensions beyond the reference: storage precision and population layout at
``Init``, periodic y/z, FP64 force sums (``Forces``), exact state
dump/restore (``Populations``/``SetPopulations``), on-device field export, and
FieldOverlay's aerodynamic-force summary computed on the device
(``ComputeAerodynamicForces``).
"""
from __future__ import annotations

import ctypes
import enum
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib as L


class LatticeModel(enum.IntEnum):
    """FlowCraft/LatticeModel.h:12 (same numbering)."""
    D3Q15 = L.D3Q15
    D3Q19 = L.D3Q19


class Storage(enum.IntEnum):
    """Population storage (fc_lbm.h FC_LBM_STORE_*): the _SR formats round the step's
    stores stochastically instead of to nearest."""
    F32 = L.STORE_F32
    F16 = L.STORE_F16
    BF16 = L.STORE_BF16
    F64 = L.STORE_F64
    F16_SR = L.STORE_F16_SR
    BF16_SR = L.STORE_BF16_SR


class Layout(enum.IntEnum):
    PINGPONG = L.LAYOUT_PINGPONG
    INPLACE = L.LAYOUT_INPLACE


@dataclass
class Params:
    """LbmSolver::Params (FlowCraft/LbmSolver.h:61-100), same defaults."""
    tau: float = 0.8
    inletVelocity: float = 0.05
    reverseFlow: bool = False
    freeSlipY: bool = True
    freeSlipZ: bool = True
    smagorinskyCs: float = 0.12

    def _c(self) -> L.FcLbmParams:
        return L.FcLbmParams(self.tau, self.inletVelocity, int(self.reverseFlow), int(self.freeSlipY),
                             int(self.freeSlipZ), self.smagorinskyCs)


@dataclass
class PacingState:
    """StepPaced's state carried between calls (FlowCraft keeps these as statics of
    StepSolverPaced, main.cpp:2376-2379, 2448-2449)."""
    chunkZDepth: int = 32
    probePhase: int = 0
    lastProbe: float = float("-inf")
    diverged: bool = False


@dataclass
class AerodynamicForces:
    """FieldOverlay::AerodynamicForces (FlowCraft/FieldOverlay.h:186-195); vectors in the
    VoxelGrid's world frame (grid-y along world X, grid-x along world Y)."""
    forceX: float = 0.0
    forceY: float = 0.0
    forceZ: float = 0.0
    minStress: float = 0.0
    maxStress: float = 0.0
    centerOfPressure: tuple = (0.0, 0.0, 0.0)
    liftArrowBase: tuple = (0.0, 0.0, 0.0)
    hasNaN: bool = False


@dataclass
class Stats:
    """LbmSolver::Stats (FlowCraft/LbmSolver.h:172-184) + totalMass (LbmReference.h:117)."""
    avgDensity: float = 0.0
    minDensity: float = 0.0
    maxDensity: float = 0.0
    avgSpeed: float = 0.0
    maxSpeed: float = 0.0
    forceX: float = 0.0
    forceY: float = 0.0
    forceZ: float = 0.0
    maxStress: float = 0.0
    totalMass: float = 0.0
    fluidCells: int = 0
    hasNaN: bool = False
    valid: bool = False


@dataclass
class VoxelGrid:
    """FlowCraft/VoxelGrid.h:13-61.  ``solid`` and ``signedDistance`` are indexed
    [z, y, x] (flat order (z*ny + y)*nx + x, VoxelGrid::Index); ``signedDistance``
    may be None (the reference's empty vector: plain bounce-back everywhere)."""
    nx: int
    ny: int
    nz: int
    solid: np.ndarray
    signedDistance: Optional[np.ndarray] = None
    cellSize: float = 1.0
    origin: tuple = (0.0, 0.0, 0.0)

    def CellCount(self) -> int:
        return self.nx * self.ny * self.nz

    def Index(self, x: int, y: int, z: int) -> int:
        return (z * self.ny + y) * self.nx + x


def _check_tensor(t, name, dtypes, numel=None, min_numel=None):
    """The library reads raw device pointers: anything but a contiguous GPU tensor of the
    expected element type and size would be read as the wrong bytes or overrun."""
    if not (hasattr(t, "data_ptr") and hasattr(t, "dtype") and hasattr(t, "is_cuda")):
        raise ValueError(f"{name} must be a torch tensor")
    if t.dtype not in dtypes:
        raise ValueError(f"{name} must have dtype {' or '.join(str(d) for d in dtypes)}, not {t.dtype}")
    if numel is not None and t.numel() != numel:
        raise ValueError(f"{name} must hold {numel} values, not {t.numel()}")
    if min_numel is not None and t.numel() < min_numel:
        raise ValueError(f"{name} must hold at least {min_numel} values, not {t.numel()}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


def _check_on_gpu(**tensors) -> None:
    for name, t in tensors.items():
        if t is not None and not t.is_cuda:
            raise ValueError(f"{name.replace('_', ' ')} must be a GPU tensor")


def check_geometry_tensors(cells: int, solid, sdf=None) -> None:
    """Arguments of LbmSolver.SetGeometryDevice: a uint8/bool mask and an optional float32
    signed distance of `cells` values each, contiguous, on the GPU."""
    import torch
    _check_tensor(solid, "solid mask", (torch.uint8, torch.bool), numel=cells)
    if sdf is not None:
        _check_tensor(sdf, "signed distance", (torch.float32,), numel=cells)
    _check_on_gpu(solid_mask=solid, signed_distance=sdf)


def check_field_at_tensors(cells, out) -> None:
    """Arguments of LbmSolver.FieldAt: int32 cell indices and a float32 output of at least
    4 values per cell, contiguous, on the GPU."""
    import torch
    _check_tensor(cells, "cell list", (torch.int32,))
    _check_tensor(out, "field output", (torch.float32,), min_numel=4 * cells.numel())
    _check_on_gpu(cell_list=cells, field_output=out)


class LbmSolver:
    """Drop-in for FlowCraft's LbmSolver on one MI355X (see module docstring)."""

    def __init__(self):
        self._h = None
        self._lib = None
        self._error = ""
        self._params = Params()
        self._solid = np.zeros(0, np.uint8)
        self._dims = (0, 0, 0)
        self._model = LatticeModel.D3Q19
        self._storage = Storage.F32
        self._layout = Layout.PINGPONG

    # ------------------------------------------------------------ lifecycle
    def Init(self, device=None, grid: VoxelGrid = None, model: LatticeModel = LatticeModel.D3Q19, *,
             storage: Storage = Storage.F32, layout: Layout = Layout.PINGPONG, periodicY: bool = False,
             periodicZ: bool = False, stream=None) -> bool:
        """LbmSolver::Init (FlowCraft/LbmSolver.cpp:205-366).  ``device`` is ignored."""
        self.Shutdown()
        if isinstance(device, VoxelGrid) and grid is None:
            grid, device = device, None
        self._lib = L.load()
        if grid is None or grid.CellCount() == 0:
            self._error = "empty voxel grid"
            return False
        solid = np.array(grid.solid, dtype=np.uint8).reshape(-1)  # a copy: SolidMask() is the solver's own
        if solid.size != grid.CellCount():
            self._error = "voxel grid solid mask has the wrong size"
            return False
        sdf = None
        if grid.signedDistance is not None and np.asarray(grid.signedDistance).size == grid.CellCount():
            sdf = np.ascontiguousarray(np.asarray(grid.signedDistance, dtype=np.float32).reshape(-1))
        desc = L.FcLbmDesc(grid.nx, grid.ny, grid.nz, int(model), int(storage), int(layout), int(periodicY),
                           int(periodicZ), stream)
        h = ctypes.c_void_p()
        rc = self._lib.fc_lbm_create(ctypes.byref(desc), ctypes.byref(h))
        if rc != L.FC_LBM_OK:
            self._error = (self._lib.fc_lbm_last_error(h) or b"").decode()
            if h:
                self._lib.fc_lbm_destroy(h)
            return False
        rc = self._lib.fc_lbm_set_geometry(h, solid.ctypes.data,
                                           None if sdf is None else sdf.ctypes.data)
        if rc != L.FC_LBM_OK:
            self._error = self._lib.fc_lbm_last_error(h).decode()
            self._lib.fc_lbm_destroy(h)
            return False
        self._h = h
        self._error = ""
        self._solid = solid
        self._dims = (grid.nx, grid.ny, grid.nz)
        self._model = LatticeModel(model)
        self._storage = Storage(storage)
        self._layout = Layout(layout)
        return True

    def Shutdown(self) -> None:
        if self._h is not None:
            self._lib.fc_lbm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.Shutdown()
        except Exception:
            pass

    def LastError(self) -> str:
        return self._error

    def IsValid(self) -> bool:
        return self._h is not None

    def Model(self) -> LatticeModel:
        return self._model

    def Nx(self) -> int:
        return self._dims[0]

    def Ny(self) -> int:
        return self._dims[1]

    def Nz(self) -> int:
        return self._dims[2]

    def CellCount(self) -> int:
        return self._dims[0] * self._dims[1] * self._dims[2]

    def SolidMask(self) -> np.ndarray:
        """LbmSolver::SolidMask (FlowCraft/LbmSolver.h:216-218).  After SetGeometryDevice the
        host copy is made on first use from the library's own device copy of the mask."""
        if self._solid is None and self.IsValid():
            out = np.empty(self.CellCount(), np.uint8)
            self._check(self._lib.fc_lbm_solid_mask(self._h, out.ctypes.data))
            self._solid = out
        return self._solid

    # ---------------------------------------------------------------- state
    def _check(self, rc: int) -> None:
        if rc != L.FC_LBM_OK:
            msg = self._lib.fc_lbm_last_error(self._h).decode()
            self._error = msg
            raise L.FcLbmError(rc, msg)

    def Reset(self, params: Params = None, seedAtMidSpeed: bool = False) -> None:
        """LbmSolver::Reset (FlowCraft/LbmSolver.cpp:389-427): sets params, then seeds."""
        self._params = Params(**vars(params)) if params is not None else Params()
        if not self.IsValid():
            return
        p = self._params._c()
        self._check(self._lib.fc_lbm_set_params(self._h, ctypes.byref(p)))
        self._check(self._lib.fc_lbm_reset_uniform(self._h, int(seedAtMidSpeed)))

    def ResetFromField(self, field4) -> None:
        """LbmSolver::ResetFromField (FlowCraft/LbmSolver.cpp:429-472); params untouched."""
        if not self.IsValid():
            return
        if hasattr(field4, "data_ptr"):  # a device tensor: hand it over in HBM
            if field4.numel() != self.CellCount() * 4:
                return
            t = field4.contiguous()
            self._check(self._lib.fc_lbm_reset_from_field_device(self._h, t.data_ptr()))
            return
        f = np.ascontiguousarray(np.asarray(field4, dtype=np.float32).reshape(-1))
        if f.size != self.CellCount() * 4:
            return
        self._check(self._lib.fc_lbm_reset_from_field(self._h, f.ctypes.data))

    def SetParams(self, params: Params) -> None:
        """LbmSolver::SetParams (FlowCraft/LbmSolver.h:135): state untouched."""
        self._params = Params(**vars(params))
        if self.IsValid():
            p = self._params._c()
            self._check(self._lib.fc_lbm_set_params(self._h, ctypes.byref(p)))

    def GetParams(self) -> Params:
        return Params(**vars(self._params))

    # ------------------------------------------------------------- stepping
    def Step(self, cmd=None, numSubsteps: int = 1) -> None:
        """LbmSolver::Step (FlowCraft/LbmSolver.cpp:556-562); ``cmd`` ignored.
        ``Step(n)`` is accepted as shorthand for ``Step(None, n)``."""
        if isinstance(cmd, int) and numSubsteps == 1:
            cmd, numSubsteps = None, cmd
        if not self.IsValid():
            return
        self._check(self._lib.fc_lbm_step(self._h, int(numSubsteps)))

    def StepRange(self, cmd=None, zStart: int = 0, zCount: int = 0) -> None:
        """LbmSolver::StepRange (FlowCraft/LbmSolver.cpp:543-554)."""
        if not self.IsValid():
            return
        self._check(self._lib.fc_lbm_step_range(self._h, int(zStart), int(zCount)))

    def AdvanceSubstep(self) -> None:
        """LbmSolver::AdvanceSubstep (FlowCraft/LbmSolver.h:170)."""
        if self.IsValid():
            self._check(self._lib.fc_lbm_advance(self._h))

    def Sync(self) -> None:
        if self.IsValid():
            self._check(self._lib.fc_lbm_sync(self._h))

    def StepCount(self) -> int:
        return int(self._lib.fc_lbm_step_count(self._h)) if self.IsValid() else 0

    # ------------------------------------------------------------ readbacks
    def ReadbackField(self) -> np.ndarray:
        """LbmSolver::ReadbackField (FlowCraft/LbmSolver.cpp:722-737): 4 floats/cell."""
        if not self.IsValid():
            return np.zeros(0, np.float32)
        out = np.empty(self.CellCount() * 4, np.float32)
        self._check(self._lib.fc_lbm_readback_field(self._h, out.ctypes.data))
        return out

    def ReadbackFieldRange(self, zStart: int, zCount: int, outField: np.ndarray) -> None:
        """LbmSolver::ReadbackFieldRange (FlowCraft/LbmSolver.cpp:584-651): outField
        must already hold CellCount()*4 float32 values; layers land at global offsets."""
        if not self.IsValid() or outField.dtype != np.float32 or outField.size != self.CellCount() * 4:
            return
        if not outField.flags["C_CONTIGUOUS"]:
            return
        self._check(self._lib.fc_lbm_readback_field_range(self._h, int(zStart), int(zCount), outField.ctypes.data))

    def FieldDevice(self, out_tensor) -> None:
        """Write the 4-floats/cell field into a device tensor (no host round trip)."""
        if self.IsValid():
            self._check(self._lib.fc_lbm_field_device(self._h, out_tensor.data_ptr()))

    def SetGeometryDevice(self, solid_tensor, sdf_tensor=None) -> None:
        """Extension: replace the geometry from device tensors (uint8/bool mask and float32
        SDF of CellCount() values, contiguous, on the GPU; e.g. Voxelizer.VoxelizeDevice's);
        a reset must follow.  Raises ValueError for any other tensor.  The library copies
        both, so the tensors may be reused afterwards."""
        if not self.IsValid():
            return
        check_geometry_tensors(self.CellCount(), solid_tensor, sdf_tensor)
        import torch
        torch.cuda.current_stream(solid_tensor.device).synchronize()
        self._check(self._lib.fc_lbm_set_geometry_device(self._h, solid_tensor.data_ptr(),
                                                         None if sdf_tensor is None else sdf_tensor.data_ptr()))
        # the host copy of the mask is made when SolidMask() is first asked for
        self._solid = None

    def FieldAt(self, cells_tensor, out_tensor) -> None:
        """The field at listed cells (contiguous int32 device tensor of global indices) into a
        contiguous float32 device tensor of at least 4 values per cell: fluid cells as
        FieldDevice, solid cells zero.  Raises ValueError for any other tensor.  Asynchronous
        on the solver's stream (Sync before reading on another)."""
        if self.IsValid():
            check_field_at_tensors(cells_tensor, out_tensor)
            self._check(self._lib.fc_lbm_field_at(self._h, cells_tensor.data_ptr(), int(cells_tensor.numel()),
                                                  out_tensor.data_ptr()))

    def ReadbackStats(self) -> Stats:
        """LbmSolver::ReadbackStats (FlowCraft/LbmSolver.cpp:739-788), reduced on device."""
        if not self.IsValid():
            return Stats()
        s = L.FcLbmStats()
        self._check(self._lib.fc_lbm_readback_stats(self._h, ctypes.byref(s)))
        return Stats(s.avg_density, s.min_density, s.max_density, s.avg_speed, s.max_speed, s.force_x, s.force_y,
                     s.force_z, s.max_stress, s.total_mass, s.fluid_cells, bool(s.has_nan), True)

    def Forces(self) -> np.ndarray:
        """(Fx, Fy, Fz, maxStress): FP64 momentum-exchange sums over all solid cells."""
        out = np.zeros(4, np.float64)
        if self.IsValid():
            self._check(self._lib.fc_lbm_forces(self._h, out.ctypes.data))
        return out

    def ForcesDevice(self, out_tensor) -> None:
        """Same sums into a float64 device tensor of 4, without synchronising."""
        if self.IsValid():
            self._check(self._lib.fc_lbm_forces_device(self._h, out_tensor.data_ptr()))

    def StepPaced(self, substeps: int, state: Optional["PacingState"] = None, cancel=None,
                  targetChunkMs: float = 3.0, probeIntervalMs: float = 400.0, progress=None) -> int:
        """FlowCraft's StepSolverPaced (FlowCraft/main.cpp:2372-2486): each substep as
        z-chunks of adaptive depth (halved above targetChunkMs, doubled under a quarter of
        it, within [4, 256] layers), waiting for each chunk; AdvanceSubstep; a slice health
        probe (nz/2, nz/3, 2nz/3 in turn) at most every probeIntervalMs.  Returns the
        substeps done: fewer when cancel() turns true between chunks, or i+1 after the
        substep whose probe found a divergence (state.diverged).  progress(f) receives
        (i + z/nz) / substeps.  In the in-place layout the chunks of a substep overwrite
        the state, so a started substep is finished and cancel() is only asked at
        substep boundaries."""
        import time
        if not self.IsValid():
            return 0
        st = state if state is not None else PacingState()
        nz = self.Nz()
        undoable = self._layout == Layout.PINGPONG
        if progress:
            progress(0.0)
        for i in range(substeps):
            z0 = 0
            while z0 < nz:
                if (z0 == 0 or undoable) and cancel and cancel():
                    return i
                zc = min(st.chunkZDepth, nz - z0)
                t0 = time.perf_counter()
                self._check(self._lib.fc_lbm_step_range(self._h, z0, zc))
                self._check(self._lib.fc_lbm_sync(self._h))
                ms = (time.perf_counter() - t0) * 1e3
                if ms > targetChunkMs and st.chunkZDepth > 4:
                    st.chunkZDepth = max(4, st.chunkZDepth // 2)
                elif ms < targetChunkMs * 0.25 and st.chunkZDepth < 256:
                    st.chunkZDepth = min(256, st.chunkZDepth * 2)
                z0 += zc
                if progress:
                    progress((i + z0 / nz) / substeps)
            if undoable and cancel and cancel():
                return i
            self._check(self._lib.fc_lbm_advance(self._h))
            now = time.perf_counter()
            if (now - st.lastProbe) * 1e3 >= probeIntervalMs and nz > 0:
                st.lastProbe = now
                pz = nz // 2 if st.probePhase == 0 else (nz // 3 if st.probePhase == 1 else (2 * nz) // 3)
                st.probePhase = (st.probePhase + 1) % 3
                if self.ProbeSliceHealth(pz):
                    st.diverged = True
                    return i + 1
        return substeps

    def SnapshotBegin(self) -> None:
        """Start an asynchronous field snapshot of the current state (fc_lbm_snapshot_begin):
        later steps run while it is copied to pinned host memory."""
        if self.IsValid():
            self._check(self._lib.fc_lbm_snapshot_begin(self._h))

    def SnapshotReady(self) -> bool:
        r = ctypes.c_int(0)
        if self.IsValid():
            self._check(self._lib.fc_lbm_snapshot_poll(self._h, ctypes.byref(r)))
        return bool(r.value)

    def SnapshotWait(self, copy: bool = True) -> np.ndarray:
        """The snapshot's field (4 floats per cell).  copy=False returns a view of the
        pinned buffer, valid until the next SnapshotBegin or Shutdown."""
        if not self.IsValid():
            return np.zeros(0, np.float32)
        p = ctypes.POINTER(ctypes.c_float)()
        self._check(self._lib.fc_lbm_snapshot_wait(self._h, ctypes.byref(p)))
        view = np.ctypeslib.as_array(p, shape=(self.CellCount() * 4,))
        return view.copy() if copy else view

    def ComputeAerodynamicForces(self, cellSize: float = 1.0, origin=(0.0, 0.0, 0.0)) -> AerodynamicForces:
        """FieldOverlay::ComputeAerodynamicForces (FlowCraft/FieldOverlay.cpp:469-621) of the
        current state, on the device (the reference takes the ReadbackField() result).
        cellSize/origin: the VoxelGrid's world frame."""
        if not self.IsValid():
            return AerodynamicForces()
        a = L.FcLbmAero()
        o = (ctypes.c_float * 3)(*[float(v) for v in origin])
        self._check(self._lib.fc_lbm_aero_forces(self._h, float(cellSize), o, ctypes.byref(a)))
        return AerodynamicForces(a.force_x, a.force_y, a.force_z, a.min_stress, a.max_stress,
                                 tuple(a.center_of_pressure), tuple(a.lift_arrow_base), bool(a.has_nan))

    def ProbeSliceHealth(self, zLayer: int) -> bool:
        """LbmSolver::ProbeSliceHealth (FlowCraft/LbmSolver.cpp:653-720): True = diverged."""
        if not self.IsValid() or self.Nz() <= 0:
            return False
        d = ctypes.c_int(0)
        self._check(self._lib.fc_lbm_probe_health(self._h, int(zLayer), ctypes.byref(d)))
        return bool(d.value)

    def HasDiverged(self) -> bool:
        """LbmReference::HasDiverged (FlowCraft/LbmReference.h:129)."""
        if not self.IsValid():
            return False
        d = ctypes.c_int(0)
        self._check(self._lib.fc_lbm_has_diverged(self._h, ctypes.byref(d)))
        return bool(d.value)

    # ----------------------------------------------------- state dump/restore
    def _pop_dtype(self):
        return np.float64 if self._storage == Storage.F64 else np.float32

    def Populations(self) -> np.ndarray:
        """Current post-collision populations, shape (Q, nz, ny, nx)."""
        q = 15 if self._model == LatticeModel.D3Q15 else 19
        out = np.empty((q, self.Nz(), self.Ny(), self.Nx()), self._pop_dtype())
        self._check(self._lib.fc_lbm_get_populations(self._h, out.ctypes.data))
        return out

    def SetPopulations(self, pops: np.ndarray) -> None:
        p = np.ascontiguousarray(pops, dtype=self._pop_dtype())
        self._check(self._lib.fc_lbm_set_populations(self._h, p.ctypes.data))

    def BoundaryLinks(self):
        """(cell, direction, kind, q) of every fluid->solid link of the Bouzidi table."""
        cnt = ctypes.c_longlong(0)
        self._check(self._lib.fc_lbm_boundary_links(self._h, 0, None, None, None, None, ctypes.byref(cnt)))
        n = cnt.value
        cell = np.empty(n, np.int64)
        d = np.empty(n, np.int32)
        kind = np.empty(n, np.int32)
        q = np.empty(n, np.float32)
        self._check(self._lib.fc_lbm_boundary_links(self._h, n, cell.ctypes.data, d.ctypes.data, kind.ctypes.data,
                                                    q.ctypes.data, ctypes.byref(cnt)))
        return cell, d, kind, q

    def SurfaceCells(self) -> int:
        return int(self._lib.fc_lbm_surface_cells(self._h)) if self.IsValid() else 0

    def PopulationBytes(self) -> int:
        return int(self._lib.fc_lbm_population_bytes(self._h)) if self.IsValid() else 0

    def Stream(self) -> int:
        return int(self._lib.fc_lbm_stream(self._h) or 0) if self.IsValid() else 0
