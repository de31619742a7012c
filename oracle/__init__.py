"""ORACLE — TEST INFRASTRUCTURE ONLY.

ctypes front-end for the CPU oracle built from ``oracle/lbm_oracle.cpp``
(``liblbm_oracle.so``) and, where it was built, FlowCraft's own LbmReference
(``oracle/_ref/libref_lbm.so``).  Only ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s CPU-baseline leg may import this package, and only as the
checker: the engine in ``flowcraft_amd`` never routes through it.

The restated functions and the reference lines they follow are listed in the
header of ``lbm_oracle.cpp``; every function there is pinned by tests in
``tests/test_oracle*.py``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(_HERE, "liblbm_oracle.so")
REF_SO = os.path.join(_HERE, "_ref", "libref_lbm.so")
REF_OVERLAY_SO = os.path.join(_HERE, "_ref", "libref_overlay.so")

_c_float_p = ctypes.POINTER(ctypes.c_float)
_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_u8_p = ctypes.POINTER(ctypes.c_uint8)
_c_u16_p = ctypes.POINTER(ctypes.c_uint16)

STORAGE_NATIVE, STORAGE_F16, STORAGE_BF16, STORAGE_F16_SR, STORAGE_BF16_SR = 0, 1, 2, 3, 4
STATS_FIELDS = ("avgDensity", "minDensity", "maxDensity", "avgSpeed", "maxSpeed", "forceX", "forceY",
                "forceZ", "maxStress", "totalMass", "hasNaN")


def build(reference: bool = False) -> None:
    """Compile the oracle (and, when asked and the reference is present, oracle/_ref)."""
    subprocess.run(["make", "-s", "-C", _HERE, "liblbm_oracle.so"], check=True)
    if reference and os.path.isdir("/root/reference/FlowCraft"):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = ctypes.CDLL(ORACLE_SO)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int] * 9
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_set_geometry.argtypes = [ctypes.c_void_p, _c_u8_p, _c_float_p]
        L.orc_set_params.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_float, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_float]
        L.orc_reset.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_reset_from_field.argtypes = [ctypes.c_void_p, _c_float_p]
        L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_readback_field_f32.argtypes = [ctypes.c_void_p, _c_float_p]
        L.orc_readback_field_f64.argtypes = [ctypes.c_void_p, _c_double_p]
        L.orc_stats.argtypes = [ctypes.c_void_p, _c_double_p]
        L.orc_has_diverged.argtypes = [ctypes.c_void_p]
        L.orc_has_diverged.restype = ctypes.c_int
        L.orc_get_pops.argtypes = [ctypes.c_void_p, _c_double_p]
        L.orc_set_pops.argtypes = [ctypes.c_void_p, _c_double_p]
        L.orc_f32_to_f16.argtypes = [ctypes.c_float]
        L.orc_f32_to_f16.restype = ctypes.c_uint16
        L.orc_f32_to_f16_flush.argtypes = [ctypes.c_float]
        L.orc_f32_to_f16_flush.restype = ctypes.c_uint16
        L.orc_f16_to_f32.argtypes = [ctypes.c_uint16]
        L.orc_f16_to_f32.restype = ctypes.c_float
        L.orc_f32_to_bf16.argtypes = [ctypes.c_float]
        L.orc_f32_to_bf16.restype = ctypes.c_uint16
        L.orc_bf16_to_f32.argtypes = [ctypes.c_uint16]
        L.orc_bf16_to_f32.restype = ctypes.c_float
        L.orc_f32_to_f16_array.argtypes = [_c_float_p, _c_u16_p, ctypes.c_size_t]
        L.orc_f32_to_bf16_array.argtypes = [_c_float_p, _c_u16_p, ctypes.c_size_t]
        L.orc_f16_scale.restype = ctypes.c_float
        L.orc_sr_hash.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.orc_sr_hash.restype = ctypes.c_uint32
        L.orc_f32_to_f16_sr.argtypes = [ctypes.c_float, ctypes.c_uint32]
        L.orc_f32_to_f16_sr.restype = ctypes.c_uint16
        L.orc_f32_to_bf16_sr.argtypes = [ctypes.c_float, ctypes.c_uint32]
        L.orc_f32_to_bf16_sr.restype = ctypes.c_uint16
        L.orc_aero_forces.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, _c_float_p, _c_u8_p,
                                      _c_float_p, _c_float_p, ctypes.POINTER(ctypes.c_int)]
        L.orc_vox_frame.argtypes = [_c_float_p, ctypes.c_longlong, ctypes.c_longlong, ctypes.c_int, _c_float_p,
                                    ctypes.POINTER(ctypes.c_int), _c_float_p]
        L.orc_vox_frame.restype = ctypes.c_int
        L.orc_voxelize.argtypes = [_c_float_p, ctypes.c_longlong, ctypes.POINTER(ctypes.c_uint32), ctypes.c_longlong,
                                   ctypes.c_int, _c_float_p, _c_u8_p, _c_float_p]
        L.orc_time_steps.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.orc_time_steps.restype = ctypes.c_double
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref_lib():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError("oracle/_ref/libref_lbm.so is not built (make -C oracle ref)")
        L = ctypes.CDLL(REF_SO)
        L.ref_create.restype = ctypes.c_void_p
        L.ref_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_u8_p, _c_float_p,
                                 ctypes.c_float, ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_float, ctypes.c_int]
        L.ref_destroy.argtypes = [ctypes.c_void_p]
        L.ref_set_params.argtypes = [ctypes.c_void_p, ctypes.c_float, ctypes.c_float, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_float]
        L.ref_reset.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.ref_reset_from_field.argtypes = [ctypes.c_void_p, _c_float_p]
        L.ref_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.ref_readback_field.argtypes = [ctypes.c_void_p, _c_float_p]
        L.ref_stats.argtypes = [ctypes.c_void_p, _c_double_p]
        L.ref_has_diverged.argtypes = [ctypes.c_void_p]
        L.ref_has_diverged.restype = ctypes.c_int
        L.ref_time_steps.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.ref_time_steps.restype = ctypes.c_double
        _ref = L
    return _ref


AERO_FIELDS = ("forceX", "forceY", "forceZ", "minStress", "maxStress", "cop", "liftArrowBase", "hasNaN")


def _aero_call(fn, solid, field, cell_size, origin) -> dict:
    sol = np.ascontiguousarray(solid, dtype=np.uint8)
    nz, ny, nx = sol.shape
    f = np.ascontiguousarray(field, dtype=np.float32).reshape(-1)
    assert f.size == sol.size * 4
    o = np.ascontiguousarray(origin, dtype=np.float32)
    out = np.zeros(11, np.float32)
    nan = ctypes.c_int(0)
    fn(nx, ny, nz, float(cell_size), _ptr(o, _c_float_p), _ptr(sol, _c_u8_p), _ptr(f, _c_float_p),
       _ptr(out, _c_float_p), ctypes.byref(nan))
    return dict(forceX=out[0], forceY=out[1], forceZ=out[2], minStress=out[3], maxStress=out[4],
                cop=out[5:8].copy(), liftArrowBase=out[8:11].copy(), hasNaN=bool(nan.value))


def aero_forces(solid: np.ndarray, field: np.ndarray, cell_size=1.0, origin=(0.0, 0.0, 0.0)) -> dict:
    """Restated FieldOverlay::ComputeAerodynamicForces (FieldOverlay.cpp:469-621) on a
    readback field; solid indexed [z, y, x]."""
    return _aero_call(lib().orc_aero_forces, solid, field, cell_size, origin)


def ref_overlay_available() -> bool:
    return os.path.exists(REF_OVERLAY_SO)


_ref_overlay = None


def ref_aero_forces(solid: np.ndarray, field: np.ndarray, cell_size=1.0, origin=(0.0, 0.0, 0.0)) -> dict:
    """FlowCraft's own FieldOverlay::ComputeAerodynamicForces, compiled into oracle/_ref."""
    global _ref_overlay
    if _ref_overlay is None:
        if not ref_overlay_available():
            raise RuntimeError("oracle/_ref/libref_overlay.so is not built (make -C oracle ref)")
        L = ctypes.CDLL(REF_OVERLAY_SO)
        L.ref_aero_forces.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, _c_float_p, _c_u8_p,
                                      _c_float_p, _c_float_p, ctypes.POINTER(ctypes.c_int)]
        _ref_overlay = L
    return _aero_call(_ref_overlay.ref_aero_forces, solid, field, cell_size, origin)


REF_VOXELIZER_SO = os.path.join(_HERE, "_ref", "libref_voxelizer.so")
VOX_DEFAULT_PADDING = (0.3, 1.0, 0.5)  # Voxelizer::Params::paddingFactor (Voxelizer.h:29)


def _mesh_args(positions, indices):
    p = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
    i = np.ascontiguousarray(indices, dtype=np.uint32).reshape(-1)
    return p, i


def vox_frame(positions, indices, target=640, padding=VOX_DEFAULT_PADDING):
    """(nx, ny, nz), cellSize, origin of Voxelizer::Voxelize's grid for a mesh."""
    L = lib()
    p, i = _mesh_args(positions, indices)
    dims = np.zeros(3, np.int32)
    fr = np.zeros(4, np.float32)
    pad = np.ascontiguousarray(padding, dtype=np.float32)
    L.orc_vox_frame(_ptr(p, _c_float_p), len(p), len(i), int(target), _ptr(pad, _c_float_p),
                    dims.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _ptr(fr, _c_float_p))
    return tuple(int(v) for v in dims), float(fr[0]), tuple(float(v) for v in fr[1:])


def voxelize(positions, indices, target=640, padding=VOX_DEFAULT_PADDING):
    """Restated Voxelizer::Voxelize: (solid uint8 [nz, ny, nx], sdf float32 [nz, ny, nx],
    cellSize, origin)."""
    L = lib()
    p, i = _mesh_args(positions, indices)
    (nx, ny, nz), cs, org = vox_frame(p, i, target, padding)
    solid = np.zeros((nz, ny, nx), np.uint8)
    sdf = np.zeros((nz, ny, nx), np.float32)
    pad = np.ascontiguousarray(padding, dtype=np.float32)
    if solid.size:
        L.orc_voxelize(_ptr(p, _c_float_p), len(p), _ptr(i, ctypes.POINTER(ctypes.c_uint32)), len(i), int(target),
                       _ptr(pad, _c_float_p), _ptr(solid, _c_u8_p), _ptr(sdf, _c_float_p))
    return solid, sdf, cs, org


def ref_voxelizer_available() -> bool:
    return os.path.exists(REF_VOXELIZER_SO)


_ref_vox = None


def ref_voxelize(positions, indices, target=640, padding=VOX_DEFAULT_PADDING):
    """FlowCraft's own Voxelizer::Voxelize (oracle/_ref), same return as voxelize()."""
    global _ref_vox
    if _ref_vox is None:
        if not ref_voxelizer_available():
            raise RuntimeError("oracle/_ref/libref_voxelizer.so is not built (make -C oracle ref)")
        L = ctypes.CDLL(REF_VOXELIZER_SO)
        L.ref_voxelize.restype = ctypes.c_void_p
        L.ref_voxelize.argtypes = [_c_float_p, ctypes.c_longlong, ctypes.POINTER(ctypes.c_uint32), ctypes.c_longlong,
                                   ctypes.c_int, _c_float_p]
        L.ref_vox_info.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), _c_float_p]
        L.ref_vox_copy.argtypes = [ctypes.c_void_p, _c_u8_p, _c_float_p]
        L.ref_vox_free.argtypes = [ctypes.c_void_p]
        _ref_vox = L
    L = _ref_vox
    p, i = _mesh_args(positions, indices)
    pad = np.ascontiguousarray(padding, dtype=np.float32)
    h = L.ref_voxelize(_ptr(p, _c_float_p), len(p), _ptr(i, ctypes.POINTER(ctypes.c_uint32)), len(i), int(target),
                       _ptr(pad, _c_float_p))
    dims = np.zeros(3, np.int32)
    fr = np.zeros(4, np.float32)
    L.ref_vox_info(h, dims.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _ptr(fr, _c_float_p))
    nx, ny, nz = (int(v) for v in dims)
    solid = np.zeros((nz, ny, nx), np.uint8)
    sdf = np.zeros((nz, ny, nx), np.float32)
    L.ref_vox_copy(h, _ptr(solid, _c_u8_p), _ptr(sdf, _c_float_p))
    L.ref_vox_free(h)
    return solid, sdf, float(fr[0]), tuple(float(v) for v in fr[1:])


REF_SMOKE_SO = os.path.join(_HERE, "_ref", "libref_smoke.so")


def ref_smoke_available() -> bool:
    return os.path.exists(REF_SMOKE_SO)


class _SmokeBase:
    """Common front end of the restated (orc_smoke_*) and FlowCraft's own (ref_smoke_*)
    SmokeTrailSystem.  confine: None or (axis 0/1/2 = grid x/y/z, index)."""
    prefix = "orc_smoke_"

    def __init__(self, L, solid, cell_size=1.0, origin=(0.0, 0.0, 0.0)):
        self.L = L
        for name, argt in (("create", [ctypes.c_int] * 3 + [ctypes.c_float, _c_float_p, _c_u8_p]),
                           ("destroy", [ctypes.c_void_p]),
                           ("reset", [ctypes.c_void_p] + [ctypes.c_int] * 5),
                           ("scatter", [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
                           ("shift_plane", [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
                           ("step", [ctypes.c_void_p, _c_float_p, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int]),
                           ("trails", [ctypes.c_void_p, _c_float_p, ctypes.POINTER(ctypes.c_int)])):
            getattr(L, self.prefix + name).argtypes = argt
        L_ = getattr(L, self.prefix + "create")
        L_.restype = ctypes.c_void_p
        self._solid = np.ascontiguousarray(solid, dtype=np.uint8)
        nz, ny, nx = self._solid.shape
        o = np.ascontiguousarray(origin, dtype=np.float32)
        self.h = L_(nx, ny, nz, float(cell_size), _ptr(o, _c_float_p), _ptr(self._solid, _c_u8_p))
        self.P, self.T = 0, 2

    def _f(self, name):
        return getattr(self.L, self.prefix + name)

    def __del__(self):
        if getattr(self, "h", None):
            self._f("destroy")(self.h)
            self.h = None

    @staticmethod
    def _c(confine):
        return (-1, 0) if confine is None else (int(confine[0]), int(confine[1]))

    def reset(self, reverse, count, trail, confine=None):
        self.P, self.T = max(count, 0), max(trail, 2)
        self._f("reset")(self.h, int(reverse), int(count), int(trail), *self._c(confine))

    def scatter(self, confine=None):
        self._f("scatter")(self.h, *self._c(confine))

    def shift_plane(self, confine):
        self._f("shift_plane")(self.h, *self._c(confine))

    def step(self, field, reverse, mmps, dt, suspend=False, confine=None):
        f = np.ascontiguousarray(field, dtype=np.float32)
        self._f("step")(self.h, _ptr(f, _c_float_p), int(reverse), float(mmps), float(dt), int(suspend),
                        *self._c(confine))

    def trails(self):
        """(points float32 [P, T, 3] with unused slots zero, lengths int32 [P])."""
        out = np.zeros((self.P, self.T, 3), np.float32)
        n = np.zeros(self.P, np.int32)
        if self.P:
            self._f("trails")(self.h, _ptr(out, _c_float_p), n.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return out, n


class OracleSmoke(_SmokeBase):
    """The restated SmokeTrailSystem (oracle/smoke_oracle.cpp)."""

    def __init__(self, solid, cell_size=1.0, origin=(0.0, 0.0, 0.0)):
        super().__init__(lib(), solid, cell_size, origin)


class RefSmoke(_SmokeBase):
    """FlowCraft's own SmokeTrailSystem, compiled into oracle/_ref."""
    prefix = "ref_smoke_"

    def __init__(self, solid, cell_size=1.0, origin=(0.0, 0.0, 0.0)):
        if not ref_smoke_available():
            raise RuntimeError("oracle/_ref/libref_smoke.so is not built (make -C oracle ref)")
        L = ctypes.CDLL(REF_SMOKE_SO)
        L.ref_smoke_load_field.argtypes = [ctypes.c_void_p, _c_float_p]
        L.ref_smoke_step_loaded.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                            ctypes.c_int, ctypes.c_int, ctypes.c_int]
        super().__init__(L, solid, cell_size, origin)

    def load_field(self, field):
        """The field into the std::vector the reference steps from (FlowCraft's
        ReadbackField result)."""
        f = np.ascontiguousarray(field, dtype=np.float32)
        self.L.ref_smoke_load_field(self.h, _ptr(f, _c_float_p))

    def step_loaded(self, rev, mmps, dt, suspend=False, confine=None):
        self.L.ref_smoke_step_loaded(self.h, int(rev), float(mmps), float(dt), int(suspend), *self._c(confine))


class OracleLbm:
    """The restated LbmReference.  bits=32 (parity twin) or 64 (FP64 oracle)."""

    def __init__(self, solid: np.ndarray, sdf: np.ndarray | None, q: int = 19, bits: int = 32,
                 storage: int = STORAGE_NATIVE, periodic_y: bool = False, periodic_z: bool = False,
                 nthreads: int = 0):
        L = lib()
        nz, ny, nx = solid.shape
        self.shape = (nz, ny, nx)
        self.q = q
        self.bits = bits
        self.h = L.orc_create(nx, ny, nz, q, bits, storage, int(periodic_y), int(periodic_z), nthreads)
        if not self.h:
            raise ValueError("orc_create rejected the configuration")
        self._solid = np.ascontiguousarray(solid, dtype=np.uint8)
        self._sdf = None if sdf is None else np.ascontiguousarray(sdf, dtype=np.float32)
        L.orc_set_geometry(self.h, _ptr(self._solid, _c_u8_p),
                           None if self._sdf is None else _ptr(self._sdf, _c_float_p))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def set_params(self, tau=0.8, u=0.05, reverse=False, free_slip_y=True, free_slip_z=True, cs=0.12):
        lib().orc_set_params(self.h, tau, u, int(reverse), int(free_slip_y), int(free_slip_z), cs)

    def reset(self, seed_mid=False):
        lib().orc_reset(self.h, int(seed_mid))

    def reset_from_field(self, field: np.ndarray):
        f = np.ascontiguousarray(field, dtype=np.float32)
        lib().orc_reset_from_field(self.h, _ptr(f, _c_float_p))

    def step(self, n=1):
        lib().orc_step(self.h, n)

    def field(self) -> np.ndarray:
        n = int(np.prod(self.shape))
        if self.bits == 64:
            out = np.empty(n * 4, np.float64)
            lib().orc_readback_field_f64(self.h, _ptr(out, _c_double_p))
        else:
            out = np.empty(n * 4, np.float32)
            lib().orc_readback_field_f32(self.h, _ptr(out, _c_float_p))
        return out

    def stats(self) -> dict:
        out = np.zeros(11, np.float64)
        lib().orc_stats(self.h, _ptr(out, _c_double_p))
        return dict(zip(STATS_FIELDS, out.tolist()))

    def diverged(self) -> bool:
        return bool(lib().orc_has_diverged(self.h))

    def pops(self) -> np.ndarray:
        out = np.empty(self.q * int(np.prod(self.shape)), np.float64)
        lib().orc_get_pops(self.h, _ptr(out, _c_double_p))
        return out.reshape(self.q, *self.shape)

    def set_pops(self, pops: np.ndarray):
        p = np.ascontiguousarray(pops, dtype=np.float64)
        lib().orc_set_pops(self.h, _ptr(p, _c_double_p))

    def time_steps(self, warmup: int, steps: int) -> float:
        return lib().orc_time_steps(self.h, warmup, steps)


class RefLbm:
    """FlowCraft's own LbmReference, compiled from its sources into oracle/_ref by `make ref`
    (built where the reference sources are present; the built library travels with the tree)."""

    def __init__(self, solid: np.ndarray, sdf: np.ndarray | None, q: int = 19, tau=0.8, u=0.05,
                 reverse=False, free_slip_y=True, free_slip_z=True, cs=0.12, seed_mid=False):
        L = ref_lib()
        nz, ny, nx = solid.shape
        self.shape = (nz, ny, nx)
        self._solid = np.ascontiguousarray(solid, dtype=np.uint8)
        self._sdf = None if sdf is None else np.ascontiguousarray(sdf, dtype=np.float32)
        self.h = L.ref_create(nx, ny, nz, q, _ptr(self._solid, _c_u8_p),
                              None if self._sdf is None else _ptr(self._sdf, _c_float_p),
                              tau, u, int(reverse), int(free_slip_y), int(free_slip_z), cs, int(seed_mid))

    def __del__(self):
        if getattr(self, "h", None):
            ref_lib().ref_destroy(self.h)
            self.h = None

    def set_params(self, tau=0.8, u=0.05, reverse=False, free_slip_y=True, free_slip_z=True, cs=0.12):
        ref_lib().ref_set_params(self.h, tau, u, int(reverse), int(free_slip_y), int(free_slip_z), cs)

    def reset(self, seed_mid=False):
        ref_lib().ref_reset(self.h, int(seed_mid))

    def reset_from_field(self, field):
        f = np.ascontiguousarray(field, dtype=np.float32)
        ref_lib().ref_reset_from_field(self.h, _ptr(f, _c_float_p))

    def step(self, n=1):
        ref_lib().ref_step(self.h, n)

    def field(self) -> np.ndarray:
        out = np.empty(int(np.prod(self.shape)) * 4, np.float32)
        ref_lib().ref_readback_field(self.h, _ptr(out, _c_float_p))
        return out

    def stats(self) -> dict:
        out = np.zeros(11, np.float64)
        ref_lib().ref_stats(self.h, _ptr(out, _c_double_p))
        return dict(zip(STATS_FIELDS, out.tolist()))

    def diverged(self) -> bool:
        return bool(ref_lib().ref_has_diverged(self.h))

    def time_steps(self, warmup: int, steps: int) -> float:
        return ref_lib().ref_time_steps(self.h, warmup, steps)
