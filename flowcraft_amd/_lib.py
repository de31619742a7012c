"""ctypes binding of libfc_lbm.so (the C ABI in include/flowcraft/fc_lbm.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``python -m flowcraft_amd.build``) into ``flowcraft_amd/lib/``.  There is no
fallback: if the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
LIB_PATH = os.path.join(LIB_DIR, "libfc_lbm.so")

FC_LBM_OK, FC_LBM_EINVAL, FC_LBM_ENOMEM, FC_LBM_EDEVICE, FC_LBM_ESTATE = 0, 1, 2, 3, 4
D3Q15, D3Q19 = 0, 1
STORE_F32, STORE_F16, STORE_BF16, STORE_F64, STORE_F16_SR, STORE_BF16_SR = 0, 1, 2, 3, 4, 5
LAYOUT_PINGPONG, LAYOUT_INPLACE = 0, 1

# Every symbol include/flowcraft/fc_lbm.h declares (tests check the export table).
EXPORTS = (
    "fc_lbm_create", "fc_lbm_destroy", "fc_lbm_last_error", "fc_lbm_params_default", "fc_lbm_set_geometry", "fc_lbm_set_geometry_device", "fc_lbm_solid_mask",
    "fc_lbm_set_params", "fc_lbm_get_params", "fc_lbm_reset_uniform", "fc_lbm_reset_from_field",
    "fc_lbm_reset_from_field_device", "fc_lbm_step", "fc_lbm_step_range", "fc_lbm_advance",
    "fc_lbm_readback_field_range", "fc_lbm_readback_field", "fc_lbm_readback_field_range_noalloc", "fc_lbm_field_device", "fc_lbm_field_at", "fc_lbm_readback_stats",
    "fc_lbm_forces", "fc_lbm_forces_device", "fc_lbm_probe_health", "fc_lbm_has_diverged",
    "fc_lbm_get_populations", "fc_lbm_set_populations", "fc_lbm_sync", "fc_lbm_info", "fc_lbm_step_count",
    "fc_lbm_population_bytes", "fc_lbm_stream", "fc_lbm_boundary_links", "fc_lbm_surface_cells",
    "fc_lbm_aero_forces", "fc_vox_params_default", "fc_voxelize_frame", "fc_voxelize", "fc_voxelize_device", "fc_voxelize_last_error",
    "fc_smoke_create", "fc_smoke_destroy", "fc_smoke_last_error", "fc_smoke_reset", "fc_smoke_scatter",
    "fc_smoke_shift_plane", "fc_smoke_step", "fc_smoke_step_field", "fc_smoke_trails", "fc_smoke_info",
    "fc_lbm_snapshot_begin", "fc_lbm_snapshot_poll", "fc_lbm_snapshot_wait",
)


class FcLbmDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("model", ctypes.c_int),
                ("storage", ctypes.c_int), ("layout", ctypes.c_int), ("periodic_y", ctypes.c_int),
                ("periodic_z", ctypes.c_int), ("stream", ctypes.c_void_p)]


class FcLbmParams(ctypes.Structure):
    _fields_ = [("tau", ctypes.c_float), ("inlet_velocity", ctypes.c_float), ("reverse_flow", ctypes.c_int),
                ("free_slip_y", ctypes.c_int), ("free_slip_z", ctypes.c_int), ("smagorinsky_cs", ctypes.c_float)]


class FcLbmStats(ctypes.Structure):
    _fields_ = [("avg_density", ctypes.c_double), ("min_density", ctypes.c_double),
                ("max_density", ctypes.c_double), ("avg_speed", ctypes.c_double), ("max_speed", ctypes.c_double),
                ("force_x", ctypes.c_double), ("force_y", ctypes.c_double), ("force_z", ctypes.c_double),
                ("max_stress", ctypes.c_double), ("total_mass", ctypes.c_double),
                ("fluid_cells", ctypes.c_longlong), ("has_nan", ctypes.c_int)]


class FcLbmAero(ctypes.Structure):
    _fields_ = [("force_x", ctypes.c_float), ("force_y", ctypes.c_float), ("force_z", ctypes.c_float),
                ("min_stress", ctypes.c_float), ("max_stress", ctypes.c_float),
                ("center_of_pressure", ctypes.c_float * 3), ("lift_arrow_base", ctypes.c_float * 3),
                ("has_nan", ctypes.c_int)]


class FcVoxParams(ctypes.Structure):
    _fields_ = [("target_resolution", ctypes.c_int), ("padding", ctypes.c_float * 3)]


class FcVoxGrid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("cell_size", ctypes.c_float),
                ("origin", ctypes.c_float * 3)]


_lib = None


def load_hip_runtime() -> None:
    """Make torch's HIP runtime the process's one.  torch bundles its own
    libamdhip64.so.7, with the same soname as ROCm's, so whichever is loaded first serves
    every later library: loading torch before libfc_lbm.so lets torch tensors and this
    library share the runtime torch was built against (torch initialised on a runtime
    loaded by another library sees no GPU)."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def load() -> ctypes.CDLL:
    """Load libfc_lbm.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(python -c 'import __graft_entry__ as g; g.build()')")
    load_hip_runtime()
    L = ctypes.CDLL(LIB_PATH)
    h = ctypes.c_void_p
    i = ctypes.c_int
    vp = ctypes.c_void_p
    L.fc_lbm_create.argtypes = [ctypes.POINTER(FcLbmDesc), ctypes.POINTER(h)]
    L.fc_lbm_destroy.argtypes = [h]
    L.fc_lbm_last_error.argtypes = [h]
    L.fc_lbm_last_error.restype = ctypes.c_char_p
    L.fc_lbm_params_default.argtypes = [ctypes.POINTER(FcLbmParams)]
    L.fc_lbm_params_default.restype = None
    L.fc_lbm_set_geometry.argtypes = [h, vp, vp]
    L.fc_lbm_set_geometry_device.argtypes = [h, vp, vp]
    L.fc_lbm_solid_mask.argtypes = [h, vp]
    L.fc_lbm_set_params.argtypes = [h, ctypes.POINTER(FcLbmParams)]
    L.fc_lbm_get_params.argtypes = [h, ctypes.POINTER(FcLbmParams)]
    L.fc_lbm_reset_uniform.argtypes = [h, i]
    L.fc_lbm_reset_from_field.argtypes = [h, vp]
    L.fc_lbm_reset_from_field_device.argtypes = [h, vp]
    L.fc_lbm_step.argtypes = [h, i]
    L.fc_lbm_step_range.argtypes = [h, i, i]
    L.fc_lbm_advance.argtypes = [h]
    L.fc_lbm_readback_field_range.argtypes = [h, i, i, vp]
    L.fc_lbm_readback_field.argtypes = [h, vp]
    L.fc_lbm_readback_field_range_noalloc.argtypes = [h, i, i, vp]
    L.fc_lbm_field_device.argtypes = [h, vp]
    L.fc_lbm_field_at.argtypes = [h, vp, ctypes.c_longlong, vp]
    L.fc_lbm_readback_stats.argtypes = [h, ctypes.POINTER(FcLbmStats)]
    L.fc_lbm_forces.argtypes = [h, vp]
    L.fc_lbm_forces_device.argtypes = [h, vp]
    L.fc_lbm_probe_health.argtypes = [h, i, ctypes.POINTER(i)]
    L.fc_lbm_has_diverged.argtypes = [h, ctypes.POINTER(i)]
    L.fc_lbm_get_populations.argtypes = [h, vp]
    L.fc_lbm_set_populations.argtypes = [h, vp]
    L.fc_lbm_sync.argtypes = [h]
    L.fc_lbm_info.argtypes = [h] + [ctypes.POINTER(i)] * 6
    L.fc_lbm_step_count.argtypes = [h]
    L.fc_lbm_step_count.restype = ctypes.c_longlong
    L.fc_lbm_population_bytes.argtypes = [h]
    L.fc_lbm_population_bytes.restype = ctypes.c_size_t
    L.fc_lbm_stream.argtypes = [h]
    L.fc_lbm_stream.restype = ctypes.c_void_p
    L.fc_lbm_boundary_links.argtypes = [h, ctypes.c_longlong, vp, vp, vp, vp, ctypes.POINTER(ctypes.c_longlong)]
    L.fc_lbm_aero_forces.argtypes = [h, ctypes.c_float, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(FcLbmAero)]
    ll = ctypes.c_longlong
    L.fc_vox_params_default.argtypes = [ctypes.POINTER(FcVoxParams)]
    L.fc_voxelize_frame.argtypes = [vp, ll, vp, ll, ctypes.POINTER(FcVoxParams), ctypes.POINTER(FcVoxGrid)]
    L.fc_voxelize.argtypes = [vp, ll, vp, ll, ctypes.POINTER(FcVoxParams), vp, vp]
    L.fc_voxelize_device.argtypes = [vp, ll, vp, ll, ctypes.POINTER(FcVoxParams), vp, vp, ll]
    L.fc_voxelize_last_error.argtypes = []
    L.fc_voxelize_last_error.restype = ctypes.c_char_p
    fl = ctypes.c_float
    L.fc_smoke_create.argtypes = [i, i, i, fl, ctypes.POINTER(fl), vp, ctypes.POINTER(h)]
    L.fc_smoke_destroy.argtypes = [h]
    L.fc_smoke_last_error.argtypes = [h]
    L.fc_smoke_last_error.restype = ctypes.c_char_p
    L.fc_smoke_reset.argtypes = [h, i, i, i, i, i]
    L.fc_smoke_scatter.argtypes = [h, i, i]
    L.fc_smoke_shift_plane.argtypes = [h, i, i]
    L.fc_smoke_step.argtypes = [h, h, i, fl, fl, i, i, i]
    L.fc_smoke_step_field.argtypes = [h, vp, i, fl, fl, i, i, i]
    L.fc_smoke_trails.argtypes = [h, vp, vp]
    L.fc_smoke_info.argtypes = [h, ctypes.POINTER(i), ctypes.POINTER(i)]
    L.fc_lbm_snapshot_begin.argtypes = [h]
    L.fc_lbm_snapshot_poll.argtypes = [h, ctypes.POINTER(i)]
    L.fc_lbm_snapshot_wait.argtypes = [h, ctypes.POINTER(ctypes.POINTER(ctypes.c_float))]
    L.fc_lbm_surface_cells.argtypes = [h]
    L.fc_lbm_surface_cells.restype = ctypes.c_longlong
    _lib = L
    return L


class FcLbmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"fc_lbm error {code}: {msg}")
        self.code = code
