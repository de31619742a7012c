"""The C-ABI library loads (no GPU needed) and exports every symbol the header declares."""
from __future__ import annotations

import ctypes
import os
import re

import pytest

from flowcraft_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "flowcraft", "fc_lbm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fc_(?:lbm|vox|voxelize|smoke)_?[a-z_0-9]*)\s*\(", text)))


def test_header_and_binding_agree():
    assert sorted(_lib.EXPORTS) == header_functions()


def test_library_loads_and_exports_all_symbols():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail(f"{_lib.LIB_PATH} not built: run __graft_entry__.build()")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in header_functions():
        assert hasattr(lib, name), name


def test_params_default_matches_reference_defaults():
    """LbmSolver::Params defaults (FlowCraft/LbmSolver.h:61-100); no GPU call."""
    L = _lib.load()
    p = _lib.FcLbmParams()
    L.fc_lbm_params_default(ctypes.byref(p))
    assert abs(p.tau - 0.8) < 1e-7 and abs(p.inlet_velocity - 0.05) < 1e-7
    assert p.reverse_flow == 0 and p.free_slip_y == 1 and p.free_slip_z == 1
    assert abs(p.smagorinsky_cs - 0.12) < 1e-7


def test_null_handle_errors_are_reported():
    L = _lib.load()
    assert L.fc_lbm_step(None, 1) == _lib.FC_LBM_EINVAL
    assert L.fc_lbm_destroy(None) == _lib.FC_LBM_EINVAL
    assert L.fc_lbm_last_error(None) == b"null fc_lbm handle"
    desc = _lib.FcLbmDesc(2, 1, 1, 1, 0, 0, 0, 0, None)
    h = ctypes.c_void_p()
    assert L.fc_lbm_create(ctypes.byref(desc), ctypes.byref(h)) == _lib.FC_LBM_EINVAL
    assert b"at least" in L.fc_lbm_last_error(h)
    L.fc_lbm_destroy(h)
