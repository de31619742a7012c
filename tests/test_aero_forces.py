"""FieldOverlay::ComputeAerodynamicForces (FlowCraft/FieldOverlay.cpp:469-621) — the
aerodynamic summary FlowCraft draws (force sums, stress colour range, centre of
pressure, lift-arrow base) — restated in the oracle and pinned three ways:

* closed forms: planted fields with hand-computed results, one per branch of the
  centre-of-pressure rule (|Fz| weights, |F| weights, solid centroid, nothing) and of
  the lift-arrow base (top of the 5x5 columns under the centre, else the top layer);
* FlowCraft's own function, compiled from FieldOverlay.cpp into oracle/_ref, on the
  same planted fields and on fields of real flows (the reference sphere case and the
  inclined-wing case), in two world frames;
* the symmetric-body centroid: a sphere's solid cells average to its centre.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from tests._cases import case_a_grid, case_a_params, make_oracle, oracle_params, ulp_diff32, wing_grid

needs_ref = pytest.mark.skipif(not oracle.ref_overlay_available(),
                               reason="oracle/_ref/libref_overlay.so is not built (make -C oracle ref)")

FRAMES = [(1.0, (0.0, 0.0, 0.0)), (2.5, (-30.0, 12.25, -7.0))]


def _same(a: dict, b: dict, ulps: int = 2):
    """Equal up to the rounding of differently ordered FP64 sums: float results within
    `ulps` units in the last place, the stress range and the flags exactly."""
    for k in ("forceX", "forceY", "forceZ"):
        assert ulp_diff32(np.float32([a[k]]), np.float32([b[k]])) <= ulps, (k, a[k], b[k])
    assert ulp_diff32(a["cop"], b["cop"]) <= ulps, (a["cop"], b["cop"])
    assert ulp_diff32(a["liftArrowBase"], b["liftArrowBase"]) <= ulps, (a["liftArrowBase"], b["liftArrowBase"])
    assert a["minStress"] == b["minStress"] and a["maxStress"] == b["maxStress"]
    assert a["hasNaN"] == b["hasNaN"]


def _two_cell_case():
    solid = np.zeros((6, 8, 10), np.uint8)  # [z, y, x]
    field = np.zeros((6, 8, 10, 4), np.float32)
    solid[2, 5, 3] = 1  # A: x=3, y=5, z=2
    solid[4, 1, 7] = 1  # B: x=7, y=1, z=4
    field[2, 5, 3] = (0.5, 0.0, 0.0, 1.0)
    field[4, 1, 7] = (2.0, 0.0, 0.0, 3.0)
    return solid, field


def test_closed_form_lift_weighted_centre():
    # cell centres, world = origin + ((y+.5)cs, (x+.5)cs, (z+.5)cs) with cs 2, origin (1,2,3):
    # A (12, 9, 8), B (4, 17, 12); |Fz| weights 1 and 3 -> cop (6, 15, 11).  The 5x5 columns
    # around grid (x=6, y=2) contain B (top z 4): base z = 3 + (4+1)*2 = 13.
    solid, field = _two_cell_case()
    r = oracle.aero_forces(solid, field, 2.0, (1.0, 2.0, 3.0))
    assert (r["forceX"], r["forceY"], r["forceZ"]) == (0.0, 0.0, 4.0)
    assert r["minStress"] == 0.0 and r["maxStress"] == 2.0 and not r["hasNaN"]
    assert r["cop"].tolist() == [6.0, 15.0, 11.0]
    assert r["liftArrowBase"].tolist() == [6.0, 15.0, 13.0]


def test_closed_form_force_weighted_and_centroid_fallbacks():
    solid, field = _two_cell_case()
    field[..., 3] = 0.0
    field[2, 5, 3, 1:3] = (3.0, 4.0)  # |F| = 5
    field[4, 1, 7, 1:3] = (0.0, 5.0)  # |F| = 5: equal weights -> the midpoint (8, 13, 10)
    r = oracle.aero_forces(solid, field, 2.0, (1.0, 2.0, 3.0))
    assert r["cop"].tolist() == [8.0, 13.0, 10.0]
    field[..., 1:] = 0.0  # no force anywhere: the centroid of the solid cells, the same midpoint
    r = oracle.aero_forces(solid, field, 2.0, (1.0, 2.0, 3.0))
    assert r["cop"].tolist() == [8.0, 13.0, 10.0]
    assert (r["forceX"], r["forceY"], r["forceZ"]) == (0.0, 0.0, 0.0)


def test_closed_form_hollow_column_and_empty_grid():
    # a ring in the x-y plane at z = 3: its centre column is empty, so the base sits on the
    # top solid layer of the whole body (z = 3) instead
    nz, ny, nx = 6, 21, 21
    solid = np.zeros((nz, ny, nx), np.uint8)
    yy, xx = np.mgrid[0:ny, 0:nx]
    ring = (np.hypot(xx - 10, yy - 10) > 6) & (np.hypot(xx - 10, yy - 10) < 8)
    solid[3][ring] = 1
    field = np.zeros((nz, ny, nx, 4), np.float32)
    r = oracle.aero_forces(solid, field)
    assert r["cop"].tolist() == [10.5, 10.5, 3.5]  # centroid: the ring is symmetric about (10.5, 10.5)
    assert r["liftArrowBase"].tolist() == [10.5, 10.5, 4.0]
    assert r["maxStress"] == np.float32(1e-6) and r["minStress"] == 0.0
    e = oracle.aero_forces(np.zeros((3, 4, 5), np.uint8), np.zeros((3, 4, 5, 4), np.float32), 1.5, (1, 2, 3))
    assert e["cop"].tolist() == [0.0, 0.0, 0.0] and e["liftArrowBase"].tolist() == [0.0, 0.0, 4.5]


def test_nonfinite_cells_are_flagged_and_skipped():
    solid, field = _two_cell_case()
    field[2, 5, 3, 1] = np.nan
    r = oracle.aero_forces(solid, field, 2.0, (1.0, 2.0, 3.0))
    assert r["hasNaN"] and r["forceZ"] == 3.0 and r["cop"].tolist() == [4.0, 17.0, 12.0]


def test_sphere_centroid_is_its_centre():
    g = case_a_grid()  # r = 5 at (20, 12, 12): voxel set symmetric about the centre
    r = oracle.aero_forces(g.solid, np.zeros(g.solid.size * 4, np.float32))
    assert r["cop"].tolist() == [12.0, 20.0, 12.0]  # world X = grid y, world Y = grid x


@needs_ref
def test_planted_fields_match_flowcraft():
    rng = np.random.default_rng(7)
    cases = []
    solid, field = _two_cell_case()
    cases.append((solid, field.copy()))
    f2 = field.copy()
    f2[..., 3] = 0.0
    f2[2, 5, 3, 1:3] = (3.0, 4.0)
    cases.append((solid, f2))
    cases.append((solid, np.zeros_like(field)))
    f3 = field.copy()
    f3[4, 1, 7, 0] = np.inf
    cases.append((solid, f3))
    sol = (rng.random((9, 11, 13)) < 0.3).astype(np.uint8)
    fr = rng.normal(size=(9, 11, 13, 4)).astype(np.float32)
    cases.append((sol, fr))
    cases.append((np.zeros((3, 4, 5), np.uint8), np.zeros((3, 4, 5, 4), np.float32)))
    for s, f in cases:
        for cs, o in FRAMES:
            _same(oracle.aero_forces(s, f, cs, o), oracle.ref_aero_forces(s, f, cs, o))


@needs_ref
@pytest.mark.parametrize("case", ["sphere", "wing"])
def test_flow_fields_match_flowcraft(case):
    g = case_a_grid() if case == "sphere" else wing_grid()
    orc = make_oracle(g)
    oracle_params(orc, case_a_params())
    orc.reset(True)
    orc.step(30)
    field = orc.field()
    for cs, o in FRAMES:
        a = oracle.aero_forces(g.solid, field, cs, o)
        _same(a, oracle.ref_aero_forces(g.solid, field, cs, o))
        assert a["forceX"] > 0.0 and a["maxStress"] > 0.0
