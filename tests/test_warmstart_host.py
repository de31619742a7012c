"""Host-side logic of the warm-start study (no GPU): the convergence rules, the
trailing mean, the summary's fallback accounting, and the weight file format."""
from __future__ import annotations

import numpy as np
import pytest

from flowcraft_amd import warmstart as ws


def test_trailing_mean_matches_direct_window():
    x = np.random.default_rng(0).normal(size=57)
    t = ws.trailing(x, 10)
    for i in range(len(x)):
        lo = max(0, i + 1 - 10)
        assert t[i] == pytest.approx(x[lo:i + 1].mean(), rel=1e-12, abs=1e-12)


def test_convergence_rules():
    band = dict(ref=1.0, tol=0.03)
    x = np.array([5, 4, 3, 1.0, 1.0, 1.0, 1.02, 1.05, 1.1, 1.2])
    # horizon: never stays to the end
    assert ws.converged_index(x, **band) is None
    # hold 3 samples: enters at 3 and stays for 4
    assert ws.converged_index(x, **band, hold=3) == 3
    # hold 5: no run that long, and the run does not reach the end
    assert ws.converged_index(x, **band, hold=5) is None
    # reach: first in-band sample
    assert ws.converged_index(x, **band, hold=1) == 3
    y = np.array([5, 4, 3, 1.0, 1.0, 1.0, 1.05, 1.0, 1.0, 1.0])
    assert ws.converged_index(y, **band) == 7
    # a run still in the band at the end counts for 'hold' even when shorter than hold
    assert ws.converged_index(y, **band, hold=100) == 7
    assert ws.converged_index(np.ones(4), **band) == 0


def test_arm_steps_rules_and_nonfinite():
    cd = np.concatenate([np.linspace(3, 1, 50), np.ones(50)])
    steps, cda = ws._arm_steps(cd, 1.0, 0.01, 1, 10, 200)
    assert steps["reach"] <= steps["hold"] <= steps["horizon"]
    bad = cd.copy()
    bad[70] = np.nan
    steps, cda = ws._arm_steps(bad, 1.0, 0.01, 1, 10, 200)
    assert cda is None and all(v is None for v in steps.values())


def _rec(steps_by_base, ms=0.1, infer=3.0, coarse=5.0, cd=None):
    return {"steps": steps_by_base, "ms_per_step": ms, "infer_ms": infer, "coarse_ms": coarse,
            "cd_final": cd or {"plain": 1.0, "plain_mid": 1.02, "warm": 1.01, "coarse": 0.97}}


def _all(v):
    return {k: v for k in ws.RULES}


def test_summary_charges_fallbacks_the_plain_time():
    none = _all(None)
    steps1 = {b: {"plain": _all(1000), "plain_mid": _all(900), "warm": _all(250), "coarse": _all(500)}
              for b in ws.PLAIN_ARMS}
    steps2 = {b: {"plain": _all(2000), "plain_mid": _all(2000), "warm": dict(none), "coarse": dict(none)}
              for b in ws.PLAIN_ARMS}
    # the mid-speed baseline converges to its own value at 900 steps in the first shape
    steps1["plain_mid"]["plain_mid"] = _all(900)
    s = ws.summarize([_rec(steps1), _rec(steps2)], 6000)
    for rule in ws.RULES:
        o = s["plain"][rule]
        assert o["warm"]["fallbacks"] == 1 and o["coarse"]["fallbacks"] == 1
        # warm wall: 250*0.1 + 3 = 28 ms and (fallback) 2000*0.1 + 3 = 203 ms
        assert o["warm"]["ms"]["mean"] == pytest.approx((28.0 + 203.0) / 2)
        assert o["plain_ms"]["mean"] == pytest.approx((100.0 + 200.0) / 2)
        assert o["warm"]["speedup_wall_of_means"] == pytest.approx(150.0 / 115.5)
        # coarse: 500*0.1 + 5 = 55 ms and 2000*0.1 + 5 = 205 ms
        assert o["coarse"]["ms"]["mean"] == pytest.approx((55.0 + 205.0) / 2)
        m = s["plain_mid"][rule]
        assert m["plain_ms"]["mean"] == pytest.approx((90.0 + 200.0) / 2)
    d = s["final_cd_rel_diff"]
    assert d["warm_vs_plain"]["mean"] == pytest.approx(0.01)
    assert d["plain_mid_vs_plain"]["mean"] == pytest.approx(0.02)
    assert d["coarse_vs_plain_mid"]["mean"] == pytest.approx(0.05 / 1.02)


def test_summary_counts_an_unconverged_baseline():
    steps = {b: {"plain": _all(None), "plain_mid": _all(None), "warm": _all(100), "coarse": _all(100)}
             for b in ws.PLAIN_ARMS}
    s = ws.summarize([_rec(steps)], 6000)
    for base in ws.PLAIN_ARMS:
        for rule in ws.RULES:
            assert s[base][rule]["plain_unconverged"] == 1 and "warm" not in s[base][rule]


# ----------------------------------------------------- FlowCraft's coarse warm start
def _grid(nx, ny, nz, cs=1.0, origin=(0.0, 0.0, 0.0), solid=None):
    from flowcraft_amd.lbm import VoxelGrid
    sol = np.zeros((nz, ny, nx), np.uint8) if solid is None else solid
    return VoxelGrid(nx, ny, nz, sol, None, cellSize=cs, origin=origin)


def _numbered_field(g):
    """A field whose value names its cell: (index, x, y, z) per cell."""
    n = g.nx * g.ny * g.nz
    idx = np.arange(n)
    x, y, z = idx % g.nx, (idx // g.nx) % g.ny, idx // (g.nx * g.ny)
    return np.stack([idx, x, y, z], axis=1).astype(np.float32).reshape(-1)


def test_sampling_onto_the_same_frame_is_the_identity():
    """nx != ny: without the grid/world axis swap of VoxelGrid::CellCenter and its inverse
    a fine cell would read another cell (or clamp), so identity pins the swap."""
    g = _grid(7, 4, 3, cs=0.5, origin=(1.0, -2.0, 3.0))
    f = _numbered_field(g)
    assert np.array_equal(ws.sample_coarse_field(g, g, f), f)


def test_sampling_a_coarser_aligned_frame_takes_the_containing_cell():
    fine = _grid(8, 6, 4)
    coarse = _grid(4, 3, 2, cs=2.0)
    out = ws.sample_coarse_field(fine, coarse, _numbered_field(coarse)).reshape(4, 6, 8, 4)
    z, y, x = np.meshgrid(np.arange(4), np.arange(6), np.arange(8), indexing="ij")
    assert np.array_equal(out[..., 1], x // 2) and np.array_equal(out[..., 2], y // 2)
    assert np.array_equal(out[..., 3], z // 2)


def test_sampling_shifts_with_the_world_frame():
    """Moving the coarse origin by one coarse cell along world X (the grid-y axis) moves the
    sampled coarse grid-y index, clamped at the edge; grid-x is untouched."""
    fine = _grid(8, 6, 4)
    coarse = _grid(4, 3, 2, cs=2.0, origin=(2.0, 0.0, 0.0))
    out = ws.sample_coarse_field(fine, coarse, _numbered_field(coarse)).reshape(4, 6, 8, 4)
    z, y, x = np.meshgrid(np.arange(4), np.arange(6), np.arange(8), indexing="ij")
    assert np.array_equal(out[..., 2], np.clip((y - 2) // 2, 0, 2))
    assert np.array_equal(out[..., 1], x // 2)
    # and along world Y (grid-x)
    coarse2 = _grid(4, 3, 2, cs=2.0, origin=(0.0, -4.0, 0.0))
    out2 = ws.sample_coarse_field(fine, coarse2, _numbered_field(coarse2)).reshape(4, 6, 8, 4)
    assert np.array_equal(out2[..., 1], np.clip((x + 4) // 2, 0, 3)) and np.array_equal(out2[..., 2], y // 2)


def test_sampling_a_solid_coarse_cell_gives_quiescent_fluid():
    solid = np.zeros((2, 3, 4), np.uint8)
    solid[1, 2, 3] = 1
    fine = _grid(8, 6, 4)
    coarse = _grid(4, 3, 2, cs=2.0, solid=solid)
    out = ws.sample_coarse_field(fine, coarse, _numbered_field(coarse) + 5.0).reshape(4, 6, 8, 4)
    blk = out[2:4, 4:6, 6:8].reshape(-1, 4)
    assert np.all(blk == np.array([1, 0, 0, 0], np.float32))
    assert np.all(out[0:2, 0:2, 0:2, 0] == 5.0)


def test_coarse_grid_frame_and_shell():
    """FlowCraft's frame at 1/10 of the 256x128x128 lattice: resolution 25 along the longest
    axis, cell 10.24, 25x13x13 cells; every coarse cell within sqrt(3)/2 cells of the
    surface is solid, on both sides of it, and no fluid cell is nearer."""
    from flowcraft_amd import geometry
    rng = ws.shape_stream(7)
    for _ in range(4):
        g, d = geometry.random_shape_grid(256, 128, 128, rng)
        cg = ws.flowcraft_coarse_grid(d, (256, 128, 128), 10)
        assert (cg.nx, cg.ny, cg.nz) == (25, 13, 13)
        assert cg.cellSize == pytest.approx(10.24, rel=1e-6) and cg.origin == (0.0, 0.0, 0.0)
        sdf = cg.signedDistance
        r = 0.8660254 * cg.cellSize
        assert np.all(sdf[cg.solid == 1] <= r + 1e-4) and np.all(sdf[cg.solid == 0] > r - 1e-4)
        assert np.any((cg.solid == 1) & (sdf > 0))  # the shell's fluid side
        assert cg.solid.sum() > 0


def test_weight_file_keeps_bf16_bits(tmp_path):
    """The saved conv weights are the bf16 values the kernels consume (round to nearest
    even of the fp32 master weights)."""
    class FakeNet:
        def __init__(self, sd):
            self.sd = sd

        def state_dict(self):
            return self.sd

        def load_state_dict(self, sd):
            self.sd = sd

    rng = np.random.default_rng(3)
    w = rng.normal(size=(16, 27, 8)).astype(np.float32)
    sd = {"e0.w": w, "e0.b": rng.normal(size=16).astype(np.float32), "head.w": rng.normal(size=(4, 16)).astype(np.float32),
          "head.b": np.zeros(4, np.float32)}
    p = tmp_path / "w.npz"
    ws.save_weights(FakeNet(sd), str(p))
    net = FakeNet({})
    ws.load_weights(net, str(p))
    import torch
    expect = torch.tensor(w).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(net.sd["e0.w"], expect)
    assert np.array_equal(net.sd["head.w"], sd["head.w"]) and np.array_equal(net.sd["e0.b"], sd["e0.b"])


def test_committed_weights_cover_every_layer():
    from flowcraft_amd.surrogate import LAYERS
    with np.load(ws.DEFAULT_WEIGHTS, allow_pickle=False) as z:
        keys = set(z.files)
        for name, mode, ins, kin, c1, cout, stride, lvl in LAYERS:
            assert z[f"{name}.w.bf16"].shape == (cout, 27, kin)
            assert z[f"{name}.b"].shape == (cout,)
        assert z["head.w"].shape == (4, 16)
    assert "head.b" in keys
