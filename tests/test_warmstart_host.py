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


def _rec(plain, warm, coarse, ms=0.1, infer=3.0):
    return {"plain": plain, "warm": warm, "coarse": coarse, "ms_per_step": ms, "infer_ms": infer, "coarse_ms": 5.0,
            "cd_ref": 1.0, "warm_cd_final": 1.01, "cl_ref": 0.0, "warm_cl_final": 0.0}


def test_summary_charges_fallbacks_the_plain_time():
    r = {k: None for k in ws.RULES}
    recs = [_rec({k: 1000 for k in ws.RULES}, {k: 250 for k in ws.RULES}, {k: 500 for k in ws.RULES}),
            _rec({k: 2000 for k in ws.RULES}, dict(r), dict(r))]
    s = ws.summarize(recs, 6000)
    for rule in ws.RULES:
        o = s[rule]
        assert o["warm_fallbacks"] == 1 and o["coarse_fallbacks"] == 1
        # warm wall: 250*0.1 + 3 = 28 ms and (fallback) 2000*0.1 + 3 = 203 ms
        assert o["warm_ms"]["mean"] == pytest.approx((28.0 + 203.0) / 2)
        assert o["plain_ms"]["mean"] == pytest.approx((100.0 + 200.0) / 2)
        assert o["speedup_wall_of_means"] == pytest.approx(150.0 / 115.5)
    assert s["final_cd_rel_diff_warm_vs_plain"]["mean"] == pytest.approx(0.01)


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
