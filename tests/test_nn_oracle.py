"""Pins the FP64 NumPy surrogate oracle (oracle/nn_oracle.py): central finite
differences of the whole network's loss against its hand-written backward, and
closed-form special cases of the convolution and its adjoint."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import nn_oracle as no
from flowcraft_amd.surrogate import LAYERS

# a slimmer copy of the production layer list (same structure, fewer channels) keeps
# finite differences cheap
SMALL = [
    ("e0", 0, ("x",), 8, 0, 8, 1, 0),
    ("e1", 0, ("e0",), 8, 0, 8, 2, 1),
    ("e1b", 0, ("e1",), 8, 0, 8, 1, 1),
    ("e2", 0, ("e1b",), 8, 0, 8, 2, 2),
    ("e2b", 0, ("e2",), 8, 0, 8, 1, 2),
    ("e3", 0, ("e2b",), 8, 0, 8, 2, 3),
    ("e3b", 0, ("e3",), 8, 0, 8, 1, 3),
    ("e3c", 0, ("e3b",), 8, 0, 8, 1, 3),
    ("d2", 1, ("e3c", "e2b"), 16, 8, 8, 1, 2),
    ("d1", 1, ("d2", "e1b"), 16, 8, 8, 1, 1),
    ("d0", 1, ("d1", "e0"), 16, 8, 8, 1, 0),
]


def make_params(layers, rng, scale=0.3):
    params = {}
    for name, mode, ins, kin, c1, cout, stride, lvl in layers:
        params[name] = (rng.normal(0, scale / np.sqrt(27 * kin), (cout, 27, kin)), rng.normal(0, 0.05, cout))
    hw = rng.normal(0, 0.3, (4, layers[-1][5]))
    hb = rng.normal(0, 0.1, 4)
    return params, hw, hb


def test_structure_matches_production_layers():
    assert [l[:3] for l in SMALL] == [l[:3] for l in LAYERS]
    for name, mode, ins, kin, c1, cout, stride, lvl in LAYERS:
        assert kin % 8 == 0 and cout % 8 == 0 and (mode == 0 or (c1 % 8 == 0 and (kin - c1) % 8 == 0))


def test_identity_kernel_and_adjoint_dot_product():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(1, 6, 4, 8, 8))
    w = np.zeros((8, 27, 8))
    w[:, 13, :] = np.eye(8)  # centre tap
    assert np.allclose(no.conv_fwd(x, w, 0.0, 1), x)
    # <conv(x), y> == <x, conv_adj(y)> for both strides (adjointness)
    for s in (1, 2):
        w = rng.normal(size=(5, 27, 8))
        y = rng.normal(size=(1, no.out_dims(6, s), no.out_dims(4, s), no.out_dims(8, s), 5))
        lhs = (no.conv_fwd(x, w, 0.0, s) * y).sum()
        rhs = (x * no.conv_adj(y, w, x.shape, s)).sum()
        assert abs(lhs - rhs) < 1e-9 * abs(lhs)
        # wgrad is the derivative of <conv(x, w), y> in w
        dw = no.conv_wgrad(x, y, s)
        e = rng.normal(size=w.shape)
        h = 1e-6
        fd = ((no.conv_fwd(x, w + h * e, 0.0, s) * y).sum() - (no.conv_fwd(x, w - h * e, 0.0, s) * y).sum()) / (2 * h)
        assert abs(fd - (dw * e).sum()) < 1e-6 * abs(fd)
    # up2 / down2_sum are adjoint too
    a = rng.normal(size=(1, 3, 2, 4, 8))
    b = rng.normal(size=(1, 6, 4, 8, 8))
    assert abs((no.up2(a) * b).sum() - (a * no.down2_sum(b)).sum()) < 1e-9


@pytest.mark.parametrize("leak", [1.0, no.LEAK])
def test_network_backward_matches_finite_differences(leak, monkeypatch):
    """leak = 1 makes the network linear (finite differences exact to rounding); with the
    production slope a perturbation can cross a ReLU kink, so the step is kept tiny."""
    monkeypatch.setattr(no, "LEAK", leak)
    rng = np.random.default_rng(1)
    params, hw, hb = make_params(SMALL, rng)
    x = rng.normal(size=(1, 16, 8, 8, 8))
    tgt = rng.normal(size=(1, 16, 8, 8, 4))
    wm = (rng.random((1, 16, 8, 8)) > 0.2).astype(np.float64)
    acts, pred = no.forward(params, x, SMALL, hw, hb)
    grads = no.backward(params, acts, pred, tgt, wm, SMALL, hw)

    def loss_with(name, kind, delta):
        p2 = dict(params)
        hw2, hb2 = hw, hb
        if name == "head":
            if kind == "w":
                hw2 = hw + delta
            else:
                hb2 = hb + delta
        else:
            w, b = params[name]
            p2[name] = (w + delta, b) if kind == "w" else (w, b + delta)
        _, pr = no.forward(p2, x, SMALL, hw2, hb2)
        return no.loss_fn(pr, tgt, wm)

    h = 1e-6
    for name in ["e0", "e1", "e2b", "e3", "e3c", "d2", "d1", "d0", "head"]:
        for kind in ("w", "b"):
            shape = (hw if kind == "w" else hb).shape if name == "head" else params[name][0 if kind == "w" else 1].shape
            e = rng.normal(size=shape)
            fd = (loss_with(name, kind, h * e) - loss_with(name, kind, -h * e)) / (2 * h)
            an = (grads[f"{name}.{kind}"] * e).sum()
            assert abs(fd - an) <= 1e-4 * abs(fd) + 2e-10, (name, kind, fd, an)
