"""ORACLE — TEST INFRASTRUCTURE ONLY.

FP64 NumPy restatement of the warm-start surrogate (flowcraft_amd/surrogate.py):
3x3x3 convolutions (padding 1, stride 1 or 2) on channels-last tensors, the fused
nearest-upsample + concat input of the decoder convolutions, leaky ReLU, the 1x1
head and the masked MSE loss — forward and hand-written backward.  The network
has no counterpart in FlowCraft; this module is pinned by central finite
differences (tests/test_nn_oracle.py) and by closed-form special cases, and the
GPU kernels are checked against it.

Tensors: x[n, z, y, x, c]; conv weights w[cout, 27, cin] with tap = (kz*3+ky)*3+kx.
"""
from __future__ import annotations

import numpy as np

LEAK = 0.05


def bf16(x):
    """Round float values to bfloat16 (round-to-nearest-even), returned as float64."""
    a = np.asarray(x, np.float64).astype(np.float32)
    b = a.view(np.uint32).astype(np.uint64)
    r = (b + 0x7FFF + ((b >> 16) & 1)) >> 16
    out = (r << 16).astype(np.uint32).view(np.float32)
    return out.astype(np.float64)


def f32(x):
    return np.asarray(x, np.float64).astype(np.float32).astype(np.float64)


def _ident(x):
    return x


def taps():
    for t in range(27):
        yield t, t // 9, (t // 3) % 3, t % 3


def out_dims(d, s):
    return (d - 1) // s + 1


def conv_fwd(x, w, b, s=1):
    n, D, H, W, C = x.shape
    Do, Ho, Wo = out_dims(D, s), out_dims(H, s), out_dims(W, s)
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((n, Do, Ho, Wo, w.shape[0]), np.float64)
    for t, kz, ky, kx in taps():
        patch = xp[:, kz:kz + s * (Do - 1) + 1:s, ky:ky + s * (Ho - 1) + 1:s, kx:kx + s * (Wo - 1) + 1:s, :]
        y += np.einsum("ndhwc,oc->ndhwo", patch, w[:, t, :])
    return y + b


def conv_adj(dy, w, in_shape, s=1):
    """Data gradient of conv_fwd: the adjoint gather."""
    n, D, H, W, C = in_shape
    _, Do, Ho, Wo, _ = dy.shape
    dxp = np.zeros((n, D + 2, H + 2, W + 2, C), np.float64)
    for t, kz, ky, kx in taps():
        dxp[:, kz:kz + s * (Do - 1) + 1:s, ky:ky + s * (Ho - 1) + 1:s, kx:kx + s * (Wo - 1) + 1:s, :] += \
            np.einsum("ndhwo,oc->ndhwc", dy, w[:, t, :])
    return dxp[:, 1:D + 1, 1:H + 1, 1:W + 1, :]


def conv_wgrad(x, dy, s=1):
    n, D, H, W, C = x.shape
    _, Do, Ho, Wo, Co = dy.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    dw = np.zeros((Co, 27, C), np.float64)
    for t, kz, ky, kx in taps():
        patch = xp[:, kz:kz + s * (Do - 1) + 1:s, ky:ky + s * (Ho - 1) + 1:s, kx:kx + s * (Wo - 1) + 1:s, :]
        dw[:, t, :] = np.einsum("ndhwo,ndhwc->oc", dy, patch)
    return dw


def up2(x):
    return x.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)


def down2_sum(x):
    n, D, H, W, C = x.shape
    return x.reshape(n, D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4, 6))


def leaky(x):
    return np.where(x > 0, x, LEAK * x)


def leaky_grad(y, g):
    return np.where(y > 0, g, LEAK * g)


def forward(params, x, layers, head_w, head_b, quant=False):
    """params: {name: (w, b)}; x: [n, D, H, W, 8].  Returns (acts, pred).
    quant=True rounds every activation to bf16 where the GPU stores one (the weights
    passed in are expected to be bf16 values already)."""
    q = bf16 if quant else _ident
    acts = {"x": x}
    for name, mode, ins, kin, c1, cout, stride, lvl in layers:
        w, b = params[name]
        if mode == 0:
            acts[name] = q(leaky(conv_fwd(acts[ins[0]], w, b, stride)))
        else:
            cat = np.concatenate([up2(acts[ins[0]]), acts[ins[1]]], axis=-1)
            acts[name] = q(leaky(conv_fwd(cat, w, b, 1)))
    pred = np.einsum("ndhwc,oc->ndhwo", acts["d0"], head_w) + head_b
    return acts, pred


def loss_fn(pred, tgt, wmask):
    V = pred.shape[0] * pred.shape[1] * pred.shape[2] * pred.shape[3]
    e = pred - tgt
    return float((wmask[..., None] * e * e).sum() / (4.0 * V))


def backward(params, acts, pred, tgt, wmask, layers, head_w, quant=False):
    """Gradients of loss_fn(forward(...)) for every parameter.  quant=True rounds each
    activation gradient to bf16 where the GPU backward stores one: the head's dh, dz
    after the leaky-ReLU derivative, each data gradient, each accumulation, and the
    two halves of a concatenated input's gradient."""
    q = bf16 if quant else _ident
    V = pred.shape[0] * pred.shape[1] * pred.shape[2] * pred.shape[3]
    g = 2.0 * wmask[..., None] * (pred - tgt) / (4.0 * V)
    grads = {"head.w": np.einsum("ndhwo,ndhwc->oc", g, acts["d0"]), "head.b": g.sum(axis=(0, 1, 2, 3))}
    dact = {"d0": q(np.einsum("ndhwo,oc->ndhwc", g, head_w))}

    def accumulate(key, v):
        dact[key] = q(dact[key] + v) if key in dact else v

    for name, mode, ins, kin, c1, cout, stride, lvl in reversed(layers):
        w, b = params[name]
        dz = q(leaky_grad(acts[name], dact[name]))
        grads[name + ".b"] = dz.sum(axis=(0, 1, 2, 3))
        if mode == 0:
            src = acts[ins[0]]
            grads[name + ".w"] = conv_wgrad(src, dz, stride)
            if name != "e0":
                accumulate(ins[0], q(conv_adj(dz, w, src.shape, stride)))
        else:
            coarse, skip = acts[ins[0]], acts[ins[1]]
            cat = np.concatenate([up2(coarse), skip], axis=-1)
            grads[name + ".w"] = conv_wgrad(cat, dz, 1)
            dcat = q(conv_adj(dz, w, cat.shape, 1))
            dact[ins[1]] = q(dact.get(ins[1], 0) + dcat[..., c1:])
            accumulate(ins[0], q(down2_sum(dcat[..., :c1])))
    return grads
