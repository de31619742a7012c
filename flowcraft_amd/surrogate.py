"""Learned warm start: a 3D convolutional encoder-decoder that maps an obstacle
(solid mask + signed distance) and the inflow speed to an initial (rho, u) field,
which seeds the LBM through ``LbmSolver.ResetFromField`` — the hook FlowCraft's
own coarse-grid warm start uses (FlowCraft/main.cpp:1496-1564).

Everything numeric runs in libfc_nn.so (hand-written BF16 MFMA kernels); torch
only allocates device memory and provides the stream.  No autograd: the backward
pass is written out layer by layer.

Network (channels-last bf16 activations, fp32 master weights, fp32 accumulate):
  level 0 (full res)   e0  conv 8->16
  level 1 (1/2)        e1  conv s2 16->32,  e1b conv 32->32
  level 2 (1/4)        e2  conv s2 32->64,  e2b conv 64->64
  level 3 (1/8)        e3  conv s2 64->128, e3b, e3c conv 128->128
  decoder              d2 = conv(up2(e3c) ++ e2b) 192->64
                       d1 = conv(up2(d2)  ++ e1b)  96->32
                       d0 = conv(up2(d1)  ++ e0)   48->16
  head                 1x1 conv 16->4 (fp32) -> ((rho-1)*3/u^2, u/u_in)
All 3x3x3 convs: padding 1, leaky ReLU (slope 0.05).  ~1.75M parameters.
Input features (8 channels): solid, clip(sdf, +-16)/16, u_in/0.1, x, y, z
(normalised cell-centre coordinates), 0, 0.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

from . import _lib

NN_PATH = os.path.join(_lib.LIB_DIR, "libfc_nn.so")
FWD, UPCAT, ADJ = 0, 1, 2

# name, mode, input(s), kin, c1 (upcat coarse channels), cout, stride, level of the output
LAYERS = [
    ("e0", FWD, ("x",), 8, 0, 16, 1, 0),
    ("e1", FWD, ("e0",), 16, 0, 32, 2, 1),
    ("e1b", FWD, ("e1",), 32, 0, 32, 1, 1),
    ("e2", FWD, ("e1b",), 32, 0, 64, 2, 2),
    ("e2b", FWD, ("e2",), 64, 0, 64, 1, 2),
    ("e3", FWD, ("e2b",), 64, 0, 128, 2, 3),
    ("e3b", FWD, ("e3",), 128, 0, 128, 1, 3),
    ("e3c", FWD, ("e3b",), 128, 0, 128, 1, 3),
    ("d2", UPCAT, ("e3c", "e2b"), 192, 128, 64, 1, 2),
    ("d1", UPCAT, ("d2", "e1b"), 96, 64, 32, 1, 1),
    ("d0", UPCAT, ("d1", "e0"), 48, 32, 16, 1, 0),
]
HEAD_IN = 16
IN_CH = 8

_nn = None


def nn_lib() -> ctypes.CDLL:
    global _nn
    if _nn is not None:
        return _nn
    if not os.path.exists(NN_PATH):
        raise RuntimeError(f"{NN_PATH} is missing: build with __graft_entry__.build()")
    _lib.load_hip_runtime()
    L = ctypes.CDLL(NN_PATH)
    i, ll, vp, f = ctypes.c_int, ctypes.c_longlong, ctypes.c_void_p, ctypes.c_float
    L.fc_nn_conv.argtypes = [i, vp, vp, vp, vp, vp, vp] + [i] * 12 + [vp]
    L.fc_nn_wgrad.argtypes = [i, vp, vp, vp, vp, vp, i, i] + [i] * 11 + [vp]
    L.fc_nn_colsum.argtypes = [vp, ll, i, vp, vp, i, vp]
    L.fc_nn_leaky_backward.argtypes = [vp, vp, vp, ll, vp]
    L.fc_nn_leaky_backward_colsum.argtypes = [vp, vp, vp, ll, i, vp, vp, i, vp]
    L.fc_nn_col_blocks.argtypes = []
    L.fc_nn_transpose_w.argtypes = [vp, vp, i, i, vp]
    L.fc_nn_flip_transpose_w.argtypes = [vp, vp, i, i, vp]
    L.fc_nn_upcat_split.argtypes = [vp, i, i, i, i, i, i, vp, vp, vp]
    L.fc_nn_head_loss.argtypes = [vp, i, vp, vp, vp, vp, ll, f, vp, vp, vp, vp, vp]
    L.fc_nn_adam.argtypes = [vp, vp, vp, vp, vp, ll, f, f, f, f, i, f, vp]
    L.fc_nn_add_bf16.argtypes = [vp, vp, ll, vp]
    L.fc_nn_f32_to_bf16.argtypes = [vp, vp, ll, vp]
    L.fc_nn_bf16_to_f32.argtypes = [vp, vp, ll, vp]
    L.fc_nn_features.argtypes = [vp, vp, f, i, i, i, i, vp, vp]
    L.fc_nn_targets.argtypes = [vp, vp, f, ll, vp, vp, vp]
    L.fc_nn_pred_to_field.argtypes = [vp, f, ll, vp, vp]
    _nn = L
    return L


def wgrad_splits(V: int, K: int, cout: int) -> int:
    """Voxel splits of a weight-gradient launch: enough workgroups (>= 8192, or the 256
    splits a 64x64 tile gets) to hide the gather latency, at least 512 voxels per split.
    The kernel's output tile is 16x256, 32x128 or 64x64 (co x k) by the channel count."""
    tco = 16 if cout <= 16 else (32 if cout <= 32 else 64)
    tiles = -(-K // (4096 // tco)) * -(-cout // tco)
    want = max(min(256, V // 2048), -(-8192 // tiles))
    return int(max(1, min(want, V // 512)))


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with code {rc}")


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


class Surrogate:
    """The warm-start network: parameters, forward, backward, Adam."""

    def __init__(self, seed: int = 0, device: str = "cuda"):
        import torch
        self.torch = torch
        self.dev = device
        rng = np.random.Generator(np.random.PCG64(seed))
        self.params = {}
        for name, mode, ins, kin, c1, cout, stride, lvl in LAYERS:
            fan_in = 27 * kin
            bound = math.sqrt(6.0 / ((1 + 0.05 ** 2) * fan_in))
            w = rng.uniform(-bound, bound, size=(cout, 27, kin)).astype(np.float32)
            self._add(name, w, np.zeros(cout, np.float32))
        wh = rng.uniform(-1 / math.sqrt(HEAD_IN), 1 / math.sqrt(HEAD_IN), size=(4, HEAD_IN)).astype(np.float32)
        self.head_w = torch.tensor(wh, device=device)
        self.head_b = torch.zeros(4, dtype=torch.float32, device=device)
        self.head_gw = torch.zeros(4 * HEAD_IN + 5, dtype=torch.float32, device=device)
        self.head_state = [torch.zeros_like(self.head_w), torch.zeros_like(self.head_w),
                           torch.zeros_like(self.head_b), torch.zeros_like(self.head_b)]
        self.step_count = 0
        self._ws = None

    def _add(self, name, w, b):
        t = self.torch
        wt = t.tensor(w, device=self.dev)
        p = {"w": wt, "wbf": wt.to(t.bfloat16), "b": t.tensor(b, device=self.dev),
             "gw": t.zeros_like(wt), "gb": t.zeros(b.shape, dtype=t.float32, device=self.dev),
             "mw": t.zeros_like(wt), "vw": t.zeros_like(wt), "mb": t.zeros(b.shape, dtype=t.float32, device=self.dev),
             "vb": t.zeros(b.shape, dtype=t.float32, device=self.dev)}
        self.params[name] = p

    def num_parameters(self) -> int:
        n = sum(p["w"].numel() + p["b"].numel() for p in self.params.values())
        return n + self.head_w.numel() + self.head_b.numel()

    # ----------------------------------------------------------------- io
    def state_dict(self) -> dict:
        out = {f"{k}.w": p["w"].cpu().numpy() for k, p in self.params.items()}
        out.update({f"{k}.b": p["b"].cpu().numpy() for k, p in self.params.items()})
        out["head.w"] = self.head_w.cpu().numpy()
        out["head.b"] = self.head_b.cpu().numpy()
        return out

    def load_state_dict(self, sd: dict) -> None:
        t = self.torch
        for k, p in self.params.items():
            p["w"].copy_(t.tensor(np.asarray(sd[f"{k}.w"], np.float32)))
            p["b"].copy_(t.tensor(np.asarray(sd[f"{k}.b"], np.float32)))
            p["wbf"].copy_(p["w"].to(t.bfloat16))
        self.head_w.copy_(t.tensor(np.asarray(sd["head.w"], np.float32)))
        self.head_b.copy_(t.tensor(np.asarray(sd["head.b"], np.float32)))

    def save(self, path: str) -> None:
        np.savez(path, **self.state_dict())

    def load(self, path: str) -> None:
        with np.load(path, allow_pickle=False) as z:
            self.load_state_dict({k: z[k] for k in z.files})

    # ------------------------------------------------------------ helpers
    def _stream(self):
        return ctypes.c_void_p(self.torch.cuda.current_stream().cuda_stream)

    @staticmethod
    def level_dims(dims, lvl):
        d, h, w = dims
        s = 1 << lvl
        return d // s, h // s, w // s

    def features(self, solid, sdf, u_in: float, nb: int, dims):
        """solid (uint8) and sdf (float32) device tensors of nb*D*H*W -> [V][8] bf16."""
        t = self.torch
        d, h, w = dims
        x = t.empty(nb * d * h * w * IN_CH, dtype=t.bfloat16, device=self.dev)
        _chk(nn_lib().fc_nn_features(_p(solid), _p(sdf), u_in, nb, d, h, w, _p(x), self._stream()), "features")
        return x

    # ------------------------------------------------------------ forward
    def forward(self, x, nb: int, dims):
        """x: [nb*D*H*W*8] bf16.  Returns the activations dict (incl. 'x')."""
        t = self.torch
        L = nn_lib()
        acts = {"x": x}
        s = self._stream()
        for name, mode, ins, kin, c1, cout, stride, lvl in LAYERS:
            dd, hd, wd = self.level_dims(dims, lvl)
            y = t.empty(nb * dd * hd * wd * cout, dtype=t.bfloat16, device=self.dev)
            p = self.params[name]
            if mode == FWD:
                src = acts[ins[0]]
                si = self.level_dims(dims, lvl - (1 if stride == 2 else 0))
                _chk(L.fc_nn_conv(FWD, _p(src), None, _p(p["wbf"]), _p(p["b"]), _p(y), None, nb, si[0], si[1],
                                  si[2], dd, hd, wd, kin, 0, cout, stride, 1, s), name)
            else:
                coarse, skip = acts[ins[0]], acts[ins[1]]
                _chk(L.fc_nn_conv(UPCAT, _p(skip), _p(coarse), _p(p["wbf"]), _p(p["b"]), _p(y), None, nb, dd, hd, wd,
                                  dd, hd, wd, kin, c1, cout, 1, 1, s), name)
            acts[name] = y
        return acts

    def head(self, acts, nb, dims, tgt=None, wmask=None, want_grad=False, pred_out=None):
        """Head + masked MSE.  Returns (loss tensor [1] fp32 on device, dh or None)."""
        t = self.torch
        V = nb * dims[0] * dims[1] * dims[2]
        if self._ws is None or self._ws.numel() < 512 * (4 * HEAD_IN + 5):
            self._ws = t.empty(512 * (4 * HEAD_IN + 5), dtype=t.float32, device=self.dev)
        dh = t.empty(V * HEAD_IN, dtype=t.bfloat16, device=self.dev) if want_grad else None
        inv_norm = 1.0 / (4.0 * V)
        _chk(nn_lib().fc_nn_head_loss(_p(acts["d0"]), HEAD_IN, _p(self.head_w), _p(self.head_b), _p(tgt), _p(wmask),
                                      V, inv_norm, _p(dh), _p(self._ws), _p(self.head_gw), _p(pred_out),
                                      self._stream()), "head")
        return self.head_gw[4 * HEAD_IN + 4:], dh

    def predict(self, x, nb, dims):
        """Forward + head: fp32 prediction [V][4] in target units."""
        t = self.torch
        acts = self.forward(x, nb, dims)
        V = nb * dims[0] * dims[1] * dims[2]
        pred = t.empty(V * 4, dtype=t.float32, device=self.dev)
        self.head(acts, nb, dims, pred_out=pred)
        return pred

    # ----------------------------------------------------------- backward
    def backward(self, acts, dh, nb, dims):
        """Accumulates parameter gradients (gw, gb overwritten) from dL/d(d0) = dh."""
        t = self.torch
        L = nn_lib()
        s = self._stream()
        grads = {"d0": dh}
        for name, mode, ins, kin, c1, cout, stride, lvl in reversed(LAYERS):
            p = self.params[name]
            dd, hd, wd = self.level_dims(dims, lvl)
            V = nb * dd * hd * wd
            y = acts[name]
            dz = t.empty_like(y)
            # dZ and the bias gradient (its per-channel sums) in one pass
            cws = t.empty(L.fc_nn_col_blocks() * cout, dtype=t.float32, device=self.dev)
            fused = L.fc_nn_leaky_backward_colsum(_p(y), _p(grads[name]), _p(dz), V, cout, _p(p["gb"]), _p(cws), 0,
                                                   s) == 0
            if not fused:
                _chk(L.fc_nn_leaky_backward(_p(y), _p(grads[name]), _p(dz), V * cout, s), "leaky")
            # weight and bias gradients
            K = 27 * kin
            splits = wgrad_splits(V, K, cout)
            ws = t.empty(splits * cout * K, dtype=t.float32, device=self.dev)
            if mode == FWD:
                si = self.level_dims(dims, lvl - (1 if stride == 2 else 0))
                _chk(L.fc_nn_wgrad(FWD, _p(acts[ins[0]]), None, _p(dz), _p(p["gw"]), _p(ws), splits, 0, nb, si[0],
                                   si[1], si[2], dd, hd, wd, kin, 0, cout, stride, s), name + " wgrad")
            else:
                _chk(L.fc_nn_wgrad(UPCAT, _p(acts[ins[1]]), _p(acts[ins[0]]), _p(dz), _p(p["gw"]), _p(ws), splits, 0,
                                   nb, dd, hd, wd, dd, hd, wd, kin, c1, cout, 1, s), name + " wgrad")
            if not fused:
                _chk(L.fc_nn_colsum(_p(dz), V, cout, _p(p["gb"]), _p(cws), 0, s), name + " bias")
            if name == "e0":
                continue  # no gradient needed for the input features
            # data gradient through the transposed weights: the stride-2 adjoint gathers with
            # them; a stride-1 adjoint is the forward convolution of dZ with the transposed,
            # mirrored weights, and runs on the forward kernels
            wt = t.empty_like(p["wbf"])
            if stride == 2:
                _chk(L.fc_nn_transpose_w(_p(p["wbf"]), _p(wt), cout, kin, s), "transpose")
            else:
                _chk(L.fc_nn_flip_transpose_w(_p(p["wbf"]), _p(wt), cout, kin, s), "transpose")
            if mode == FWD:
                src = ins[0]
                si = self.level_dims(dims, lvl - (1 if stride == 2 else 0))
                dx = t.empty(nb * si[0] * si[1] * si[2] * kin, dtype=t.bfloat16, device=self.dev)
                if stride == 2:
                    _chk(L.fc_nn_conv(ADJ, _p(dz), None, _p(wt), None, _p(dx), None, nb, dd, hd, wd, si[0], si[1],
                                      si[2], cout, 0, kin, stride, 0, s), name + " dgrad")
                else:
                    _chk(L.fc_nn_conv(FWD, _p(dz), None, _p(wt), None, _p(dx), None, nb, dd, hd, wd, dd, hd, wd,
                                      cout, 0, kin, 1, 0, s), name + " dgrad")
                if src in grads:
                    _chk(L.fc_nn_add_bf16(_p(dx), _p(grads[src]), dx.numel(), s), "add")
                else:
                    grads[src] = dx
            else:
                coarse, skip = ins
                dcat = t.empty(V * kin, dtype=t.bfloat16, device=self.dev)
                _chk(L.fc_nn_conv(FWD, _p(dz), None, _p(wt), None, _p(dcat), None, nb, dd, hd, wd, dd, hd, wd,
                                  cout, 0, kin, 1, 0, s), name + " dgrad")
                if skip not in grads:
                    grads[skip] = t.zeros(V * (kin - c1), dtype=t.bfloat16, device=self.dev)
                dcoarse = t.empty(V // 8 * c1, dtype=t.bfloat16, device=self.dev)
                _chk(L.fc_nn_upcat_split(_p(dcat), c1, kin - c1, nb, dd, hd, wd, _p(grads[skip]), _p(dcoarse), s),
                     name + " split")
                if coarse in grads:
                    _chk(L.fc_nn_add_bf16(_p(dcoarse), _p(grads[coarse]), dcoarse.numel(), s), "add")
                else:
                    grads[coarse] = dcoarse
        return grads

    # --------------------------------------------------------------- adam
    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        L = nn_lib()
        s = self._stream()
        self.step_count += 1
        k = self.step_count
        for name, p in self.params.items():
            _chk(L.fc_nn_adam(_p(p["w"]), _p(p["gw"]), _p(p["mw"]), _p(p["vw"]), _p(p["wbf"]), p["w"].numel(), lr, b1,
                              b2, eps, k, 1.0, s), "adam")
            _chk(L.fc_nn_adam(_p(p["b"]), _p(p["gb"]), _p(p["mb"]), _p(p["vb"]), None, p["b"].numel(), lr, b1, b2,
                              eps, k, 1.0, s), "adam")
        gw = self.head_gw[: 4 * HEAD_IN]
        gb = self.head_gw[4 * HEAD_IN: 4 * HEAD_IN + 4]
        m, v, mb, vb = self.head_state
        _chk(L.fc_nn_adam(_p(self.head_w), _p(gw), _p(m), _p(v), None, self.head_w.numel(), lr, b1, b2, eps, k, 1.0,
                          s), "adam")
        _chk(L.fc_nn_adam(_p(self.head_b), _p(gb), _p(mb), _p(vb), None, 4, lr, b1, b2, eps, k, 1.0, s), "adam")

    def train_step(self, x, tgt, wmask, nb, dims, lr=1e-3):
        """One Adam step on a batch; returns the loss as a device tensor."""
        acts = self.forward(x, nb, dims)
        loss, dh = self.head(acts, nb, dims, tgt=tgt, wmask=wmask, want_grad=True)
        loss = loss.clone()
        self.backward(acts, dh, nb, dims)
        self.adam_step(lr)
        return loss


def targets_from_field(field_dev, solid_dev, u_in, V):
    """LBM field (float4 per cell, device) -> (targets fp32 [V][4], fluid weights [V])."""
    import torch
    tgt = torch.empty(V * 4, dtype=torch.float32, device=field_dev.device)
    wm = torch.empty(V, dtype=torch.float32, device=field_dev.device)
    _chk(nn_lib().fc_nn_targets(_p(field_dev), _p(solid_dev), u_in, V, _p(tgt), _p(wm),
                                ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)), "targets")
    return tgt, wm


def field_from_prediction(pred, u_in, V):
    """Surrogate prediction [V][4] -> LBM seed field (rho, ux, uy, uz) on the device."""
    import torch
    field = torch.empty(V * 4, dtype=torch.float32, device=pred.device)
    _chk(nn_lib().fc_nn_pred_to_field(_p(pred), u_in, V, _p(field),
                                      ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)), "pred_to_field")
    return field


def time_layers(nb=4, dims=(128, 128, 256), reps=10, train_reps=5, log=print):
    """Times each forward convolution of the network (HIP events over `reps` launches),
    the forward pass (the convolutions back to back, `reps` passes) and a whole training
    step on random inputs of nb samples of dims = (D, H, W); returns {"layers": [(name,
    voxels, kin, cout, ms, TFLOP/s)], "forward_ms", "train_step_ms", "fwd_flops"}."""
    import torch
    L = nn_lib()
    net = Surrogate(seed=0)
    V0 = nb * dims[0] * dims[1] * dims[2]
    x = torch.randn(V0 * IN_CH, device=net.dev).to(torch.bfloat16)
    acts = net.forward(x, nb, dims)
    torch.cuda.synchronize()
    s = net._stream()
    rows, tot_f = [], 0.0
    for name, mode, ins, kin, c1, cout, stride, lvl in LAYERS:
        dd, hd, wd = net.level_dims(dims, lvl)
        V = nb * dd * hd * wd
        y = torch.empty(V * cout, dtype=torch.bfloat16, device=net.dev)
        p = net.params[name]
        if mode == FWD:
            si = net.level_dims(dims, lvl - (1 if stride == 2 else 0))
            args = (FWD, _p(acts[ins[0]]), None, _p(p["wbf"]), _p(p["b"]), _p(y), None, nb, si[0], si[1], si[2], dd,
                    hd, wd, kin, 0, cout, stride, 1, s)
        else:
            args = (UPCAT, _p(acts[ins[1]]), _p(acts[ins[0]]), _p(p["wbf"]), _p(p["b"]), _p(y), None, nb, dd, hd, wd,
                    dd, hd, wd, kin, c1, cout, 1, 1, s)
        _chk(L.fc_nn_conv(*args), name)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            L.fc_nn_conv(*args)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        flops = 2.0 * V * 27 * kin * cout
        tot_f += flops
        rows.append((name, V, kin, cout, ms, flops / ms / 1e9))
        log(f"{name:4s} V={V:9d} K={27 * kin:5d} N={cout:4d}  {ms:8.3f} ms  {flops / ms / 1e9:8.1f} TFLOP/s")
    # the forward pass as the network runs it: the 11 convolutions back to back
    for _ in range(2):
        net.forward(x, nb, dims)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        net.forward(x, nb, dims)
    e1.record()
    torch.cuda.synchronize()
    fwd_ms = e0.elapsed_time(e1) / reps
    log(f"forward pass (11 convolutions back to back): {fwd_ms:.3f} ms  {tot_f / fwd_ms / 1e9:8.1f} TFLOP/s; "
        f"layers timed one by one: {sum(r[4] for r in rows):.3f} ms")
    if train_reps <= 0:
        return {"layers": rows, "train_step_ms": None, "fwd_flops": tot_f, "forward_ms": fwd_ms}
    tgt = torch.zeros(V0 * 4, device=net.dev)
    wm = torch.ones(V0, device=net.dev)
    for _ in range(2):
        net.train_step(x, tgt, wm, nb, dims)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(train_reps):
        net.train_step(x, tgt, wm, nb, dims)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / train_reps
    log(f"train step (batch {nb}): {ms:.2f} ms; 3 x forward FLOPs -> {3 * tot_f / ms / 1e9:.1f} TFLOP/s")
    return {"layers": rows, "train_step_ms": ms, "fwd_flops": tot_f, "forward_ms": fwd_ms}


if __name__ == "__main__":
    import argparse
    import json
    ap = argparse.ArgumentParser(prog="python -m flowcraft_amd.surrogate",
                                 description="Time the surrogate's convolution kernels and a training step.")
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--dims", type=int, nargs=3, default=[128, 128, 256], metavar=("D", "H", "W"))
    ap.add_argument("--reps", type=int, default=10, help="launches per layer timing")
    ap.add_argument("--train-reps", type=int, default=5, help="training steps timed (0: layers only)")
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    import torch
    torch.cuda.init()
    r = time_layers(a.batch, tuple(a.dims), reps=a.reps, train_reps=a.train_reps)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(r, f)
