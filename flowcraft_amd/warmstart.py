"""Warm-start pipeline: training data from the engine itself, surrogate training,
and the warm-vs-plain convergence study.

Dataset: random obstacles (half NACA 4-digit wings extruded over z, half
ellipsoids; ``geometry.random_shape_grid``) on a D3Q19 Smagorinsky lattice,
each run from the uniform rest start for a fixed number of steps; the target is
the final (rho, u) field in normalised units ((rho-1)*3/u^2, u/u_in).

Convergence study: for each held-out shape, four runs of ``horizon`` steps — FlowCraft's
two uniform seeds (at rest, and at half the inlet speed, its default: LbmSolver::Reset,
FlowCraft/main.cpp:431), the surrogate's field, and FlowCraft's own coarse-grid warm start
(ApplyWarmStart: the model at 1/``coarse_div`` resolution, ``coarse_steps`` coarse steps,
nearest-sampled onto the fine grid; FlowCraft/main.cpp:1413-1564).  Cd(t) is the
momentum-exchange drag Fx / (0.5 U^2 A) averaged over the trailing ``window`` steps
(sampled every ``every`` steps).  Each plain run's value at the horizon is a reference;
every run's steps to convergence to it are counted under three rules (first entry into the
``tol`` band, staying in it for ``hold`` steps, staying to the end).  A seeded run that
diverges or never converges falls back to the plain start (counted, charged its time).
"""
from __future__ import annotations

import math
import os
import time

import numpy as np

from . import geometry
from .lbm import LatticeModel, Layout, LbmSolver, Params, Storage, VoxelGrid
from . import surrogate as sg


# Surrogate weights trained on the 256x128x128 warm-start distribution (600 engine-generated
# fields, 3000 steps each, targets the mean of the last 1000 steps; `train` below), for
# the study's 6000-step horizon.
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "surrogate_256x128x128.npz")
TRAIN_STEPS = 3000
HORIZON = 6000
# On this channel the plain start's trailing-mean drag still drifts 2-7% per 1000 steps at
# step 6000 and settles (mostly under 1% per 1000 steps) between 12000 and 18000 steps.
# The settled variant of the study: weights trained on 320 fields of SETTLED_TRAIN_STEPS
# steps (mean of the last 1000), reference = the plain run's value at SETTLED_HORIZON.
SETTLED_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights",
                               "surrogate_256x128x128_settled.npz")
SETTLED_TRAIN_STEPS = 15000
SETTLED_HORIZON = 24000


def make_solver(grid, storage=Storage.F16, layout=Layout.PINGPONG):
    s = LbmSolver()
    if not s.Init(None, grid, LatticeModel.D3Q19, storage=storage, layout=layout):
        raise RuntimeError(s.LastError())
    return s


def shape_stream(seed):
    return np.random.Generator(np.random.PCG64(seed))


def run_to_mean_field(s, steps, avg_window, avg_every, V):
    """Steps `steps` times from the solver's current state; returns the (rho, u) field
    averaged over the last `avg_window` steps (sampled every `avg_every` steps; the
    final field when avg_window is 0) as a device tensor."""
    import torch
    field = torch.zeros(V * 4, dtype=torch.float32, device="cuda")
    tmp = torch.empty(V * 4, dtype=torch.float32, device="cuda")
    torch.cuda.current_stream().synchronize()
    start = steps - avg_window if avg_window > 0 else steps
    s.Step(None, start)
    n = 0
    if avg_window > 0:
        for _ in range(avg_window // avg_every):
            s.Step(None, avg_every)
            s.FieldDevice(tmp)
            s.Sync()
            field += tmp
            n += 1
        field /= n
    else:
        s.FieldDevice(field)
        s.Sync()
    return field


def generate_dataset(n, dims, steps, u_in, tau, cs, seed, storage=Storage.F16, log=print, time_budget_s=None,
                     avg_window=1000, avg_every=50):
    """Runs n random shapes from the uniform rest start; returns a list of device samples
    (solid, sdf, tgt, wmask) and their descriptions.  dims = (nx, ny, nz).  The target is
    the (rho, u) field averaged over the last `avg_window` of the `steps` steps: the
    acoustic waves the impulsive start leaves in the domain average out, the part of the
    field the geometry determines remains."""
    import torch
    nx, ny, nz = dims
    rng = shape_stream(seed)
    params = Params(tau=tau, inletVelocity=u_in, smagorinskyCs=cs)
    out, desc = [], []
    V = nx * ny * nz
    t0 = time.time()
    for i in range(n):
        g, d = geometry.random_shape_grid(nx, ny, nz, rng)
        s = make_solver(g, storage)
        s.Reset(params, False)
        field = run_to_mean_field(s, steps, avg_window, avg_every, V)
        solid = torch.from_numpy(np.ascontiguousarray(g.solid.reshape(-1))).cuda()
        sdf = torch.from_numpy(np.ascontiguousarray(g.signedDistance.reshape(-1))).cuda()
        tgt, wm = sg.targets_from_field(field, solid, u_in, V)
        st = s.ReadbackStats()
        s.Shutdown()
        if st.hasNaN:
            log(f"  shape {i} diverged; skipped")
            continue
        out.append((solid, sdf, tgt, wm))
        desc.append(d)
        if (i + 1) % 50 == 0 or i == n - 1:
            log(f"  dataset {i + 1}/{n} shapes, {time.time() - t0:.1f}s")
        if time_budget_s is not None and time.time() - t0 > time_budget_s:
            log(f"  dataset time budget reached at {i + 1} shapes")
            break
    return out, desc


def batch_tensors(samples, idx, u_in, net, dims):
    """Concatenates the samples idx into one batch: features (bf16), targets, weights."""
    import torch
    solid = torch.cat([samples[i][0] for i in idx])
    sdf = torch.cat([samples[i][1] for i in idx])
    tgt = torch.cat([samples[i][2] for i in idx])
    wm = torch.cat([samples[i][3] for i in idx])
    nx, ny, nz = dims
    x = net.features(solid, sdf, u_in, len(idx), (nz, ny, nx))
    return x, tgt, wm


def train(net, samples, dims, u_in, epochs=20, batch=4, lr=1e-3, seed=0, time_budget_s=None, log=print):
    """Adam on shuffled batches; returns the per-epoch mean loss history."""
    import torch
    nx, ny, nz = dims
    rng = np.random.Generator(np.random.PCG64(seed))
    hist = []
    t0 = time.time()
    t_log = t0
    for ep in range(epochs):
        order = rng.permutation(len(samples))
        losses = []
        for b in range(0, len(order) - batch + 1, batch):
            idx = order[b:b + batch]
            x, tgt, wm = batch_tensors(samples, idx, u_in, net, dims)
            losses.append(net.train_step(x, tgt, wm, len(idx), (nz, ny, nx), lr=lr))
            if time.time() - t_log > 45.0:
                torch.cuda.synchronize()
                t_log = time.time()
                log(f"  epoch {ep + 1}: {b + batch}/{len(order)} samples ({t_log - t0:.1f}s)")
            if time_budget_s is not None and time.time() - t0 > time_budget_s:
                break
        torch.cuda.synchronize()
        mean = float(torch.stack(losses).mean().cpu())
        hist.append(mean)
        log(f"  epoch {ep + 1}: loss {mean:.5f}  ({time.time() - t0:.1f}s)")
        if time_budget_s is not None and time.time() - t0 > time_budget_s:
            log("  training time budget reached")
            break
    return hist


def field_l2(net, samples, dims, u_in, idx):
    """Relative L2 error of the surrogate's normalised field over fluid cells."""
    import torch
    nx, ny, nz = dims
    errs = []
    for i in idx:
        x, tgt, wm = batch_tensors(samples, [i], u_in, net, dims)
        pred = net.predict(x, 1, (nz, ny, nx))
        w = wm.repeat_interleave(4)
        e = ((pred - tgt) ** 2 * w).sum() / ((tgt ** 2 * w).sum() + 1e-30)
        errs.append(float(torch.sqrt(e).cpu()))
    return errs


def frontal_area(desc, dims):
    return float(desc["area"])


def run_cd_series(solver, steps, every, area, u_in):
    """Steps the solver, sampling the FP64 force every `every` steps on the device;
    returns (Cd samples, Cl samples, ms per step)."""
    import torch
    n = steps // every
    buf = torch.empty((n, 4), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(n):
        solver.Step(None, every)
        solver.ForcesDevice(buf[k])
    solver.Sync()
    dt = time.perf_counter() - t0
    f = buf.cpu().numpy()
    q = 0.5 * u_in * u_in * area
    return f[:, 0] / q, f[:, 1] / q, dt * 1e3 / steps


def trailing(x, w):
    c = np.cumsum(np.concatenate([[0.0], x]))
    out = np.empty_like(x)
    for i in range(len(x)):
        lo = max(0, i + 1 - w)
        out[i] = (c[i + 1] - c[lo]) / (i + 1 - lo)
    return out


def converged_index(cd_avg, ref, tol, hold=None):
    """First index from which |cd_avg - ref| <= tol*|ref| holds for the next `hold`
    samples (to the end of the series when hold is None, or when fewer remain), or None."""
    ok = np.abs(cd_avg - ref) <= tol * abs(ref)
    n = len(ok)
    if hold is None:
        if not ok[-1]:
            return None
        bad = np.nonzero(~ok)[0]
        return 0 if bad.size == 0 else int(bad[-1]) + 1
    # run length of consecutive in-band samples starting at each index
    run = np.zeros(n + 1, int)
    for i in range(n - 1, -1, -1):
        run[i] = run[i + 1] + 1 if ok[i] else 0
    for i in range(n):
        if ok[i] and (run[i] >= hold or i + run[i] == n):
            return i
    return None


def flowcraft_coarse_grid(desc, dims, div=10):
    """The coarse grid of FlowCraft's ApplyWarmStart (main.cpp:1496-1510) for a shape of
    the study: the same model voxelized at 1/div of the fine grid's resolution over the
    same padded domain.  Frame as Voxelizer::Voxelize builds it (Voxelizer.cpp:136-170):
    the fine grid's origin, cell size = longest domain axis / max(resolution/div, 4),
    nx = ceil(world-Y extent / cell), ny = ceil(world-X extent / cell), nz likewise; the
    fine grid here is the study's lattice with cell size 1 at the origin.  Classification
    by the voxelizer's rule: every cell whose centre lies within sqrt(3)/2 cells of the
    surface is solid (the conservative shell, on both sides of it), so is everything
    enclosed; the signed distance (world units) is kept for the Bouzidi links.  The
    surface is the study's analytic one (geometry.shape_sdf_at), not a triangulation of
    it."""
    nx, ny, nz = dims
    f32 = np.float32
    res = max(nx, ny, nz)
    res_c = max(res // max(div, 1), 4)
    cs = max(f32(res) / f32(res_c), f32(1e-6))
    nxc = max(int(math.ceil(f32(nx) / cs)), 1)  # grid-x spans world Y
    nyc = max(int(math.ceil(f32(ny) / cs)), 1)
    nzc = max(int(math.ceil(f32(nz) / cs)), 1)
    z, y, x = np.meshgrid((np.arange(nzc, dtype=f32) + f32(0.5)) * cs, (np.arange(nyc, dtype=f32) + f32(0.5)) * cs,
                          (np.arange(nxc, dtype=f32) + f32(0.5)) * cs, indexing="ij")
    d = geometry.shape_sdf_at(desc, dims, x, y, z).astype(f32)
    solid = (d <= cs * f32(0.8660254)).astype(np.uint8)
    return VoxelGrid(nxc, nyc, nzc, solid, d, cellSize=float(cs), origin=(0.0, 0.0, 0.0))


def sample_coarse_field(fine: VoxelGrid, coarse: VoxelGrid, coarse_field) -> np.ndarray:
    """FlowCraft's SampleCoarseFieldOntoFineGrid (main.cpp:1413-1461): each fine cell takes
    the field of the coarse cell that contains its centre in the world frame — world centre
    by VoxelGrid::CellCenter (grid-y along world X, grid-x along world Y, VoxelGrid.h:53-55),
    coarse indices by the inverse mapping with the same swap, each clamped to the grid;
    a coarse solid cell gives quiescent fluid (1, 0, 0, 0).  Float32 arithmetic in the
    reference's order.  Returns 4 floats per fine cell."""
    f32 = np.float32
    cs_f, cs_c = f32(fine.cellSize), f32(coarse.cellSize)
    of, oc = np.asarray(fine.origin, f32), np.asarray(coarse.origin, f32)
    inv = f32(1.0) / cs_c
    gx = (np.arange(fine.nx, dtype=f32) + f32(0.5)) * cs_f  # world Y of grid-x
    gy = (np.arange(fine.ny, dtype=f32) + f32(0.5)) * cs_f  # world X of grid-y
    gz = (np.arange(fine.nz, dtype=f32) + f32(0.5)) * cs_f
    # local = (worldCenter - coarse.origin) * invCoarseCellSize, per world axis
    lx = ((of[0] + gy) - oc[0]) * inv   # world X: from the fine grid-y index
    ly = ((of[1] + gx) - oc[1]) * inv   # world Y: from the fine grid-x index
    lz = ((of[2] + gz) - oc[2]) * inv
    cx = np.clip(np.floor(ly).astype(np.int64), 0, coarse.nx - 1)  # coarse grid-x from world Y
    cy = np.clip(np.floor(lx).astype(np.int64), 0, coarse.ny - 1)  # coarse grid-y from world X
    cz = np.clip(np.floor(lz).astype(np.int64), 0, coarse.nz - 1)
    cell = (cz[:, None, None] * coarse.ny + cy[None, :, None]) * coarse.nx + cx[None, None, :]
    cf = np.asarray(coarse_field, f32).reshape(-1, 4)
    seed = cf[cell.reshape(-1)].copy()
    seed[np.asarray(coarse.solid).reshape(-1)[cell.reshape(-1)] != 0] = np.array([1.0, 0.0, 0.0, 0.0], f32)
    return seed.reshape(-1)


def flowcraft_warm_seed(desc, dims, params, div=10, steps=1000, seed_mid=True):
    """FlowCraft's ApplyWarmStart (main.cpp:1496-1564) on a shape of the study: the coarse
    grid (flowcraft_coarse_grid), a coarse solver seeded as FlowCraft's Reset seeds
    (mid-speed by default, AppState::lbmSeedMidSpeed, main.cpp:431) and stepped `steps`
    times (warmStartSteps, main.cpp:444), its field read back and nearest-sampled onto the
    fine grid (sample_coarse_field).  Returns the fine grid's seed field."""
    nx, ny, nz = dims
    cg = flowcraft_coarse_grid(desc, dims, div)
    s = make_solver(cg, Storage.F32)
    s.Reset(params, seed_mid)
    s.Step(None, steps)
    cf = s.ReadbackField()
    s.Shutdown()
    fine = VoxelGrid(nx, ny, nz, np.zeros(0, np.uint8), None, cellSize=1.0, origin=(0.0, 0.0, 0.0))
    return sample_coarse_field(fine, cg, cf)


RULES = ("reach", "hold", "horizon")


def _arm_steps(cd, ref, tol, w, every, hold):
    """Steps to convergence of one arm under three rules, with the band
    |Cd - ref| <= tol*|ref| on the trailing-mean Cd: 'horizon' — it stays in the band to
    the end of the run (the literal rule); 'hold' — it enters the band and stays for
    `hold` steps (or to the end of the run); 'reach' — it first enters the band."""
    if not np.all(np.isfinite(cd)):
        return {r: None for r in RULES}, None
    cda = trailing(cd, w)
    out = {}
    for rule in RULES:
        h = {"reach": 1, "hold": max(1, hold // every), "horizon": None}[rule]
        k = converged_index(cda, ref, tol, hold=h)
        out[rule] = None if k is None else (k + 1) * every
    return out, cda


# The study's arms: two plain starts (FlowCraft's uniform seeds: at rest, and at half the
# inlet speed with a Bernoulli density, its default — LbmSolver::Reset, main.cpp:431), the
# learned warm start and FlowCraft's coarse-grid warm start.  The plain arms are the
# baselines: each seeded arm is scored against each of them.
PLAIN_ARMS = ("plain", "plain_mid")
SEEDED_ARMS = ("warm", "coarse")


def evaluate(net, dims, n_shapes, seed, horizon, every, window, tol, u_in, tau, cs, coarse_div=10, coarse_steps=1000,
             storage=Storage.F16_SR, log=print, time_budget_s=None, hold=1000, series_every=50):
    """The warm-vs-plain study on n_shapes held-out shapes.  Each shape runs `horizon`
    steps from each arm's start.  For each plain arm, its trailing-mean Cd at the horizon is
    a reference value, and every arm's steps to convergence to that value are found under
    the three rules (reach, hold, horizon).  Returns per-shape records."""
    import torch
    nx, ny, nz = dims
    V = nx * ny * nz
    rng = shape_stream(seed)
    params = Params(tau=tau, inletVelocity=u_in, smagorinskyCs=cs)
    w = max(1, window // every)
    ds = max(1, series_every // every)
    recs = []
    t_start = time.time()
    for i in range(n_shapes):
        g, d = geometry.random_shape_grid(nx, ny, nz, rng)
        area = frontal_area(d, dims)
        rec = {"kind": d["kind"], "area": area,
               "shape": {k: (list(v) if isinstance(v, tuple) else v) for k, v in d.items() if k not in ("kind", "area")}}
        s = make_solver(g, storage)
        raw, cl_final = {}, {}
        for arm in PLAIN_ARMS:
            s.Reset(params, arm == "plain_mid")
            cd, cl, ms = run_cd_series(s, horizon, every, area, u_in)
            raw[arm] = cd
            cl_final[arm] = float(trailing(cl, w)[-1]) if np.all(np.isfinite(cl)) else None
            if arm == "plain":
                rec["ms_per_step"] = ms
        # learned arm: the surrogate's field, inference timed
        solid = torch.from_numpy(np.ascontiguousarray(g.solid.reshape(-1))).cuda()
        sdf = torch.from_numpy(np.ascontiguousarray(g.signedDistance.reshape(-1))).cuda()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x = net.features(solid, sdf, u_in, 1, (nz, ny, nx))
        pred = net.predict(x, 1, (nz, ny, nx))
        seed_field = sg.field_from_prediction(pred, u_in, V)
        torch.cuda.synchronize()
        rec["infer_ms"] = (time.perf_counter() - t0) * 1e3
        s.SetParams(params)
        s.ResetFromField(seed_field)
        raw["warm"], clw, _ = run_cd_series(s, horizon, every, area, u_in)
        cl_final["warm"] = float(trailing(clw, w)[-1]) if np.all(np.isfinite(clw)) else None
        # FlowCraft's coarse-grid warm start, its whole cost timed
        t0 = time.perf_counter()
        cseed = flowcraft_warm_seed(d, dims, params, coarse_div, coarse_steps)
        rec["coarse_ms"] = (time.perf_counter() - t0) * 1e3
        s.SetParams(params)
        s.ResetFromField(cseed)
        raw["coarse"], clc, _ = run_cd_series(s, horizon, every, area, u_in)
        cl_final["coarse"] = float(trailing(clc, w)[-1]) if np.all(np.isfinite(clc)) else None
        s.Shutdown()
        series = {arm: (trailing(cd, w) if np.all(np.isfinite(cd)) else None) for arm, cd in raw.items()}
        rec["cd_final"] = {arm: (None if v is None else float(v[-1])) for arm, v in series.items()}
        rec["cl_final"] = cl_final
        rec["series"] = {arm: (None if v is None else [float(c) for c in v[ds - 1::ds]]) for arm, v in series.items()}
        rec["steps"] = {}
        for base in PLAIN_ARMS:
            ref = rec["cd_final"][base]
            rec["steps"][base] = {arm: (_arm_steps(raw[arm], ref, tol, w, every, hold)[0] if ref is not None
                                        else {r: None for r in RULES}) for arm in raw}
        recs.append(rec)
        st = rec["steps"]
        log(f"  eval {i + 1}/{n_shapes} {d['kind']}: Cd rest {rec['cd_final']['plain']:.4f} mid "
            f"{rec['cd_final']['plain_mid']:.4f} warm {rec['cd_final']['warm']} coarse {rec['cd_final']['coarse']}; "
            + "; ".join(f"vs {b}: " + " ".join(f"{r} " + "/".join(str(st[b][a][r]) for a in ("plain", "plain_mid",
                                                                                          "warm", "coarse"))
                                              for r in RULES) for b in PLAIN_ARMS))
        if time_budget_s is not None and time.time() - t_start > time_budget_s:
            log(f"  evaluation time budget reached after {i + 1} shapes")
            break
    return recs


def _stats(v):
    v = np.asarray(v, float)
    return {"mean": float(v.mean()), "p50": float(np.percentile(v, 50)), "p95": float(np.percentile(v, 95))}


def summarize(recs, horizon):
    """Per plain baseline and rule: steps and wall time to convergence (mean, p50, p95) of
    the baseline and of each seeded arm, the wall-time speedups, and the seeded runs that
    fell back to the plain start (no convergence within the horizon: charged the
    baseline's time plus the seeding cost).  Also the spread of the final (trailing-mean)
    Cd between arms: warm, coarse and the other plain start, each against each plain arm."""
    cost = {"warm": "infer_ms", "coarse": "coarse_ms"}
    out = {"shapes": len(recs), "horizon_steps": horizon}
    for base in PLAIN_ARMS:
        ob = {}
        for rule in RULES:
            o = {"plain_unconverged": 0}
            base_s, base_t = [], []
            arm_s = {a: [] for a in SEEDED_ARMS}
            arm_t = {a: [] for a in SEEDED_ARMS}
            fb = {a: 0 for a in SEEDED_ARMS}
            for r in recs:
                st = r["steps"][base]
                ps = st[base][rule]
                if ps is None:
                    o["plain_unconverged"] += 1
                    continue
                ms = r["ms_per_step"]
                base_s.append(ps)
                base_t.append(ps * ms)
                for a in SEEDED_ARMS:
                    sa = st[a][rule]
                    if sa is None:
                        fb[a] += 1
                        sa = ps
                    arm_s[a].append(sa)
                    arm_t[a].append(sa * ms + r.get(cost[a], 0.0))
            if base_s:
                o.update(plain_steps=_stats(base_s), plain_ms=_stats(base_t))
                for a in SEEDED_ARMS:
                    o[a] = {"steps": _stats(arm_s[a]), "ms": _stats(arm_t[a]), "fallbacks": fb[a],
                            "speedup_wall": _stats(np.asarray(base_t) / np.asarray(arm_t[a])),
                            "speedup_wall_of_means": float(np.mean(base_t) / np.mean(arm_t[a]))}
            ob[rule] = o
        out[base] = ob
    spread = {}
    for base in PLAIN_ARMS:
        for a in ("plain", "plain_mid") + SEEDED_ARMS:
            if a == base:
                continue
            v = [abs(r["cd_final"][a] - r["cd_final"][base]) / abs(r["cd_final"][base]) for r in recs
                 if r["cd_final"][a] is not None and r["cd_final"][base] not in (None, 0.0)]
            if v:
                spread[f"{a}_vs_{base}"] = _stats(v)
    out["final_cd_rel_diff"] = spread
    return out


# ------------------------------------------------------------------- driver
def _logger(path):
    f = open(path, "a") if path else None

    def log(msg):
        line = f"[{time.strftime('%H:%M:%S')}] {msg}"
        print(line, flush=True)
        if f is not None:
            f.write(line + "\n")
            f.flush()
    return log


def save_weights(net, path):
    """Trained weights for inference: conv weights as their bf16 bit patterns (what the
    kernels consume), biases and the fp32 head as float32."""
    sd = net.state_dict()
    out = {}
    for k, v in sd.items():
        if k.endswith(".w") and not k.startswith("head"):
            b = np.ascontiguousarray(v, np.float32).view(np.uint32).astype(np.uint64)
            out[k + ".bf16"] = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16).astype(np.uint16)
        else:
            out[k] = np.asarray(v, np.float32)
    np.savez(path, **out)


def load_weights(net, path):
    sd = {}
    with np.load(path, allow_pickle=False) as z:
        for k in z.files:
            if k.endswith(".bf16"):
                sd[k[:-5]] = (z[k].astype(np.uint32) << 16).view(np.float32)
            else:
                sd[k] = z[k]
    net.load_state_dict(sd)


def cmd_train(args):
    import json
    import os
    import torch
    os.makedirs(args.out, exist_ok=True)
    log = _logger(os.path.join(args.out, "train.log"))
    dims = tuple(args.dims)
    torch.cuda.init()
    log(f"dataset: {args.shapes} shapes on {dims}, {args.steps} steps, u={args.u} tau={args.tau} Cs={args.cs}")
    t0 = time.time()
    samples, desc = generate_dataset(args.shapes, dims, args.steps, args.u, args.tau, args.cs, args.seed,
                                     log=log, time_budget_s=args.data_budget, avg_window=args.target_window)
    t_data = time.time() - t0
    log(f"dataset done: {len(samples)} shapes in {t_data:.1f}s")
    n_val = max(1, len(samples) // 20)
    train_set, val_set = samples[:-n_val], samples[-n_val:]
    net = sg.Surrogate(seed=args.seed)
    log(f"surrogate: {net.num_parameters()} parameters; training on {len(train_set)}, validating on {len(val_set)}")
    val0 = field_l2(net, val_set, dims, args.u, range(len(val_set)))
    t1 = time.time()
    hist = train(net, train_set, dims, args.u, epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.seed,
                 time_budget_s=args.train_budget, log=log)
    torch.cuda.synchronize()
    t_train = time.time() - t1
    val = field_l2(net, val_set, dims, args.u, range(len(val_set)))
    save_weights(net, os.path.join(args.out, "surrogate.npz"))
    rep = {"dims": dims, "shapes": len(samples), "lbm_steps": args.steps, "u_in": args.u, "tau": args.tau,
           "cs": args.cs, "seed": args.seed, "target_window": args.target_window, "data_s": t_data, "train_s": t_train, "epochs": len(hist),
           "batch": args.batch, "lr": args.lr, "loss_history": hist, "parameters": net.num_parameters(),
           "val_rel_l2_untrained": float(np.mean(val0)), "val_rel_l2": float(np.mean(val)),
           "val_rel_l2_p95": float(np.percentile(val, 95))}
    with open(os.path.join(args.out, "train.json"), "w") as f:
        json.dump(rep, f, indent=1)
    log(f"trained {len(hist)} epochs in {t_train:.1f}s; validation field rel-L2 {rep['val_rel_l2']:.4f} "
        f"(untrained {rep['val_rel_l2_untrained']:.4f})")


def cmd_evaluate(args):
    import json
    import os
    import torch
    os.makedirs(args.out, exist_ok=True)
    log = _logger(os.path.join(args.out, "evaluate.log"))
    dims = tuple(args.dims)
    torch.cuda.init()
    net = sg.Surrogate(seed=0)
    load_weights(net, args.weights)
    storage = {"f16sr": Storage.F16_SR, "f16": Storage.F16, "bf16": Storage.BF16, "f32": Storage.F32}[args.storage]
    log(f"evaluate: {args.shapes} held-out shapes (seed {args.seed}) on {dims}, {args.storage} storage, "
        f"horizon {args.horizon}, Cd every {args.every} averaged over {args.window} steps, tol {args.tol}")
    recs = evaluate(net, dims, args.shapes, args.seed, args.horizon, args.every, args.window, args.tol, args.u,
                    args.tau, args.cs, coarse_div=args.coarse_div, coarse_steps=args.coarse_steps, storage=storage,
                    log=log, time_budget_s=args.budget, hold=args.hold)
    rep = {"dims": dims, "storage": args.storage, "u_in": args.u, "tau": args.tau, "cs": args.cs, "seed": args.seed,
           "horizon": args.horizon, "every": args.every, "window": args.window, "tol": args.tol, "hold": args.hold,
           "weights": os.path.basename(args.weights), "summary": summarize(recs, args.horizon), "records": recs}
    with open(os.path.join(args.out, "evaluate.json"), "w") as f:
        json.dump(rep, f, indent=1)
    log("summary: " + json.dumps(rep["summary"]))


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(prog="python -m flowcraft_amd.warmstart",
                                 description="Learned warm start: train the surrogate on fields this engine "
                                             "generates, then compare warm and plain starts.")
    sub = ap.add_subparsers(dest="cmd", required=True)

    def common(p):
        p.add_argument("--dims", type=int, nargs=3, default=[256, 128, 128], metavar=("NX", "NY", "NZ"))
        p.add_argument("--u", type=float, default=0.06)
        p.add_argument("--tau", type=float, default=0.503)
        p.add_argument("--cs", type=float, default=0.1)
        p.add_argument("--out", required=True)

    tr = sub.add_parser("train", help="generate the dataset and train")
    common(tr)
    tr.add_argument("--shapes", type=int, default=600)
    tr.add_argument("--steps", type=int, default=TRAIN_STEPS)
    tr.add_argument("--seed", type=int, default=1)
    tr.add_argument("--epochs", type=int, default=20)
    tr.add_argument("--batch", type=int, default=4)
    tr.add_argument("--lr", type=float, default=1e-3)
    tr.add_argument("--target-window", type=int, default=1000,
                    help="steps at the end of each run the target field is averaged over (0: last field)")
    tr.add_argument("--data-budget", type=float, default=None, help="seconds")
    tr.add_argument("--train-budget", type=float, default=None, help="seconds")
    tr.set_defaults(fn=cmd_train)

    ev = sub.add_parser("evaluate", help="warm vs plain (and coarse-grid) starts on held-out shapes")
    common(ev)
    ev.add_argument("--weights", required=True)
    ev.add_argument("--shapes", type=int, default=64)
    ev.add_argument("--seed", type=int, default=1001)
    ev.add_argument("--horizon", type=int, default=HORIZON)
    ev.add_argument("--every", type=int, default=10)
    ev.add_argument("--window", type=int, default=1000)
    ev.add_argument("--tol", type=float, default=0.01)
    ev.add_argument("--hold", type=int, default=1000, help="steps a run must stay in the band (hold rule)")
    ev.add_argument("--storage", default="f16sr", choices=["f16sr", "f16", "bf16", "f32"])
    ev.add_argument("--coarse-div", type=int, default=10)
    ev.add_argument("--coarse-steps", type=int, default=1000)
    ev.add_argument("--budget", type=float, default=None, help="seconds")
    ev.set_defaults(fn=cmd_evaluate)
    args = ap.parse_args(argv)
    args.fn(args)


if __name__ == "__main__":
    main()
