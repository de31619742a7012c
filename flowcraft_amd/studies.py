"""Measured studies of the step on one MI355X, beside the throughput bench.

    python -m flowcraft_amd.studies precision [--steps 5000] [--every 50] [--window 1000]
    python -m flowcraft_amd.studies large [--steps 200] [--warmup 50]
    python -m flowcraft_amd.studies capacity [--steps 10] [--warmup 3]
    python -m flowcraft_amd.studies depth [--nz 64 256 1024 2048] [--steps 10]
    python -m flowcraft_amd.studies drain [--steps 20000] [--every 1000]
    python -m flowcraft_amd.studies sensitivity [--t0 1000] [--steps 4000] [--eps 1e-9] [--tau 0.5005]

precision — the 512x256x256 NACA 0012 wing (chord 128, leading edge at x = 128,
    angle of attack 4 deg, extruded over z; D3Q19 and D3Q15 Smagorinsky Cs = 0.1,
    u = 0.08, tau = 0.5005, free-slip y/z walls; --dims scales the wing with the
    lattice), run from the uniform rest start in FP64, FP32, FP16-offset and
    BF16-offset storage (rounded to nearest, or stochastically: f16sr, bf16sr), and the
    FP64 chaos control f64p (the inlet speed raised by one part in a million).  The FP64 run (the GPU
    twin of the FP64 oracle, equal to it within 1e-12 on the parity cases) is the
    reference.  Drag and lift coefficients Cd = Fx / (0.5 u^2 A), Cl = Fy /
    (0.5 u^2 A) with A = chord * nz (lift is along +y: the angle of attack rotates
    the section in the x-y plane), F the FP64 momentum-exchange sum over all solid
    cells of the current populations (LbmReference::SolidCellForce convention,
    FlowCraft/LbmReference.cpp:127-148).  Reported: the instantaneous value at the
    last step and the mean over the trailing `window` steps, and each storage's
    relative error against FP64.

large — the 768x512x576 bandwidth lattice (226M cells): an ellipsoid of semi-axes
    (96, 40, 40) at (192, 256, 288) plus 64 spheres of radius uniform in [4, 12]
    with centres uniform over the downstream half (numpy PCG64, seed 0); D3Q19
    Smagorinsky Cs = 0.1, u = 0.06, tau = 0.501, free-slip y/z; `warmup` untimed
    steps, then `steps` steps timed with HIP events on the solver's stream.  FP32 in
    place (one 17 GB population copy) and FP16 offsets.

capacity — the largest lattices one GPU holds, in place: 2048 x 1024 x nz with nz = 2047
    in stochastically rounded FP16 (4.29e9 cells, the engine's cell limit, 189 GB) and
    nz = 1400 in FP32 (2.94e9 cells, 241 GB); a circular cylinder of radius 48 about
    (500, 512) extruded over z, plain bounce-back (no signed distance: its 4 bytes per cell
    would not fit beside the FP32 populations), D3Q19 Cs = 0.1, u = 0.05, tau = 0.6,
    free-slip y/z; the geometry is built on the device, `warmup` untimed steps, then
    `steps` steps timed with HIP events on the solver's stream.

depth — the step's rate against the lattice depth: 512 x 256 x nz around a circular
    cylinder of radius 25.6 about (128, 128) extruded over z (plain bounce-back), D3Q19 FP32
    ping-pong, Cs = 0.1, u = 0.05, tau = 0.6; free-slip y/z walls, and the same periodic in
    y and z; 3 untimed steps, then `steps` steps timed with HIP events.

sensitivity — whether the developed wing flow amplifies a small local perturbation (and
    FP32 rounding) in FP64: sensitivity_study.

drain — the precision case's D3Q15 run (tau 0.6) from rest in FP32 and FP16: mean
    density, mean speed and Cd every 1000 steps, showing how the channel's excess mass
    drains through the outlet in each format.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import geometry
from .lbm import LatticeModel, Layout, LbmSolver, Params, Storage, VoxelGrid

HBM_PEAK_GBS = 8000.0


def c2_grid():
    return wing_grid(512, 256, 256)


def c3_grid():
    return geometry.ellipsoid_and_spheres_grid(768, 512, 576, semi_axes=(96, 40, 40), n_spheres=64, r_range=(4, 12),
                                               seed=0)


def _solver(grid, model, storage, layout):
    s = LbmSolver()
    if not s.Init(None, grid, model, storage=storage, layout=layout):
        raise RuntimeError(s.LastError())
    return s


def force_series(solver, steps, every):
    """Steps `steps` times, evaluating the FP64 force sums every `every` steps on the
    device; returns an (steps // every, 4) array (Fx, Fy, Fz, maxStress)."""
    import torch
    n = steps // every
    buf = torch.zeros((n, 4), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    for k in range(n):
        solver.Step(None, every)
        solver.ForcesDevice(buf[k])
    solver.Sync()
    return buf.cpu().numpy()


# D3Q15 regularized BGK with this inlet diverges within ~100 steps for tau up to 0.55
# (u = 0.08, free-slip walls, even in an empty channel) — in FlowCraft's compiled
# LbmReference and in this engine alike, in every precision; tau = 0.6 is stable.
TAU_D3Q15 = 0.6


def wing_grid(nx=512, ny=256, nz=256):
    """The precision workload's NACA 0012 wing, scaled with the lattice: chord nx/4, leading
    edge at x = nx/4 on the centre line, AoA 4 deg, extruded over z (512x256x256: the
    benchmark wing)."""
    return geometry.naca_wing_grid(nx, ny, nz, m=0.0, p=0.0, t=0.12, chord=nx / 4.0, le_x=nx / 4.0, le_y=ny / 2.0,
                                   aoa_deg=4.0)


# A control run: FP64 with the inlet speed raised by this fraction (a change the size of
# FP32 rounding, acting at every step; the coefficients stay normalised by the nominal
# speed, which moves them by 2e-6).  If the flow amplified small perturbations this run
# would leave the unperturbed FP64 run as far as the storage formats do; on the wing it
# stays with it (Cd differences of 1e-8 to 2e-6 over 5000 steps at tau 0.5005).
# (Noise added to the rest state alone is not enough: it leaves as acoustic waves before
# the wake becomes unsteady.)
PERTURBATION = 1e-6


def precision_study(steps=5000, every=50, window=1000, models=("D3Q19", "D3Q15"), storages=("f64", "f32", "f16", "bf16"),
                    layout=Layout.PINGPONG, log=print, tau=0.5005, tau15=TAU_D3Q15, dims=(512, 256, 256)):
    """Cd/Cl per storage against FP64 on the wing (wing_grid(*dims)), from the uniform rest
    start.  Storage names: f64, f32, f16, bf16, f16sr, bf16sr, and f64p — the FP64 chaos
    control with the inlet speed raised by PERTURBATION."""
    import torch
    nx, ny, nz = dims
    grid = wing_grid(nx, ny, nz)
    u, chord = 0.08, nx / 4.0
    q = 0.5 * u * u * chord * nz
    st_map = {"f64": Storage.F64, "f32": Storage.F32, "f16": Storage.F16, "bf16": Storage.BF16,
              "f16sr": Storage.F16_SR, "bf16sr": Storage.BF16_SR, "f64p": Storage.F64}
    w = max(1, window // every)
    out = {"workload": f"{nx}x{ny}x{nz} NACA 0012 chord {chord:g} AoA 4 deg, Cs 0.1, u 0.08, free-slip y/z; "
                       f"tau {tau} (D3Q15: {tau15})",
           "steps": steps, "every": every, "window": window, "area": chord * nz, "layout": layout.name.lower(),
           "perturbation": PERTURBATION, "models": {}}
    for mname in models:
        model = LatticeModel[mname]
        t_m = tau15 if model == LatticeModel.D3Q15 else tau
        params = Params(tau=t_m, inletVelocity=u, smagorinskyCs=0.1)
        res = {"tau": t_m}
        for sname in storages:
            s = _solver(grid, model, st_map[sname], layout)
            if sname == "f64p":
                pp = Params(**vars(params))
                pp.inletVelocity = params.inletVelocity * (1.0 + PERTURBATION)
                s.Reset(pp, False)
            else:
                s.Reset(params, False)
            t0 = time.perf_counter()
            f = force_series(s, steps, every)
            dt = time.perf_counter() - t0
            div = s.HasDiverged()
            s.Shutdown()
            cd, cl = f[:, 0] / q, f[:, 1] / q
            first_bad = next((k for k in range(len(cd)) if not (math.isfinite(cd[k]) and math.isfinite(cl[k]))), None)
            res[sname] = {"cd_last": float(cd[-1]), "cl_last": float(cl[-1]), "cd_mean": float(cd[-w:].mean()),
                          "cl_mean": float(cl[-w:].mean()), "diverged": bool(div),
                          "first_nonfinite_step": None if first_bad is None else (first_bad + 1) * every,
                          "wall_s": round(dt, 2), "cd_series": [float(v) for v in cd],
                          "cl_series": [float(v) for v in cl]}
            log(f"  {mname} {sname}: Cd {res[sname]['cd_mean']:.5f} Cl {res[sname]['cl_mean']:.5f} "
                f"(trailing {window}) last Cd {cd[-1]:.5f} Cl {cl[-1]:.5f} diverged {div} {dt:.1f}s")
            torch.cuda.empty_cache()
        ref = res.get("f64")
        if ref is not None:
            for sname, r in res.items():
                if sname in ("f64", "tau"):
                    continue
                for k in ("cd_mean", "cl_mean", "cd_last", "cl_last"):
                    r[k + "_relerr_vs_f64"] = float(abs(r[k] - ref[k]) / max(abs(ref[k]), 1e-30))
        out["models"][mname] = res
    return out


def large_lattice(steps=200, warmup=50, storages=("f32", "f16"), layout=Layout.INPLACE, log=print):
    import torch
    t0 = time.time()
    grid = c3_grid()
    log(f"  geometry {time.time() - t0:.1f}s, solid cells {int(grid.solid.sum())}")
    params = Params(tau=0.501, inletVelocity=0.06, smagorinskyCs=0.1)
    st_map = {"f32": (Storage.F32, 4), "f16": (Storage.F16, 2), "bf16": (Storage.BF16, 2),
              "f16sr": (Storage.F16_SR, 2), "bf16sr": (Storage.BF16_SR, 2)}
    out = {"workload": "768x512x576 (226M cells) ellipsoid (96,40,40) + 64 spheres r 4-12 seed 0, D3Q19 Cs 0.1, "
                       "u 0.06, tau 0.501, free-slip y/z", "steps": steps, "warmup": warmup,
           "layout": layout.name.lower(), "runs": {}}
    cells = grid.CellCount()
    for sname in storages:
        st, b = st_map[sname]
        s = _solver(grid, LatticeModel.D3Q19, st, layout)
        s.Reset(params, False)
        stream = torch.cuda.ExternalStream(s.Stream())
        s.Step(None, warmup)
        s.Sync()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        s.Step(None, steps)
        e1.record(stream)
        s.Sync()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        mlups = cells * steps / (ms * 1e-3) / 1e6
        gbs = mlups * 1e6 * 2 * 19 * b / 1e9
        st_ = s.ReadbackStats()
        r = {"mlups": round(mlups, 1), "ms_per_step": round(ms / steps, 4), "achieved_GBs": round(gbs, 1),
             "frac_hbm_peak": round(gbs / HBM_PEAK_GBS, 4), "population_bytes": s.PopulationBytes(),
             "finite": not st_.hasNaN}
        s.Shutdown()
        torch.cuda.empty_cache()
        out["runs"][sname] = r
        log(f"  {sname}: {r}")
    return out


def capacity(runs=(("f16sr", 2047), ("f32", 1400)), nx=2048, ny=1024, steps=10, warmup=3, log=print):
    import torch
    st_map = {"f32": (Storage.F32, 4), "f16": (Storage.F16, 2), "f16sr": (Storage.F16_SR, 2),
              "bf16sr": (Storage.BF16_SR, 2)}
    params = Params(tau=0.6, inletVelocity=0.05, smagorinskyCs=0.1, freeSlipY=True, freeSlipZ=True)
    out = {"workload": f"{nx}x{ny}xnz circular cylinder r 48 about (500, 512) extruded over z, plain bounce-back, "
                       "D3Q19 Cs 0.1, u 0.05, tau 0.6, free-slip y/z, in place", "steps": steps, "warmup": warmup,
           "runs": {}}
    yy, xx = np.mgrid[0:ny, 0:nx]
    layer = torch.from_numpy((np.hypot(xx + 0.5 - 500.0, yy + 0.5 - 512.0) < 48.0).astype(np.uint8))
    for sname, nz in runs:
        st, b = st_map[sname]
        cells = nx * ny * nz
        t0 = time.time()
        s = LbmSolver()
        ok = s.Init(None, VoxelGrid(nx, ny, nz, np.zeros(cells, np.uint8), None), LatticeModel.D3Q19, storage=st,
                    layout=Layout.INPLACE)
        if not ok:
            raise RuntimeError(s.LastError())
        solid = layer.cuda().reshape(1, ny, nx).expand(nz, ny, nx).contiguous()
        s.SetGeometryDevice(solid.reshape(-1))
        del solid
        torch.cuda.empty_cache()
        s.Reset(params, True)
        stream = torch.cuda.ExternalStream(s.Stream())
        s.Step(None, warmup)
        s.Sync()
        setup = time.time() - t0
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        s.Step(None, steps)
        e1.record(stream)
        s.Sync()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1)
        mlups = cells * steps / (ms * 1e-3) / 1e6
        gbs = mlups * 1e6 * 2 * 19 * b / 1e9
        st_ = s.ReadbackStats()
        r = {"nz": nz, "cells": cells, "population_bytes": s.PopulationBytes(),
             "device_bytes_in_use": int(torch.cuda.mem_get_info()[1] - torch.cuda.mem_get_info()[0]),
             "mlups": round(mlups, 1), "ms_per_step": round(ms / steps, 3), "achieved_GBs": round(gbs, 1),
             "frac_hbm_peak": round(gbs / HBM_PEAK_GBS, 4), "setup_s": round(setup, 1), "finite": not st_.hasNaN}
        s.Shutdown()
        torch.cuda.empty_cache()
        out["runs"][sname] = r
        log(f"  {sname}: {r}")
    return out


def depth_sweep(nzs=(64, 256, 1024, 2048), nx=512, ny=256, steps=10, warmup=3, log=print):
    import torch
    yy, xx = np.mgrid[0:ny, 0:nx]
    layer = torch.from_numpy((np.hypot(xx + 0.5 - nx / 4, yy + 0.5 - ny / 2) < ny / 10).astype(np.uint8))
    out = {"workload": f"{nx}x{ny}xnz circular cylinder r {ny / 10} about ({nx / 4}, {ny / 2}) extruded over z, plain "
                       "bounce-back, D3Q19 FP32 ping-pong, Cs 0.1, u 0.05, tau 0.6", "steps": steps, "warmup": warmup,
           "runs": []}
    for periodic in (False, True):
        for nz in nzs:
            cells = nx * ny * nz
            s = LbmSolver()
            ok = s.Init(None, VoxelGrid(nx, ny, nz, np.zeros(cells, np.uint8), None), LatticeModel.D3Q19,
                        storage=Storage.F32, layout=Layout.PINGPONG, periodicY=periodic, periodicZ=periodic)
            if not ok:
                raise RuntimeError(s.LastError())
            solid = layer.cuda().reshape(1, ny, nx).expand(nz, ny, nx).contiguous()
            s.SetGeometryDevice(solid.reshape(-1))
            del solid
            torch.cuda.empty_cache()
            s.Reset(Params(tau=0.6, inletVelocity=0.05, smagorinskyCs=0.1, freeSlipY=True, freeSlipZ=True), True)
            stream = torch.cuda.ExternalStream(s.Stream())
            s.Step(None, warmup)
            s.Sync()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            s.Step(None, steps)
            e1.record(stream)
            s.Sync()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / steps
            r = {"nz": nz, "walls": "periodic y/z" if periodic else "free-slip y/z", "ms_per_step": round(ms, 4),
                 "glups": round(cells / ms / 1e6, 2), "TBs": round(cells * 152 / ms / 1e9, 3),
                 "us_per_layer": round(ms * 1000 / nz, 3)}
            s.Shutdown()
            torch.cuda.empty_cache()
            out["runs"].append(r)
            log(f"  {r}")
    return out


def density_drain(steps=20000, every=1000, model="D3Q15", tau=TAU_D3Q15, storages=("f32", "f16"), log=print):
    """The 512x256x256 wing from rest (u 0.08, Cs 0.1), mean density, mean speed and Cd
    every `every` steps per storage: how the channel's excess mass drains through the
    outlet in each population format."""
    grid = c2_grid()
    q = 0.5 * 0.08 * 0.08 * 128.0 * 256
    st_map = {"f64": Storage.F64, "f32": Storage.F32, "f16": Storage.F16, "bf16": Storage.BF16,
              "f16sr": Storage.F16_SR, "bf16sr": Storage.BF16_SR}
    out = {"workload": f"512x256x256 NACA 0012 chord 128 AoA 4 deg, {model} tau {tau}, Cs 0.1, u 0.08, free-slip y/z",
           "every": every, "runs": {}}
    for sname in storages:
        s = _solver(grid, LatticeModel[model], st_map[sname], Layout.PINGPONG)
        s.Reset(Params(tau=tau, inletVelocity=0.08, smagorinskyCs=0.1), False)
        rows = []
        for k in range(steps // every):
            s.Step(None, every)
            t = s.ReadbackStats()
            rows.append({"step": (k + 1) * every, "avg_density": t.avgDensity, "avg_speed": t.avgSpeed,
                         "total_mass": t.totalMass, "cd": float(s.Forces()[0] / q)})
            log(f"  {sname} {rows[-1]}")
        s.Shutdown()
        out["runs"][sname] = rows
    return out


def sensitivity_study(t0=1000, steps=4000, every=50, eps=1e-9, tau=0.5005, dims=(512, 256, 256), log=print):
    """How the developed wing flow answers a small local perturbation, in FP64.  From rest the
    FP64 engine runs t0 steps; the exact state is dumped and continued `steps` steps three
    ways: unperturbed; with the populations of a 4x4x4-cell box in the wake (half a chord
    behind the trailing edge, at mid-height and mid-span) scaled by (1 + eps); and in FP32
    from the same state (rounded once).  Reports Cd/Cl of each and their relative distance
    from the unperturbed run over time: a distance that grows exponentially means the flow
    amplifies small differences on this horizon (chaos); one that stays near its initial
    size means it does not."""
    import torch
    nx, ny, nz = dims
    grid = wing_grid(nx, ny, nz)
    u, chord = 0.08, nx / 4.0
    q = 0.5 * u * u * chord * nz
    params = Params(tau=tau, inletVelocity=u, smagorinskyCs=0.1)
    s = _solver(grid, LatticeModel.D3Q19, Storage.F64, Layout.PINGPONG)
    s.Reset(params, False)
    s.Step(None, t0)
    s.Sync()
    pops = s.Populations()
    s.Shutdown()
    x0 = int(nx / 4 + chord * 1.5)
    box = (slice(None), slice(nz // 2 - 2, nz // 2 + 2), slice(ny // 2 - 2, ny // 2 + 2), slice(x0, x0 + 4))
    runs = {}
    for name, st in (("base", Storage.F64), ("perturbed", Storage.F64), ("f32", Storage.F32)):
        r = _solver(grid, LatticeModel.D3Q19, st, Layout.PINGPONG)
        r.Reset(params, False)
        if name == "perturbed":
            p2 = pops.copy()
            p2[box] *= 1.0 + eps
            r.SetPopulations(p2)
            del p2
        else:
            r.SetPopulations(pops)
        f = force_series(r, steps, every)
        r.Shutdown()
        runs[name] = (f[:, 0] / q, f[:, 1] / q)
        log(f"  {name}: Cd {runs[name][0][-1]:.6f} Cl {runs[name][1][-1]:.6f}")
    base_cd, base_cl = runs["base"]
    out = {"workload": f"{nx}x{ny}x{nz} NACA 0012 wing, D3Q19 FP64, tau {tau}, Cs 0.1, u 0.08; state after {t0} "
                       f"steps from rest, continued {steps} steps", "eps": eps, "every": every,
           "box": [x0, ny // 2 - 2, nz // 2 - 2, 4], "cd_base": [float(v) for v in base_cd], "series": {}}
    for name in ("perturbed", "f32"):
        cd, cl = runs[name]
        dcd = np.abs(cd - base_cd) / np.abs(base_cd)
        dcl = np.abs(cl - base_cl) / np.abs(base_cl)
        out["series"][name] = {"rel_cd": [float(v) for v in dcd], "rel_cl": [float(v) for v in dcl]}
        k = len(dcd)
        marks = {str((i + 1) * every): [float(dcd[i]), float(dcl[i])] for i in (0, k // 8, k // 4, k // 2, k - 1)}
        out["series"][name]["marks"] = marks
        log(f"  {name}: relative |dCd|, |dCl| at steps " + ", ".join(f"{m}: {v[0]:.2e} {v[1]:.2e}" for m, v in marks.items()))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m flowcraft_amd.studies")
    sub = ap.add_subparsers(dest="cmd", required=True)
    p = sub.add_parser("precision")
    p.add_argument("--steps", type=int, default=5000)
    p.add_argument("--every", type=int, default=50)
    p.add_argument("--window", type=int, default=1000)
    p.add_argument("--models", nargs="+", default=["D3Q19", "D3Q15"])
    p.add_argument("--storages", nargs="+", default=["f64", "f32", "f16", "bf16"])
    p.add_argument("--tau", type=float, default=0.5005)
    p.add_argument("--tau15", type=float, default=TAU_D3Q15)
    p.add_argument("--dims", type=int, nargs=3, default=[512, 256, 256])
    p.add_argument("--out", default=None)
    g = sub.add_parser("large")
    g.add_argument("--steps", type=int, default=200)
    g.add_argument("--warmup", type=int, default=50)
    g.add_argument("--storages", nargs="+", default=["f32", "f16"])
    g.add_argument("--layout", default="inplace", choices=["inplace", "pingpong"])
    g.add_argument("--out", default=None)
    sv = sub.add_parser("sensitivity")
    sv.add_argument("--t0", type=int, default=1000)
    sv.add_argument("--steps", type=int, default=4000)
    sv.add_argument("--every", type=int, default=50)
    sv.add_argument("--eps", type=float, default=1e-9)
    sv.add_argument("--tau", type=float, default=0.5005)
    sv.add_argument("--dims", type=int, nargs=3, default=[512, 256, 256])
    sv.add_argument("--out", default=None)
    c = sub.add_parser("capacity")
    c.add_argument("--steps", type=int, default=10)
    c.add_argument("--warmup", type=int, default=3)
    c.add_argument("--out", default=None)
    dp = sub.add_parser("depth")
    dp.add_argument("--nz", type=int, nargs="+", default=[64, 256, 1024, 2048])
    dp.add_argument("--steps", type=int, default=10)
    dp.add_argument("--out", default=None)
    d = sub.add_parser("drain")
    d.add_argument("--steps", type=int, default=20000)
    d.add_argument("--every", type=int, default=1000)
    d.add_argument("--storages", nargs="+", default=["f32", "f16"])
    d.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    import torch
    torch.cuda.init()

    def log(m):
        print(m, flush=True)

    if a.cmd == "precision":
        r = precision_study(a.steps, a.every, a.window, tuple(a.models), tuple(a.storages), log=log, tau=a.tau,
                            tau15=a.tau15, dims=tuple(a.dims))
    elif a.cmd == "sensitivity":
        r = sensitivity_study(a.t0, a.steps, a.every, a.eps, a.tau, tuple(a.dims), log=log)
    elif a.cmd == "depth":
        r = depth_sweep(tuple(a.nz), steps=a.steps, log=log)
    elif a.cmd == "capacity":
        r = capacity(steps=a.steps, warmup=a.warmup, log=log)
    elif a.cmd == "drain":
        r = density_drain(a.steps, a.every, storages=tuple(a.storages), log=log)
    else:
        r = large_lattice(a.steps, a.warmup, tuple(a.storages),
                          Layout.INPLACE if a.layout == "inplace" else Layout.PINGPONG, log=log)
    s = json.dumps(r)
    if a.out:
        with open(a.out, "w") as f:
            f.write(s)
    print(s, flush=True)


if __name__ == "__main__":
    sys.exit(main())
