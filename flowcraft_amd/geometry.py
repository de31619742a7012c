"""Analytic obstacle generators producing ``VoxelGrid``s (solid mask + SDF).

The solver reads geometry only through FlowCraft's ``VoxelGrid`` convention
(FlowCraft/VoxelGrid.h:13-61): ``solid`` 1 = solid, ``signedDistance`` positive
in fluid / negative in solid, in cell units here (cellSize = 1), sampled at
cell centres ``(x + 0.5, y + 0.5, z + 0.5)``; solid is ``sdf < 0``.  Cells
outside every obstacle's bounding box (plus a margin) keep +inf, like the
reference voxelizer's "never near a triangle" sentinel — the Bouzidi scheme
only ever reads the SDF on links that cross a surface.

All arithmetic is float32 in the order the reference test writes it
(tests/gpu_vs_cpu_test.cpp:29-41 for the sphere), so masks and SDFs are
reproducible bit for bit.  Random layouts use numpy's PCG64 with an explicit
seed.
"""
from __future__ import annotations

import math

import numpy as np

from .lbm import VoxelGrid

_F = np.float32
_MARGIN = 3


def _axes(lo, hi):
    return np.arange(lo, hi, dtype=_F) + _F(0.5)


def _box(c, r, n):
    lo = max(0, int(math.floor(c - r)) - _MARGIN)
    hi = min(n, int(math.ceil(c + r)) + _MARGIN + 1)
    return lo, hi


def empty_grid(nx: int, ny: int, nz: int, sdf_fill: float = np.inf) -> VoxelGrid:
    return VoxelGrid(nx, ny, nz, np.zeros((nz, ny, nx), np.uint8), np.full((nz, ny, nx), sdf_fill, _F))


def _finish(g: VoxelGrid) -> VoxelGrid:
    g.solid = (g.signedDistance < 0).astype(np.uint8)
    return g


def sphere_sdf(nx, ny, nz, center, radius) -> np.ndarray:
    """Exact sphere SDF everywhere, float32, in the reference test's operation order:
    d = sqrt((x+.5-cx)^2 + (y+.5-cy)^2 + (z+.5-cz)^2) - r."""
    cx, cy, cz = (_F(v) for v in center)
    z, y, x = np.meshgrid(np.arange(nz, dtype=_F), np.arange(ny, dtype=_F), np.arange(nx, dtype=_F), indexing="ij")
    h = _F(0.5)
    dx = x + h - cx
    dy = y + h - cy
    dz = z + h - cz
    return (np.sqrt(dx * dx + dy * dy + dz * dz) - _F(radius)).astype(_F)


def sphere_grid(nx, ny, nz, center, radius) -> VoxelGrid:
    """Sphere with the SDF defined on every cell (tests/gpu_vs_cpu_test.cpp:28-41)."""
    g = VoxelGrid(nx, ny, nz, None, sphere_sdf(nx, ny, nz, center, radius))
    return _finish(g)


def _union(g: VoxelGrid, region, d: np.ndarray) -> None:
    (z0, z1), (y0, y1), (x0, x1) = region
    sub = g.signedDistance[z0:z1, y0:y1, x0:x1]
    np.minimum(sub, d, out=sub)


def add_sphere(g: VoxelGrid, center, radius) -> None:
    cx, cy, cz = center
    x0, x1 = _box(cx, radius, g.nx)
    y0, y1 = _box(cy, radius, g.ny)
    z0, z1 = _box(cz, radius, g.nz)
    if x0 >= x1 or y0 >= y1 or z0 >= z1:
        return
    z, y, x = np.meshgrid(_axes(z0, z1), _axes(y0, y1), _axes(x0, x1), indexing="ij")
    dx, dy, dz = x - _F(cx), y - _F(cy), z - _F(cz)
    d = (np.sqrt(dx * dx + dy * dy + dz * dz) - _F(radius)).astype(_F)
    _union(g, ((z0, z1), (y0, y1), (x0, x1)), d)


def add_ellipsoid(g: VoxelGrid, center, semi_axes, pitch_deg: float = 0.0) -> None:
    """Ellipsoid (optionally pitched about y by ``pitch_deg``, nose up for flow in +x).
    SDF: the first-order distance estimate k0*(k0-1)/k1 with k0 = |p/a|, k1 = |p/a^2|
    (exact sign, exact on spheres, within a few percent near the surface)."""
    cx, cy, cz = center
    a, b, c = semi_axes
    rmax = max(a, b, c)
    x0, x1 = _box(cx, rmax, g.nx)
    y0, y1 = _box(cy, rmax, g.ny)
    z0, z1 = _box(cz, rmax, g.nz)
    if x0 >= x1 or y0 >= y1 or z0 >= z1:
        return
    z, y, x = np.meshgrid(_axes(z0, z1), _axes(y0, y1), _axes(x0, x1), indexing="ij")
    dx, dy, dz = x - _F(cx), y - _F(cy), z - _F(cz)
    th = math.radians(pitch_deg)
    ct, st = _F(math.cos(th)), _F(math.sin(th))
    rx = dx * ct - dz * st
    rz = dx * st + dz * ct
    k0 = np.sqrt((rx / _F(a)) ** 2 + (dy / _F(b)) ** 2 + (rz / _F(c)) ** 2)
    k1 = np.sqrt((rx / _F(a * a)) ** 2 + (dy / _F(b * b)) ** 2 + (rz / _F(c * c)) ** 2)
    k1 = np.maximum(k1, _F(1e-12))
    d = (k0 * (k0 - _F(1)) / k1).astype(_F)
    _union(g, ((z0, z1), (y0, y1), (x0, x1)), d)


def naca4_outline(m: float, p: float, t: float, n: int = 256) -> np.ndarray:
    """Closed NACA 4-digit outline in chord units (leading edge at 0, trailing edge at
    1), cosine-spaced, closed trailing edge (coefficient -0.1036).  Returns (2n-1, 2)
    points going from the trailing edge over the upper surface and back along the lower."""
    beta = np.linspace(0.0, math.pi, n)
    xc = 0.5 * (1.0 - np.cos(beta))
    yt = 5.0 * t * (0.2969 * np.sqrt(xc) - 0.1260 * xc - 0.3516 * xc ** 2 + 0.2843 * xc ** 3 - 0.1036 * xc ** 4)
    if m > 0 and 0 < p < 1:
        yc = np.where(xc < p, m / p ** 2 * (2 * p * xc - xc ** 2), m / (1 - p) ** 2 * ((1 - 2 * p) + 2 * p * xc - xc ** 2))
        dyc = np.where(xc < p, 2 * m / p ** 2 * (p - xc), 2 * m / (1 - p) ** 2 * (p - xc))
    else:
        yc = np.zeros_like(xc)
        dyc = np.zeros_like(xc)
    th = np.arctan(dyc)
    xu, yu = xc - yt * np.sin(th), yc + yt * np.cos(th)
    xl, yl = xc + yt * np.sin(th), yc - yt * np.cos(th)
    upper = np.stack([xu[::-1], yu[::-1]], axis=1)  # TE -> LE
    lower = np.stack([xl[1:], yl[1:]], axis=1)      # LE -> TE
    return np.concatenate([upper, lower], axis=0)


def _polygon_sdf(px: np.ndarray, py: np.ndarray, poly: np.ndarray) -> np.ndarray:
    """Signed distance (float64) from points to a closed polygon: negative inside
    (even-odd rule).  Vectorised over points, looped over edges."""
    best = np.full(px.shape, np.inf)
    inside = np.zeros(px.shape, bool)
    n = len(poly)
    for i in range(n):
        ax, ay = poly[i]
        bx, by = poly[(i + 1) % n]
        ex, ey = bx - ax, by - ay
        wx, wy = px - ax, py - ay
        L2 = ex * ex + ey * ey
        s = np.clip((wx * ex + wy * ey) / L2, 0.0, 1.0) if L2 > 0 else np.zeros_like(px)
        qx, qy = wx - s * ex, wy - s * ey
        np.minimum(best, qx * qx + qy * qy, out=best)
        cond = (ay <= py) != (by <= py)
        with np.errstate(divide="ignore", invalid="ignore"):
            xint = ax + (py - ay) * ex / np.where(ey == 0, 1.0, ey)
        inside ^= cond & (px < xint)
    d = np.sqrt(best)
    return np.where(inside, -d, d)


def _naca_placed(m, p, t, chord, le_x, le_y, aoa_deg) -> np.ndarray:
    """The NACA outline in grid x-y: chord `chord`, leading edge at (le_x, le_y), rotated
    by the angle of attack about the leading edge (trailing edge down)."""
    poly = naca4_outline(m, p, t) * chord
    a = math.radians(aoa_deg)
    ca, sa = math.cos(a), math.sin(a)
    wx = le_x + poly[:, 0] * ca + poly[:, 1] * sa
    wy = le_y - poly[:, 0] * sa + poly[:, 1] * ca
    return np.stack([wx, wy], axis=1)


def add_naca4_extruded(g: VoxelGrid, m: float, p: float, t: float, chord: float, le_x: float, le_y: float,
                       aoa_deg: float, z_range=None) -> None:
    """NACA mpt wing, chord ``chord`` cells, leading edge at (le_x, le_y), rotated by
    the angle of attack about the leading edge (trailing edge down, lift toward +y
    for flow in +x), extruded over z (all of it by default)."""
    wpoly = _naca_placed(m, p, t, chord, le_x, le_y, aoa_deg)
    wx, wy = wpoly[:, 0], wpoly[:, 1]
    x0, x1 = max(0, int(wx.min()) - _MARGIN), min(g.nx, int(math.ceil(wx.max())) + _MARGIN + 1)
    y0, y1 = max(0, int(wy.min()) - _MARGIN), min(g.ny, int(math.ceil(wy.max())) + _MARGIN + 1)
    if x0 >= x1 or y0 >= y1:
        return
    Y, X = np.meshgrid(np.arange(y0, y1) + 0.5, np.arange(x0, x1) + 0.5, indexing="ij")
    d2 = _polygon_sdf(X, Y, wpoly).astype(_F)
    z0, z1 = (0, g.nz) if z_range is None else z_range
    sub = g.signedDistance[z0:z1, y0:y1, x0:x1]
    np.minimum(sub, d2[None, :, :], out=sub)


def naca_wing_grid(nx, ny, nz, m=0.0, p=0.0, t=0.12, chord=128.0, le_x=128.0, le_y=None, aoa_deg=4.0) -> VoxelGrid:
    g = empty_grid(nx, ny, nz)
    add_naca4_extruded(g, m, p, t, chord, le_x, ny / 2.0 if le_y is None else le_y, aoa_deg)
    return _finish(g)


def ellipsoid_grid(nx, ny, nz, center, semi_axes, pitch_deg=0.0) -> VoxelGrid:
    g = empty_grid(nx, ny, nz)
    add_ellipsoid(g, center, semi_axes, pitch_deg)
    return _finish(g)


def ellipsoid_and_spheres_grid(nx, ny, nz, semi_axes=(96, 40, 40), n_spheres=64, r_range=(4, 12), seed=0,
                               center=None) -> VoxelGrid:
    """An ellipsoid plus a seeded random field of spheres in the downstream half
    (x in [nx/2, nx)), radii uniform in r_range."""
    g = empty_grid(nx, ny, nz)
    c = center if center is not None else (nx * 0.25, ny / 2.0, nz / 2.0)
    add_ellipsoid(g, c, semi_axes)
    rng = np.random.Generator(np.random.PCG64(seed))
    for _ in range(n_spheres):
        r = rng.uniform(*r_range)
        cx = rng.uniform(nx / 2.0, nx)
        cy = rng.uniform(0.0, ny)
        cz = rng.uniform(0.0, nz)
        add_sphere(g, (cx, cy, cz), r)
    return _finish(g)


def random_shape_grid(nx, ny, nz, rng: np.random.Generator, chord=64.0):
    """One sample of the warm-start shape distribution: half NACA 4-digit wings
    (m in [0,.06], p in [.2,.6], t in [.08,.18], AoA in [-4,12] deg, chord 64,
    extruded over z), half ellipsoids (semi-axes uniform in 8..32).  Returns the
    grid and a description dict (kind, parameters, frontal area)."""
    if rng.uniform() < 0.5:
        m, p, t = rng.uniform(0, 0.06), rng.uniform(0.2, 0.6), rng.uniform(0.08, 0.18)
        aoa = rng.uniform(-4, 12)
        g = naca_wing_grid(nx, ny, nz, m, p, t, chord, le_x=nx * 0.25, le_y=ny / 2.0, aoa_deg=aoa)
        return g, {"kind": "naca", "m": m, "p": p, "t": t, "aoa": aoa, "chord": chord, "area": chord * nz}
    a, b, c = rng.uniform(8, 32, size=3)
    g = ellipsoid_grid(nx, ny, nz, (nx * 0.25, ny / 2.0, nz / 2.0), (a, b, c))
    return g, {"kind": "ellipsoid", "axes": (a, b, c), "area": math.pi * b * c}


def shape_sdf_at(desc: dict, dims, px, py, pz) -> np.ndarray:
    """Signed distance (float64, grid-cell units) of a random_shape_grid shape at points
    (px, py, pz) in grid coordinates (cell centres at i + 0.5) of the dims = (nx, ny, nz)
    lattice it was placed in: the same surface random_shape_grid voxelizes, evaluated
    anywhere (e.g. at the cell centres of a coarser grid over the same domain)."""
    nx, ny, nz = dims
    px, py, pz = (np.asarray(v, np.float64) for v in (px, py, pz))
    if desc["kind"] == "naca":
        poly = _naca_placed(desc["m"], desc["p"], desc["t"], desc.get("chord", 64.0), nx * 0.25, ny / 2.0,
                            desc["aoa"])
        return _polygon_sdf(px + 0 * pz, py + 0 * pz, poly)
    a, b, c = desc["axes"]
    dx, dy, dz = px - nx * 0.25, py - ny / 2.0, pz - nz / 2.0
    k0 = np.sqrt((dx / a) ** 2 + (dy / b) ** 2 + (dz / c) ** 2)
    k1 = np.maximum(np.sqrt((dx / (a * a)) ** 2 + (dy / (b * b)) ** 2 + (dz / (c * c)) ** 2), 1e-12)
    return k0 * (k0 - 1.0) / k1


def signed_distance_feature(g: VoxelGrid, clip: float = 16.0) -> np.ndarray:
    """SDF clipped to [-clip, clip] with the +-inf sentinel mapped to +-clip: an input
    channel for the warm-start surrogate."""
    d = np.nan_to_num(g.signedDistance.astype(_F), posinf=clip, neginf=-clip)
    return np.clip(d, -clip, clip).astype(_F)
