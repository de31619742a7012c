"""Deterministic triangle meshes (world space, millimetres) for the voxelizer: the
stand-ins for the STL models FlowCraft loads.  Every mesh is returned as
(positions float32 [n, 3], indices uint32 [3 * triangles]) with outward-facing
counter-clockwise winding (normal = (v1 - v0) x (v2 - v0) points out of the body), the
convention the voxelizer's signed distance relies on.

World axes follow FlowCraft's voxel grid (FlowCraft/VoxelGrid.h:38-49): the flow runs
along world Y (grid x), world X is lateral (grid y), world Z is up.
"""
from __future__ import annotations

import math

import numpy as np

from .geometry import naca4_outline


def _orient(pos: np.ndarray, tri: np.ndarray, inner: np.ndarray) -> np.ndarray:
    """Flip triangles whose normal points toward their interior reference point
    (`inner`, one point per triangle)."""
    a, b, c = pos[tri[:, 0]], pos[tri[:, 1]], pos[tri[:, 2]]
    n = np.cross(b - a, c - a)
    out = (a + b + c) / 3.0 - inner
    flip = (n * out).sum(axis=1) < 0
    tri = tri.copy()
    tri[flip, 1], tri[flip, 2] = tri[flip, 2].copy(), tri[flip, 1].copy()
    return tri


def _pack(pos: np.ndarray, tri: np.ndarray):
    return pos.astype(np.float32), tri.astype(np.uint32).reshape(-1)


def ellipsoid(center=(0.0, 0.0, 0.0), semi_axes=(10.0, 10.0, 10.0), n_lon: int = 64, n_lat: int = 32):
    """Closed latitude-longitude ellipsoid: n_lon * (n_lat - 1) + 2 vertices."""
    cx, cy, cz = center
    a, b, c = semi_axes
    th = np.linspace(0.0, math.pi, n_lat + 1)[1:-1]
    ph = np.linspace(0.0, 2 * math.pi, n_lon, endpoint=False)
    T, P = np.meshgrid(th, ph, indexing="ij")
    ring = np.stack([cx + a * np.sin(T) * np.cos(P), cy + b * np.sin(T) * np.sin(P), cz + c * np.cos(T)], -1)
    pos = np.concatenate([[[cx, cy, cz + c]], ring.reshape(-1, 3), [[cx, cy, cz - c]]])
    top, bot = 0, len(pos) - 1
    vid = lambda i, j: 1 + i * n_lon + (j % n_lon)
    tris = []
    for j in range(n_lon):
        tris.append((top, vid(0, j), vid(0, j + 1)))
        tris.append((bot, vid(n_lat - 2, j + 1), vid(n_lat - 2, j)))
    for i in range(n_lat - 2):
        for j in range(n_lon):
            tris.append((vid(i, j), vid(i + 1, j), vid(i + 1, j + 1)))
            tris.append((vid(i, j), vid(i + 1, j + 1), vid(i, j + 1)))
    tri = np.array(tris, np.int64)
    return _pack(pos, _orient(pos, tri, np.tile(np.array(center, float), (len(tri), 1))))


def wing(m=0.02, p=0.4, t=0.12, chord=100.0, span=400.0, root=(0.0, 0.0, 0.0), span_axis: str = "x",
         aoa_deg: float = 0.0, n_chord: int = 96, n_span: int = 64):
    """Closed straight NACA 4-digit wing: chord along +Y from the leading edge at `root`,
    thickness along Z (the section pitched nose-up by `aoa_deg`), extruded over `span`
    along `span_axis` ("x": lateral, centred on root; "z": upward fin from root)."""
    o = naca4_outline(m, p, t, n_chord)[:-1] * chord  # drop the repeated trailing-edge point
    ca, sa = math.cos(math.radians(aoa_deg)), math.sin(math.radians(aoa_deg))
    # section coordinates (s along the chord, h normal to it), pitched about the leading edge
    s = o[:, 0] * ca + o[:, 1] * sa
    h = -o[:, 0] * sa + o[:, 1] * ca
    k = len(o)
    inner2 = np.array([0.35 * chord * ca, -0.35 * chord * sa])  # an interior point of the section
    if span_axis == "x":
        ws = np.linspace(-0.5 * span, 0.5 * span, n_span + 1)
        sec = lambda w: np.stack([np.full(k, w), s, h], -1)
        inner_at = lambda w: np.array([w, inner2[0], inner2[1]])
    else:
        ws = np.linspace(0.0, span, n_span + 1)
        sec = lambda w: np.stack([h, s, np.full(k, w)], -1)
        inner_at = lambda w: np.array([inner2[1], inner2[0], w])
    pos = np.concatenate([sec(w) for w in ws]) + np.array(root, float)
    tris, inner = [], []
    for i in range(n_span):
        wm = 0.5 * (ws[i] + ws[i + 1])
        for j in range(k):
            a, b = i * k + j, i * k + (j + 1) % k
            c, d = a + k, b + k
            tris += [(a, b, d), (a, d, c)]
            inner += [inner_at(wm)] * 2
    # end caps: fans around the section's interior point
    for i, w in ((0, ws[0]), (n_span, ws[-1])):
        ci = len(pos)
        pos = np.concatenate([pos, [inner_at(w) + np.array(root, float)]])
        out = -1.0 if i == 0 else 1.0
        for j in range(k):
            tris.append((ci, i * k + j, i * k + (j + 1) % k))
            off = np.zeros(3)
            off[0 if span_axis == "x" else 2] = -out
            inner.append(inner_at(w) + np.array(root, float) + off)
    tri = np.array(tris, np.int64)
    inner = np.array(inner, float)
    inner[: 2 * n_span * k] += np.array(root, float)
    return _pack(pos, _orient(pos, tri, inner))


def merge(*meshes):
    """Concatenate meshes (indices offset); overlapping bodies stay separate shells."""
    pos, idx, base = [], [], 0
    for p, i in meshes:
        pos.append(p)
        idx.append(i.astype(np.int64) + base)
        base += len(p)
    return np.concatenate(pos).astype(np.float32), np.concatenate(idx).astype(np.uint32)


def aircraft(scale: float = 1.0, detail: int = 1):
    """A high-wing light aircraft flying toward -Y: ellipsoid fuselage, NACA 2412 main
    wing, NACA 0012 tailplane and fin.  `detail` multiplies the tessellation: 60,500
    triangles at detail=1, 241,844 at detail=2."""
    d = detail
    L = 700.0 * scale
    fus = ellipsoid((0.0, 0.0, 0.0), (55.0 * scale, 0.5 * L, 60.0 * scale), n_lon=96 * d, n_lat=48 * d)
    main = wing(0.02, 0.4, 0.12, 150.0 * scale, 1000.0 * scale, (0.0, -0.22 * L, 45.0 * scale), "x", 2.0,
                n_chord=96 * d, n_span=96 * d)
    tail = wing(0.0, 0.0, 0.10, 90.0 * scale, 320.0 * scale, (0.0, 0.33 * L, 10.0 * scale), "x", 0.0,
                n_chord=64 * d, n_span=32 * d)
    fin = wing(0.0, 0.0, 0.10, 100.0 * scale, 150.0 * scale, (0.0, 0.31 * L, 20.0 * scale), "z", 0.0,
               n_chord=64 * d, n_span=24 * d)
    return merge(fus, main, tail, fin)


def read_stl(path: str):
    """Binary or ASCII STL -> (positions, indices) with one vertex per corner (the
    voxelizer needs no vertex sharing)."""
    with open(path, "rb") as f:
        data = f.read()
    if len(data) >= 84:
        n = int.from_bytes(data[80:84], "little")
        if 84 + 50 * n == len(data):
            rec = np.frombuffer(data, dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]),
                                count=n, offset=84)
            pos = rec["v"].reshape(-1, 3).astype(np.float32)
            return pos, np.arange(len(pos), dtype=np.uint32)
    verts = [list(map(float, ln.split()[1:4])) for ln in data.decode("ascii", "replace").splitlines()
             if ln.strip().startswith("vertex")]
    pos = np.array(verts, np.float32).reshape(-1, 3)
    return pos, np.arange(len(pos), dtype=np.uint32)
