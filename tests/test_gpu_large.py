"""A lattice past 2^31 cells on one GPU.

Cell and padded-cell indices of the engine are 64-bit; the lattice limit is set by one
work-item per cell in a 32-bit grid dimension (at most 2^32 - 2^20 cells) and by memory.
The check uses a flow whose every z layer is the same: a circular cylinder extruded over
a periodic z span.  Each cell then updates exactly as in a four-layer twin of the lattice
(same x/y cross-section, same arithmetic), so every layer of the 2.2e9-cell run must equal
the twin's layer 0 bit for bit, and the twin itself equals the CPU oracle's emulation of
the storage format.  FP16 offsets rounded to nearest: a stochastically rounded format
would draw different bits per layer (its random stream depends on the cell index).
"""
from __future__ import annotations

import numpy as np
import pytest

from flowcraft_amd.lbm import LatticeModel, Layout, LbmSolver, Params, Storage, VoxelGrid
from tests._cases import make_oracle, oracle_params, ulp_diff32

pytestmark = pytest.mark.gpu

NX, NY = 2048, 512
NZ_BIG, NZ_TWIN = 2100, 4
STEPS = 6
PARAMS = Params(tau=0.6, inletVelocity=0.05, smagorinskyCs=0.1, freeSlipY=True)


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu-case] {request.node.nodeid}", flush=True)
    yield


def _cylinder_layer():
    """One z layer: a circular cylinder of radius 24 cells about (400.3, 255.7) (cell centres
    at i + 0.5), with its exact signed distance (+fluid, -solid)."""
    y, x = np.mgrid[0:NY, 0:NX].astype(np.float64)
    d = np.hypot(x + 0.5 - 400.3, y + 0.5 - 255.7) - 24.0
    sdf = d.astype(np.float32)
    return (sdf < 0).astype(np.uint8), sdf


def test_lattice_past_2_31_cells_matches_its_four_layer_twin():
    import torch

    n_big = NX * NY * NZ_BIG
    assert n_big > 2**31 and NX * (NY + 2) * (NZ_BIG + 2) > 2**31
    solid, sdf = _cylinder_layer()

    # the twin: 4 layers, engine against the oracle's FP16 emulation
    twin_grid = VoxelGrid(NX, NY, NZ_TWIN, np.ascontiguousarray(np.broadcast_to(solid, (NZ_TWIN, NY, NX))),
                          np.ascontiguousarray(np.broadcast_to(sdf, (NZ_TWIN, NY, NX))))
    twin = LbmSolver()
    assert twin.Init(None, twin_grid, LatticeModel.D3Q19, storage=Storage.F16, layout=Layout.INPLACE,
                     periodicZ=True), twin.LastError()
    twin.Reset(PARAMS, True)
    twin.Step(None, STEPS)
    twin_field = twin.ReadbackField().reshape(NZ_TWIN, NY * NX * 4)
    twin_forces = twin.Forces()
    twin_stats = twin.ReadbackStats()
    orc = make_oracle(twin_grid, LatticeModel.D3Q19, Storage.F16, periodic=(False, True))
    oracle_params(orc, PARAMS)
    orc.reset(True)
    orc.step(STEPS)
    o = orc.field().reshape(NZ_TWIN, -1)
    assert np.array_equal(twin_field.view(np.uint32), o.view(np.uint32)), f"twin vs oracle: {ulp_diff32(twin_field, o)} ulp"
    for z in range(1, NZ_TWIN):
        assert np.array_equal(twin_field[z].view(np.uint32), twin_field[0].view(np.uint32))
    del orc, o
    print(f"twin == oracle; twin force {twin_forces}", flush=True)

    # the big lattice: host grid without the obstacle, then the geometry from device tensors
    empty = VoxelGrid(NX, NY, NZ_BIG, np.zeros(n_big, np.uint8), None)
    big = LbmSolver()
    assert big.Init(None, empty, LatticeModel.D3Q19, storage=Storage.F16, layout=Layout.INPLACE,
                    periodicZ=True), big.LastError()
    del empty
    dsolid = torch.from_numpy(solid).cuda().reshape(1, NY, NX).expand(NZ_BIG, NY, NX).contiguous()
    dsdf = torch.from_numpy(sdf).cuda().reshape(1, NY, NX).expand(NZ_BIG, NY, NX).contiguous()
    big.SetGeometryDevice(dsolid.reshape(-1), dsdf.reshape(-1))
    del dsolid, dsdf
    torch.cuda.empty_cache()
    assert big.SurfaceCells() == twin.SurfaceCells() // NZ_TWIN * NZ_BIG
    big.Reset(PARAMS, True)
    big.Step(None, STEPS)
    print(f"big lattice {NX}x{NY}x{NZ_BIG} = {n_big} cells stepped {STEPS} times", flush=True)

    field = torch.empty(n_big * 4, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    big.FieldDevice(field)
    big.Sync()
    layers = field.view(torch.int32).view(NZ_BIG, NY * NX * 4)
    ref = torch.from_numpy(twin_field[0].view(np.int32)).cuda()
    for z0 in range(0, NZ_BIG, 256):
        chunk = layers[z0:z0 + 256]
        bad = (chunk != ref.unsqueeze(0)).any(dim=1)
        assert not bool(bad.any()), f"layers differing from the twin: {(torch.nonzero(bad).flatten() + z0).tolist()[:8]}"
    del field, layers

    forces = big.Forces()
    scale = NZ_BIG / NZ_TWIN
    fscale = (abs(twin_forces[0]) + abs(twin_forces[1]) + abs(twin_forces[2])) * scale
    for k in range(3):
        assert abs(forces[k] - twin_forces[k] * scale) <= 1e-9 * fscale, (k, forces, twin_forces)
    stats = big.ReadbackStats()
    assert stats.fluidCells == twin_stats.fluidCells // NZ_TWIN * NZ_BIG
    assert abs(stats.totalMass - twin_stats.totalMass * scale) <= 1e-9 * twin_stats.totalMass * scale
    print(f"every layer == twin layer 0; forces {forces}; mass {stats.totalMass}", flush=True)
    big.Shutdown()
    twin.Shutdown()
