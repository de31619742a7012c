"""The device voxelizer (fc_voxelize, csrc/voxelize.hip) against the oracle's
restatement of Voxelizer::Voxelize (pinned bit for bit to FlowCraft's own function in
tests/test_voxelizer.py): identical solid mask, signed distance and frame, on bodies that
exercise every branch (shared-edge ties on thin trailing edges, an enclosed cavity,
degenerate triangles, the seed corner inside the surface shell) and at sizes from 10^4
to 10^7 cells; and the voxelized grid feeds the solver directly."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from flowcraft_amd import meshes
from flowcraft_amd.voxelizer import Params, Voxelizer
from tests.test_voxelizer import CASES, _args

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu-case] {request.node.nodeid}", flush=True)
    yield


def _check(pos, idx, res, pad):
    g = Voxelizer.Voxelize(pos, idx, Params(res, pad))
    solid, sdf, cs, org = oracle.voxelize(pos, idx, res, pad)
    assert (g.nz, g.ny, g.nx) == solid.shape
    assert np.float32(g.cellSize) == np.float32(cs) and tuple(np.float32(g.origin)) == tuple(np.float32(org))
    assert np.array_equal(g.solid, solid), f"{int((g.solid != solid).sum())} cells differ"
    assert np.array_equal(g.signedDistance.view(np.uint32), sdf.view(np.uint32))
    return g


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_equals_restatement(name):
    (pos, idx), res, pad = _args(name)
    _check(pos, idx, res, pad)


@pytest.mark.parametrize("res", [320, 640])
def test_device_equals_restatement_aircraft(res):
    """60,500 triangles; 640 is FlowCraft's default resolution (10.7M cells here)."""
    pos, idx = meshes.aircraft()
    g = _check(pos, idx, res, oracle.VOX_DEFAULT_PADDING)
    print(f"grid {g.nx}x{g.ny}x{g.nz}, solid {int(g.solid.sum())}", flush=True)


def test_empty_mesh_gives_empty_grid():
    g = Voxelizer.Voxelize(np.zeros((0, 3), np.float32), np.zeros(0, np.uint32))
    assert g.CellCount() == 0


def test_voxelized_wing_runs_in_the_solver():
    from flowcraft_amd.lbm import LatticeModel, LbmSolver, Params as LbmParams
    pos, idx = meshes.wing(0.02, 0.4, 0.12, 100.0, 200.0, (0.0, 0.0, 0.0), "x", 4.0)
    g = Voxelizer.Voxelize(pos, idx, Params(96))
    s = LbmSolver()
    assert s.Init(None, g, LatticeModel.D3Q19), s.LastError()
    s.Reset(LbmParams(tau=0.6, inletVelocity=0.05, smagorinskyCs=0.1), False)
    s.Step(None, 50)
    st = s.ReadbackStats()
    assert not st.hasNaN and st.forceX > 0.0
    s.Shutdown()


def test_mesh_to_solver_on_the_device():
    """fc_voxelize_device + fc_lbm_set_geometry_device: the grid never leaves HBM, and a
    solver set up that way steps bit for bit like one set up from the host grid."""
    import torch
    from flowcraft_amd.lbm import LatticeModel, LbmSolver, Params
    from flowcraft_amd.voxelizer import Params as VParams, Voxelizer
    pos, idx = meshes.wing(0.04, 0.4, 0.15, 100.0, 300.0, (0.0, 0.0, 0.0), "x", 8.0)
    vp = VParams(96)
    host = Voxelizer.Voxelize(pos, idx, vp)
    dims, cs, org, dsolid, dsdf = Voxelizer.VoxelizeDevice(pos, idx, vp)
    assert dims == (host.nx, host.ny, host.nz)
    assert np.array_equal(dsolid.cpu().numpy(), host.solid)
    assert np.array_equal(dsdf.cpu().numpy().view(np.uint32), host.signedDistance.view(np.uint32))
    p = Params(tau=0.6, inletVelocity=0.05, smagorinskyCs=0.12)
    a, b = LbmSolver(), LbmSolver()
    assert a.Init(None, host, LatticeModel.D3Q19) and b.Init(None, host, LatticeModel.D3Q19)
    b.SetGeometryDevice(dsolid, dsdf)
    # the solver keeps its own copy: reusing the caller's buffers changes neither the mask
    # SolidMask() reports nor the run
    dsolid.fill_(1)
    dsdf.fill_(-1.0)
    torch.cuda.synchronize()
    assert np.array_equal(b.SolidMask(), host.solid.reshape(-1))
    for e in (a, b):
        e.Reset(p, True)
        e.Step(None, 20)
    assert np.array_equal(a.ReadbackField().view(np.uint32), b.ReadbackField().view(np.uint32))
    # and the aerodynamic summary (which reads a host copy of the mask made on first use)
    # agrees between the two set-ups
    fa, fb = a.ComputeAerodynamicForces(cs, org), b.ComputeAerodynamicForces(cs, org)
    assert fa == fb
    a.Shutdown()
    b.Shutdown()

