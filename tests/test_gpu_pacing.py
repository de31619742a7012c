"""Interactive-front-end pacing (FlowCraft/main.cpp:2372-2555) on the engine:
StepPaced, FlowCraft's StepSolverPaced loop (adaptive z-chunks, AdvanceSubstep, periodic
slice health probes, cancellation), gives exactly the plain step's state; a field
snapshot taken asynchronously while the solver keeps stepping equals a blocking readback
of the same state."""
from __future__ import annotations

import numpy as np
import pytest

from flowcraft_amd.lbm import Layout, PacingState, Storage
from tests._cases import case_a_grid, case_a_params, make_engine

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _announce(request):
    print(f"\n[gpu-case] {request.node.nodeid}", flush=True)
    yield


def _pair(layout=Layout.PINGPONG, storage=Storage.F32):
    g = case_a_grid()
    a, b = make_engine(g, storage=storage, layout=layout), make_engine(g, storage=storage, layout=layout)
    for e in (a, b):
        e.Reset(case_a_params(), True)
    return a, b


@pytest.mark.parametrize("layout", [Layout.PINGPONG, Layout.INPLACE], ids=lambda l: l.name)
@pytest.mark.parametrize("target", [1e-6, 1e6], ids=["chunks4", "chunks256"])
def test_paced_steps_equal_plain_steps(layout, target):
    a, b = _pair(layout)
    st = PacingState()
    seen = []
    assert a.StepPaced(15, st, targetChunkMs=target, progress=seen.append) == 15
    b.Step(None, 15)
    assert np.array_equal(a.ReadbackField().view(np.uint32), b.ReadbackField().view(np.uint32))
    assert st.chunkZDepth == (4 if target < 1 else 256) and not st.diverged
    assert seen[0] == 0.0 and seen[-1] == 1.0 and all(x <= y for x, y in zip(seen, seen[1:]))
    a.Shutdown()
    b.Shutdown()


def test_cancel_between_chunks_keeps_the_last_full_substep():
    a, b = _pair()
    st = PacingState(chunkZDepth=4)
    calls = {"n": 0}

    def cancel():  # asked before each chunk and before each AdvanceSubstep
        calls["n"] += 1
        return calls["n"] > 3 * 7 + 2  # 6 chunks of 4 layers (+1) per substep: cancel inside substep 3

    assert a.StepPaced(10, st, cancel=cancel, targetChunkMs=1e-6) == 3
    b.Step(None, 3)
    assert np.array_equal(a.ReadbackField().view(np.uint32), b.ReadbackField().view(np.uint32))


@pytest.mark.parametrize("layout", [Layout.PINGPONG, Layout.INPLACE], ids=lambda l: l.name)
def test_cancel_raised_inside_a_substep(layout):
    """Cancel raised after the second chunk of substep 3: ping-pong drops the chunks
    already run (3 substeps); in place the substep is finished first (4), since its
    chunks have overwritten the state."""
    a, b = _pair(layout)
    st = PacingState(chunkZDepth=4)
    flag = {"on": False}

    def progress(f):
        if f > 3.3 / 10:
            flag["on"] = True

    done = a.StepPaced(10, st, cancel=lambda: flag["on"], targetChunkMs=1e-6, progress=progress)
    assert done == (3 if layout == Layout.PINGPONG else 4)
    b.Step(None, done)
    assert np.array_equal(a.ReadbackField().view(np.uint32), b.ReadbackField().view(np.uint32))


def test_probe_stops_a_diverged_run():
    """A non-finite rest population (it stays in its cell) planted on the host in the
    layer the first probe reads (nz/2): the paced loop returns after that substep with
    diverged set."""
    a, _ = _pair()
    pops = a.Populations()
    nz = a.Nz()
    pops[0, nz // 2, 12, 3] = np.nan
    a.SetPopulations(pops)
    st = PacingState()
    assert a.StepPaced(8, st) == 1 and st.diverged


@pytest.mark.parametrize("layout", [Layout.PINGPONG, Layout.INPLACE], ids=lambda l: l.name)
def test_snapshot_while_stepping_equals_blocking_readback(layout):
    a, b = _pair(layout)
    a.Step(None, 10)
    b.Step(None, 10)
    ref10 = b.ReadbackField()
    a.SnapshotBegin()
    a.Step(None, 25)  # runs while the snapshot is copied
    snap = a.SnapshotWait()
    assert np.array_equal(snap.view(np.uint32), ref10.view(np.uint32))
    b.Step(None, 25)
    assert np.array_equal(a.ReadbackField().view(np.uint32), b.ReadbackField().view(np.uint32))
    a.SnapshotBegin()  # a second snapshot into the same buffers
    while not a.SnapshotReady():
        pass
    assert np.array_equal(a.SnapshotWait(copy=False).view(np.uint32), b.ReadbackField().view(np.uint32))
