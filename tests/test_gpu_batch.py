"""Batched step (grouped launches, hipGraph replay) against each handle run alone."""
from __future__ import annotations

import numpy as np
import pytest

from morfo70_amd import _lib
from morfo70_amd import scenarios as S

pytestmark = pytest.mark.gpu

SIZES = [(130, 260), (67, 300), (200, 129), (96, 512), (257, 257)]


def make(seed_base=50):
    return [S.seeded_beach(nx, ny, 2.0, 2.0, seed_base + q, 20, random_waves=True)
            for q, (nx, ny) in enumerate(SIZES)]


@pytest.mark.parametrize("graph", [True, False])
@pytest.mark.parametrize("steps", [7, 12])
def test_batch_bitwise_equals_individual(graph, steps):
    from morfo70_amd.batch import Batch
    from morfo70_amd.hydro import from_scenario
    scs = make()
    print(f"[batch] {len(scs)} scenarios, {steps} steps, graph={graph}", flush=True)
    ref = []
    for sc in scs:
        g = from_scenario(sc, solver=_lib.SOLVER_FUSED)
        g.Run(steps)
        ref.append(g.state())
        g.close()
    hs = [from_scenario(sc, solver=_lib.SOLVER_FUSED) for sc in scs]
    for h in hs:
        h.push()
        h.set_forcing()
    b = Batch(hs, graph=graph)
    b.run(steps)
    for q, h in enumerate(hs):
        st = h.state()
        for k in st:
            assert np.array_equal(st[k], ref[q][k]), (q, k)
    # a second Run continues from the advanced state
    b.run(3)
    for h in hs:
        assert h.check_finite()
    b.close()


def test_batch_vs_oracle():
    import oracle
    from morfo70_amd.batch import Batch
    from morfo70_amd.hydro import from_scenario
    scs = make(80)[:3]
    hs = [from_scenario(sc, solver=_lib.SOLVER_FUSED) for sc in scs]
    for h in hs:
        h.push()
        h.set_forcing()
    Batch(hs).run(15)
    from tests.helpers import compare_states
    for sc, h in zip(scs, hs):
        o = oracle.OracleHydro(sc)
        o.run(15)
        res = compare_states(h.state(), o.get_state())
        for k, r in res.items():
            assert r["nan"] == 0 and r["rel"] <= 1e-10, (sc.name, k, r)


def _individual(scs, runs, rows_block=0):
    """Each scenario alone on a FUSED handle, Run(n) for n in runs; final host states."""
    from morfo70_amd.hydro import from_scenario
    out = []
    for sc in scs:
        g = from_scenario(sc, solver=_lib.SOLVER_FUSED)
        if rows_block:
            g._chk(g._L.m70h_set_rows_block(g.handle, rows_block))
        for n in runs:
            g.Run(n)
        out.append(g.state())
        g.close()
    return out


def test_batch_applies_each_scenarios_tide_series():
    """A scenario with a time series of the offshore offset (tide) advances in a batch
    exactly as alone: its step k reads series value f0_pos + k on the device, across two
    batch runs and past the end of a short series (the last value repeats)."""
    from morfo70_amd.batch import Batch
    from morfo70_amd.hydro import from_scenario
    scs = [S.seeded_beach(130, 260, 2.0, 2.0, 61, 40, tide=True),
           S.seeded_beach(96, 512, 2.0, 2.0, 62, 40, random_waves=True),
           S.seeded_beach(201, 300, 2.0, 2.0, 63, 40, tide=True)]
    scs[2].f0_series = scs[2].f0_series[:9] * 40.0   # shorter than the 12 steps run
    assert np.ptp(scs[0].f0_series[:12]) > 0
    print("[batch] tide series, 3 scenarios, Run(7) + Run(5)", flush=True)
    ref = _individual(scs, (7, 5))
    hs = [from_scenario(sc, solver=_lib.SOLVER_FUSED) for sc in scs]
    for h in hs:
        h.push()
        h.set_forcing()
    b = Batch(hs, graph=True)
    b.run(7)
    b.run(5)
    for q, h in enumerate(hs):
        st = h.state()
        for k in st:
            assert np.array_equal(st[k], ref[q][k]), (q, k)
    b.close()


def test_batch64_scenarios_bitwise_equal_individual_runs():
    """The 64 uneven scenarios of the batched workload (nx U[128,1024], ny U[256,2048],
    seeded bathymetry and waves) advanced in one batch, under the schedule the autotuner
    picks on it (256-thread rows, graph replay) and then eagerly, equal each scenario run
    alone bit for bit."""
    from morfo70_amd.batch import Batch
    from morfo70_amd.hydro import from_scenario
    scs = S.batch_scenarios(64, steps=3)
    print(f"[batch64] {len(scs)} scenarios, {sum(s.cells for s in scs) / 1e6:.1f} M cells", flush=True)
    ref = _individual(scs, (2, 1))
    hs = [from_scenario(sc, solver=_lib.SOLVER_FUSED) for sc in scs]
    for h in hs:
        h.push()
        h.set_forcing()
    b = Batch(hs, graph=True)
    b.set_rows_block(256)
    b.run(2)
    b.set_graph(False)
    b.run(1)
    bad = []
    for q, h in enumerate(hs):
        st = h.state()
        bad += [(q, k) for k in st if not np.array_equal(st[k], ref[q][k])]
        h.close()
    b.close()
    assert not bad, bad[:10]
