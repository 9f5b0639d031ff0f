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
