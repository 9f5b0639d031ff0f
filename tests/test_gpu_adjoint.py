"""GPU adjoint of Run against the tangent-linear oracle (dot-product tests) and against
finite differences.

For the Jacobian J of Run(n) (initial state and parameters -> final state), the adjoint
must satisfy <w, J v> = <J^T w, v> for any tangent v and seed w. J v comes from the
dual-number oracle (oracle/tlm.cpp, exact derivatives of the discrete CPU step), J^T w
from the GPU kernels through the C ABI. Testing one input field against one output
field at a time localises a wrong term to one block of J.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from morfo70_amd import _lib
from morfo70_amd import scenarios as S
from morfo70_amd.hydro import STATE, from_scenario
from tests.helpers import nonuniform_beach

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

PLUS = ("Qx", "Qy", "U", "V")


def _shape(sc, k):
    return (sc.nx + 1, sc.ny) if k in PLUS else (sc.nx, sc.ny)


def _lin_point(sc, warm):
    o = oracle.OracleHydro(sc)
    o.run(warm)
    return o.get_state()


def _gpu(sc, st, solver=_lib.SOLVER_EXACT):
    hy = from_scenario(sc, solver=solver)
    for k in STATE:
        if k == "D":
            hy.bathy.D[...] = st["D"]
        else:
            setattr(hy, k, np.ascontiguousarray(st[k]))
    hy.push()
    hy.set_forcing()
    return hy


def _adjoint(sc, st, n, seed: dict, slots):
    from morfo70_amd.adjoint import RunAdjoint
    hy = _gpu(sc, st)
    adj = RunAdjoint(hy)
    adj.forward(n, slots)
    sd = {k: torch.as_tensor(np.ascontiguousarray(v), device="cuda") for k, v in seed.items()}
    grads, dp = adj.reverse(sd, want=STATE)
    out = {k: v.cpu().numpy() for k, v in grads.items()}
    stats = adj.stats()
    adj.release()
    hy.close()
    return out, dp, stats


def _tlm(sc, st, n, tangents: dict, dparams=None):
    t = oracle.TangentHydro(sc, dparams=dparams)
    t.set_state(st, tangents)
    t.run(n)
    return t.get_state()[1]


def _rel(a, b):
    return abs(a - b) / max(abs(a), abs(b), 1e-300)


@pytest.mark.parametrize("n,slots", [(1, 1), (4, 2)])
def test_dot_product_by_field_blocks(n, slots):
    sc = nonuniform_beach(nx=24, ny=16, seed=5, tide=True, steps=40)
    st = _lin_point(sc, 6)
    rng = np.random.default_rng(11)
    seeds = {k: rng.normal(size=_shape(sc, k)) for k in STATE}
    tans = {k: rng.normal(size=_shape(sc, k)) for k in STATE}
    # adjoint of every output field separately
    adj = {}
    for ko in STATE:
        g, dp, _ = _adjoint(sc, st, n, {ko: seeds[ko]}, slots)
        adj[ko] = (g, dp)
    bad, nonzero = [], 0

    def check(ki, ko, lhs, rhs):
        nonlocal nonzero
        if max(abs(lhs), abs(rhs)) > 1e-9:
            nonzero += 1
            if _rel(lhs, rhs) > 1e-10:
                bad.append((ki, ko, lhs, rhs))
        elif max(abs(lhs), abs(rhs)) > 0 and abs(lhs - rhs) > 1e-12:
            bad.append((ki, ko, lhs, rhs))

    for ki in STATE:
        jv = _tlm(sc, st, n, {ki: tans[ki]})
        for ko in STATE:
            check(ki, ko, float(np.sum(seeds[ko] * jv[ko])), float(np.sum(adj[ko][0][ki] * tans[ki])))
    for p in ("cD0", "nu0", "M"):
        jv = _tlm(sc, st, n, {}, dparams={p: 1.0})
        for ko in STATE:
            check(p, ko, float(np.sum(seeds[ko] * jv[ko])), adj[ko][1][p])
    assert not bad, "\n".join(f"d{ko}/d{ki}: tlm {l:.12e} adjoint {r:.12e}" for ki, ko, l, r in bad)
    # U_c and dQx_dx enter the step only through wet/dry masks; everything else couples
    assert nonzero >= 100, nonzero


def test_checkpoint_schedules_agree():
    """The binomial schedule only changes what is recomputed: gradients for 1, 3 and n
    buffers agree (to atomics' summation order)."""
    sc = nonuniform_beach(nx=30, ny=20, seed=7)
    st = _lin_point(sc, 4)
    n = 12
    seed = {"Zs": np.random.default_rng(2).normal(size=(sc.nx, sc.ny))}
    ref = None
    for slots in (n, 3, 1):
        g, dp, stats = _adjoint(sc, st, n, seed, slots)
        assert stats["taped_steps"] == n
        if ref is None:
            ref = (g, dp)
            assert stats["forward_steps"] == 0
            continue
        assert stats["forward_steps"] > 0
        for k in STATE:
            scale = max(np.max(np.abs(ref[0][k])), 1e-300)
            assert np.max(np.abs(g[k] - ref[0][k])) <= 1e-12 * scale, (slots, k)
        assert _rel(dp["cD0"], ref[1]["cD0"]) < 1e-12
