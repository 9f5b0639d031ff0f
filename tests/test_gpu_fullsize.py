"""Full-size (2048 x 4096, tide) checks on the GPU, where the CPU oracle is too slow:

* EXACT and PARTITION solver families agree to the 1e-10 bar after 100 steps (EXACT is
  the reference's own recurrences and is pinned bit for bit to the oracle at small sizes
  in test_gpu_parity);
* FUSED against the oracle over a few steps (the oracle's OpenMP build takes seconds for
  them);
* mass budget: sum(dx_x * dy * dZs) equals dt*dy*(sum of offshore inflow Qx[0] minus the
  periodic-free lateral terms, which cancel) step by step, i.e. the step conserves
  volume up to the boundary flux (src/Hydro.cpp:253-271 with src/HydroMeshOperations.cpp
  :99-114 and :148-159).
"""
from __future__ import annotations

import numpy as np
import pytest

from morfo70_amd import _lib
from morfo70_amd import scenarios as S
from tests.helpers import compare_states

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2():
    return S.seeded_beach(2048, 4096, 2.0, 2.0, 2, 200, tide=True, name="beach_tide_2048x4096_s2")


def test_partition_vs_exact_full_size(c2):
    from morfo70_amd.hydro import from_scenario
    print("[fullsize] 2048x4096 tide, 100 steps, exact vs partition", flush=True)
    ge = from_scenario(c2, solver=_lib.SOLVER_EXACT)
    ge.Run(100)
    se = ge.state()
    ge.close()
    gp = from_scenario(c2, solver=_lib.SOLVER_PARTITION)
    gp.Run(100)
    sp = gp.state()
    gp.close()
    res = compare_states(sp, se)
    print({k: v["rel"] for k, v in res.items()}, flush=True)
    for k, r in res.items():
        assert r["nan"] == 0 and r["ref_nan"] == 0, k
        assert r["rel"] <= 1e-10, (k, r)


def test_fused_vs_exact_full_size(c2):
    """FUSED (the default family) against EXACT at the 1e-10 bar after 100 steps."""
    from morfo70_amd.hydro import from_scenario
    out = []
    for solver in (_lib.SOLVER_EXACT, _lib.SOLVER_FUSED):
        g = from_scenario(c2, solver=solver)
        g.Run(100)
        out.append(g.state())
        g.close()
    res = compare_states(out[1], out[0])
    print({k: v["rel"] for k, v in res.items()}, flush=True)
    for k, r in res.items():
        assert r["nan"] == 0 and r["rel"] <= 1e-10, (k, r)


def test_fused_vs_oracle_full_size(c2):
    """FUSED against the oracle itself at full size, over two Runs of three steps: the paths
    only this size takes (the Qx tail's multi-row runs, the derived V_c of a Run's later
    steps, the constant-drag sweeps, 512-thread row workgroups) at the 1e-10 bar."""
    import oracle as orc
    from morfo70_amd.hydro import from_scenario
    print("[fullsize] 2048x4096 tide, 2 x Run(3), fused vs oracle", flush=True)
    o = orc.OracleHydro(c2, threads=16)
    g = from_scenario(c2, solver=_lib.SOLVER_FUSED)
    for _ in range(2):
        o.run(3)
        g.Run(3)
    res = compare_states(g.state(), o.get_state())
    g.close()
    print({k: v["rel"] for k, v in res.items()}, flush=True)
    for k, r in res.items():
        assert r["nan"] == r["ref_nan"] == 0 and r["rel"] <= 1e-10, (k, r)


def test_mass_budget_full_size(c2):
    """Volume change over one step equals dt * (inflow through the offshore face):
    sum_i,j dx_x[i]*dy*(Zs_new - Zs_old) = -dt*dy*sum_j sum_i (Qx[i+1]-Qx[i] + dx_x[i]*dQy_dy)
    = dt*dy*sum_j (Qx[0] - Qx[nx]) since the alongshore differences telescope to zero."""
    from morfo70_amd.hydro import from_scenario
    g = from_scenario(c2, solver=_lib.SOLVER_FUSED)
    g.Run(20)
    dxx = np.concatenate([[c2.x[1] - c2.x[0]], 0.5 * (np.diff(c2.x)[1:] + np.diff(c2.x)[:-1]),
                          [c2.x[-1] - c2.x[-2]]])
    before = g.state()
    g.Run(1)
    after = g.state()
    vol = float(np.sum(dxx[:, None] * c2.dy * (after["Zs"] - before["Zs"])))
    flux = c2.physics.dt * c2.dy * float(np.sum(after["Qx"][0] - after["Qx"][-1]))
    scale = c2.physics.dt * c2.dy * float(np.sum(np.abs(after["Qx"][0]))) + 1e-30
    print(f"[mass] dV={vol:.6e} flux={flux:.6e} scale={scale:.3e}", flush=True)
    assert abs(vol - flux) <= 1e-9 * max(scale, abs(vol)) + 1e-9


@pytest.fixture(scope="module")
def c2_2000():
    return S.config2(steps=2000)


def test_fused_vs_exact_2000_steps(c2_2000):
    """The plain workload as configured: 2048 x 4096, tide, 2000 steps. FUSED against
    EXACT (the reference's own recurrences, equal to the oracle bit for bit where no
    transcendental function enters, and to 1e-10 here) at the 1e-10 bar."""
    from morfo70_amd.hydro import from_scenario
    print("[fullsize] 2048x4096 tide, Run(2000), fused vs exact", flush=True)
    out = []
    for solver in (_lib.SOLVER_EXACT, _lib.SOLVER_FUSED):
        g = from_scenario(c2_2000, solver=solver)
        g.Run(2000)
        out.append(g.state())
        g.close()
        print(f"[fullsize] solver {solver} done", flush=True)
    res = compare_states(out[1], out[0])
    print({k: v["rel"] for k, v in res.items()}, flush=True)
    for k, r in res.items():
        assert r["nan"] == 0 and r["rel"] <= 1e-10, (k, r)


def test_fused_vs_oracle_200_steps(c2_2000):
    """FUSED against the oracle over the first 200 steps of the configured workload (the
    oracle's OpenMP build takes about 10 s for them on 16 host threads)."""
    import oracle as orc
    from morfo70_amd.hydro import from_scenario
    threads = orc.host_cpus()["threads"]
    print(f"[fullsize] 2048x4096 tide, Run(200), fused vs oracle on {threads} threads", flush=True)
    o = orc.OracleHydro(c2_2000, threads=threads)
    g = from_scenario(c2_2000, solver=_lib.SOLVER_FUSED)
    g.Run(200)
    o.run(200)
    res = compare_states(g.state(), o.get_state())
    g.close()
    print({k: v["rel"] for k, v in res.items()}, flush=True)
    for k, r in res.items():
        assert r["nan"] == r["ref_nan"] == 0 and r["rel"] <= 1e-10, (k, r)
