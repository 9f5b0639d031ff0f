"""GPU parity: the HIP Hydro step (through the C ABI) against the CPU oracle.

Bar: the per-step arithmetic (products, sums, IEEE division and sqrt, masks) is the
reference's, so with drag/viscosity models free of transcendental functions the two
agree bit for bit. Models that call cbrt/log once per Run (ManningStrickler, LogDepth,
Breaking* viscosity) go through the device math library there, which may differ from
glibc by an ulp; for those the bar is the stated tolerance, max-norm relative error
<= 1e-10 per field after N steps (and the masks must agree, which a 1-ulp flip would
break visibly far above 1e-10).
"""
from __future__ import annotations

import numpy as np
import pytest

from morfo70_amd import _lib
from morfo70_amd import scenarios as S
from tests.helpers import compare_states, max_rel, nonuniform_beach

pytestmark = pytest.mark.gpu

TOL = 1e-10


def run_both(sc, steps, runs=1, solver=0):
    import oracle as orc
    from morfo70_amd.hydro import from_scenario
    o = orc.OracleHydro(sc)
    g = from_scenario(sc, solver=solver)
    print(f"[parity] {sc.name} nx={sc.nx} ny={sc.ny} steps={steps}x{runs} solver={solver}", flush=True)
    for _ in range(runs):
        o.run(steps)
        g.Run(steps)
    return g.state(), o.get_state()


def assert_parity(got, ref, exact: bool):
    res = compare_states(got, ref)
    for k, r in res.items():
        assert r["nan"] == r["ref_nan"] == 0, (k, r)
        if exact:
            assert r["exact"], (k, r)
        else:
            assert r["rel"] <= TOL, (k, r)
    return res


@pytest.mark.parametrize("steps", [1, 7])
def test_exact_constant_models(steps):
    ph = S.Physics(drag_model=S.DRAG["constant"], visc_model=S.VISC["constant"], dt=0.25)
    sc = nonuniform_beach(physics=ph)
    got, ref = run_both(sc, steps)
    assert_parity(got, ref, exact=True)


@pytest.mark.parametrize("nx,ny", [(3, 3), (4, 5), (5, 17), (67, 70), (130, 129)])
def test_exact_shapes(nx, ny):
    """Edge shapes: the n=3 periodic solve, chunk tails of both solvers, line counts that
    are not multiples of the 64-lane wavefront."""
    ph = S.Physics(drag_model=S.DRAG["constant"], visc_model=S.VISC["constant"], dt=0.2)
    sc = nonuniform_beach(nx=nx, ny=ny, physics=ph, seed=nx + ny)
    got, ref = run_both(sc, 4)
    assert_parity(got, ref, exact=True)


@pytest.mark.parametrize("drag", ["constant", "logdepth", "manning"])
@pytest.mark.parametrize("visc", ["constant", "rollers", "depth", "hrms"])
def test_models(drag, visc):
    ph = S.Physics(drag_model=S.DRAG[drag], visc_model=S.VISC[visc], dt=0.25)
    sc = nonuniform_beach(physics=ph, seed=3)
    if visc == "rollers":
        sc.forcing["D_r"] = 0.3 * sc.forcing["Dw"]
    got, ref = run_both(sc, 10)
    assert_parity(got, ref, exact=(drag == "constant" and visc == "constant"))


def test_tide_series_and_multiple_runs():
    """ConstantZs offshore BC with a time-varying f0 (tide hook), over three Run calls
    (drag/viscosity/forcing are frozen per Run, src/Hydro.cpp:175-192)."""
    ph = S.Physics(offshore_bc=S.OFFSHORE_BC["constant_zs"], drag_model=S.DRAG["constant"],
                   visc_model=S.VISC["constant"], dt=0.25)
    sc = nonuniform_beach(physics=ph, tide=True, steps=30, seed=5)
    got, ref = run_both(sc, 10, runs=3)
    assert_parity(got, ref, exact=True)


def test_config1_1000_steps():
    """The 120 x 200 bar beach (Chezy 65, BreakingHrms viscosity, sponge), 1000 steps."""
    sc = S.config1(steps=1000)
    got, ref = run_both(sc, 1000)
    res = assert_parity(got, ref, exact=False)
    print({k: (v["rel"], v["mismatch"]) for k, v in res.items()})


@pytest.mark.parametrize("solver", [1, 2])
@pytest.mark.parametrize("case", ["models", "shapes", "odd", "tide", "config1"])
def test_partition_solver_parity(case, solver):
    """The PARTITION (1) and FUSED (2) line solvers against the oracle at the 1e-10 bar.
    "odd": an odd alongshore count, which takes the one-element paths of the row kernels
    and the one-column tails (even counts take the column-pair paths)."""
    if case == "models":
        ph = S.Physics(drag_model=S.DRAG["logdepth"], visc_model=S.VISC["hrms"], dt=0.25)
        sc, steps, runs = nonuniform_beach(physics=ph, seed=11), 25, 2
    elif case == "shapes":
        sc, steps, runs = nonuniform_beach(nx=67, ny=70, seed=4), 20, 1
    elif case == "odd":
        sc, steps, runs = nonuniform_beach(nx=31, ny=37, seed=9), 20, 1
    elif case == "tide":
        ph = S.Physics(offshore_bc=S.OFFSHORE_BC["constant_zs"], dt=0.25)
        sc, steps, runs = nonuniform_beach(physics=ph, tide=True, steps=40, seed=6), 20, 2
    else:
        sc, steps, runs = S.config1(steps=1000), 1000, 1
    got, ref = run_both(sc, steps, runs=runs, solver=solver)
    res = assert_parity(got, ref, exact=False)
    print({k: v["rel"] for k, v in res.items()})


@pytest.mark.parametrize("solver", [1, 2])
@pytest.mark.parametrize("nx,ny", [(3000, 16), (5000, 12), (16, 6000), (12, 8192)])
def test_long_lines_parity(nx, ny, solver):
    """Lines longer than 2048 and 4096 against the oracle at the 1e-10 bar: columns of
    2999 and 4999 unknowns (PARTITION with fewer than four columns per workgroup, FUSED's
    reduced systems of 750 and 1250 rows) and rows of 6000 and 8192 (the 1024-thread LDS
    row kernel)."""
    ph = S.Physics(drag_model=S.DRAG["constant"], visc_model=S.VISC["constant"], dt=0.25)
    sc = nonuniform_beach(nx=nx, ny=ny, physics=ph, seed=nx + ny)
    got, ref = run_both(sc, 5, solver=solver)
    assert_parity(got, ref, exact=False)
    assert np.max(np.abs(got["Qx"])) > 0


def test_lake_at_rest_matches_oracle_and_stays_at_rest():
    """Still water, zero forcing: Zs = D - Z + h is zero only to rounding, so the step
    moves it at the 1e-16 level; the GPU must match the oracle bit for bit and the
    flow must stay at roundoff."""
    sc = S.lake_at_rest()
    got, ref = run_both(sc, 20)
    assert_parity(got, ref, exact=True)
    for k in ("Qx", "Qy", "U", "V"):
        assert float(np.max(np.abs(got[k]))) < 1e-12, k
    assert float(np.max(np.abs(got["Zs"][got["D"] > 0.0101]))) < 1e-12


def test_state_roundtrip_and_finite_check():
    sc = nonuniform_beach()
    from morfo70_amd.hydro import from_scenario
    g = from_scenario(sc)
    g.Run(3)
    st = {k: v.copy() for k, v in g.state().items()}
    assert g.check_finite()
    g.Qx[5, 3] = np.nan
    g.push()
    assert not g.check_finite()
    g.Qx[5, 3] = st["Qx"][5, 3]
    g.push()
    assert g.check_finite()
    st2 = g.state()
    for k in st:
        assert np.array_equal(st[k], st2[k]), k


def test_profiling_one_kernel_leaves_results_unchanged():
    """m70h_set_profiling_only times the launches of one kernel and none of the others;
    the events do not change the step's results."""
    sc = nonuniform_beach(nx=40, ny=36, seed=5)
    from morfo70_amd import _lib
    from morfo70_amd.hydro import from_scenario
    print("[profiling] one kernel, FUSED 40x36", flush=True)
    a = from_scenario(sc, solver=_lib.SOLVER_FUSED)
    b = from_scenario(sc, solver=_lib.SOLVER_FUSED)
    a.set_profiling(True, only="qx_tail")
    a.Run(4)
    b.Run(4)
    prof = a.profile()
    assert prof["qx_tail"][1] == 4
    assert all(n == 0 for k, (ms, n) in prof.items() if k != "qx_tail")
    a.set_profiling(True)
    a.Run(1)
    assert sum(n for ms, n in a.profile().values()) > 1
    a.set_profiling(False)
    b.Run(1)
    with pytest.raises(RuntimeError):
        a.set_profiling(True, only="no_such_kernel")
    sa, sb = a.state(), b.state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k


@pytest.mark.parametrize("solver", [0, 2])
def test_run_starts_from_the_callers_derived_fields(solver):
    """A Run's first step takes U_c, V_c, dQx_dx and dQy_dy as the caller left them, also
    where they do not match U, V, Qx and Qy: the FUSED steps after a Run's first form V_c
    and U_c from V and U, and store them only on the Run's last step. Against the oracle given the same perturbed
    state, over two Runs: bit for bit for EXACT, at the 1e-10 bar for FUSED (whose column
    elimination rounds differently by a few ulps). The perturbation moves the oracle's result
    far above that bar, so a step that ignored the caller's V_c would fail."""
    ph = S.Physics(drag_model=S.DRAG["constant"], visc_model=S.VISC["constant"], dt=0.25)
    sc = nonuniform_beach(nx=40, ny=36, physics=ph, seed=6)
    import oracle as orc
    from morfo70_amd.hydro import STATE, from_scenario
    o = orc.OracleHydro(sc)
    g = from_scenario(sc, solver=solver)
    print(f"[parity] perturbed U_c/V_c between Runs, solver={solver}", flush=True)
    o.run(2)
    g.Run(2)
    st = o.get_state()
    rng = np.random.default_rng(3)
    wet = st["D"] > 0.05
    for k in ("U_c", "V_c", "dQx_dx", "dQy_dy"):
        st[k] = st[k] + 1e-4 * rng.standard_normal(st[k].shape) * wet
    ctl = orc.OracleHydro(sc)
    ctl.set_state(o.get_state())
    o.set_state(st)
    for k in STATE:
        tgt = g.bathy.D if k == "D" else getattr(g, k)
        tgt[...] = st[k]
    g.push()
    for _ in range(2):
        o.run(3)
        g.Run(3)
        assert_parity(g.state(), o.get_state(), exact=solver == 0)
    for _ in range(2):
        ctl.run(3)
    assert max_rel(ctl.get_state()["Qx"], o.get_state()["Qx"]) > 1e-7


def test_run_reports_a_nonfinite_state():
    """Urms = 0 with the fluxes at rest makes the drag term 0/0 (src/HydroParametrizations.cpp:25),
    which the reference lets pass silently. m70h_run returns M70H_ENONFINITE; the state,
    NaN included, still reaches the host mirrors; a clean Run afterwards reports success."""
    from morfo70_amd.hydro import from_scenario
    sc = nonuniform_beach(nx=30, ny=24, steps=4)
    g = from_scenario(sc, solver=_lib.SOLVER_FUSED)
    g.Run(2)                                   # clean
    urms = g.wave.Urms.copy()
    g.wave.Urms[10:14, 5:9] = 0.0
    g.Qx[...] = 0.0
    g.Qy[...] = 0.0
    g.U[...] = 0.0
    g.V[...] = 0.0
    g.U_c[...] = 0.0
    g.V_c[...] = 0.0
    with pytest.raises(_lib.M70HError) as ei:
        g.Run(3)
    assert ei.value.code == _lib.M70H_ENONFINITE
    assert np.isnan(g.Qx).any()
    g2 = from_scenario(sc, solver=_lib.SOLVER_EXACT)
    g2.wave.Urms[...] = urms
    g2.Run(3)                                  # no exception
    assert np.isfinite(g2.Qx).all()
    g.close()
    g2.close()
