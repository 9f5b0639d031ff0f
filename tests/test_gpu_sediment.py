"""GPU parity of sediment transport and the bed update (class Sediment, src/Sediment.cpp)
through the C ABI (m70s_*) against the oracle.

Bar: the flux divergence and the Adams-Bashforth bed update are products, sums and IEEE
divisions in the reference's order, so given the same fluxes they agree bit for bit. The
coefficients call tanh three times per cell (src/Sediment.cpp:194-250), which the device
library may round differently from glibc by an ulp; there the bar is max-norm relative
error <= 1e-10 per field, also after several coupled hydro/sediment outer iterations.
"""
from __future__ import annotations

import numpy as np
import pytest

from morfo70_amd import scenarios as S
from tests.helpers import compare_states, max_rel, nonuniform_beach

pytestmark = pytest.mark.gpu

TOL = 1e-10
SED_FIELDS = ("Alpha", "gamma", "Gamma", "qx", "qy", "dh_dt", "h", "Zb")


def _pair(sc, sp=None):
    import oracle as orc
    from morfo70_amd.hydro import from_scenario
    from morfo70_amd.sediment import Sediment
    sp = sp or S.SedimentParams()
    o = orc.OracleHydro(sc)
    g = from_scenario(sc)
    return o, orc.OracleSediment(o, sp), g, Sediment(g, params=sp)


def _sed_state(gs):
    out = {k: getattr(gs, k) for k in SED_FIELDS[:6]}
    out["h"], out["Zb"] = gs.bathy.h, gs.bathy.Zb
    return out


def _assert_sed(gs, osd, tol=TOL, exact=()):
    got = _sed_state(gs)
    for k in SED_FIELDS:
        r = osd.get(k)
        assert np.all(np.isfinite(got[k])), k
        if k in exact:
            assert np.array_equal(got[k], r), (k, max_rel(got[k], r))
        else:
            assert max_rel(got[k], r) <= tol, (k, max_rel(got[k], r))


@pytest.mark.parametrize("nx,ny,outer", [(24, 20, 4), (67, 70, 3), (130, 129, 2)])
def test_coupled_outer_loop(nx, ny, outer):
    """The reference's outer loop (src/Morfo70.cpp:141-163): Run, fluxes, bed update,
    repeated; the bed update feeds the next Run through Zs. dt = 2 s so the bed moves
    well above rounding."""
    ph = S.Physics(drag_model=S.DRAG["constant"], visc_model=S.VISC["constant"], dt=0.25)
    sc = nonuniform_beach(nx=nx, ny=ny, physics=ph, seed=nx)
    o, osd, g, gs = _pair(sc)
    print(f"[sediment] coupled {nx}x{ny} outer={outer}", flush=True)
    for it in range(outer):
        o.run(6)
        g.Run(6)
        osd.fluxes()
        gs.calculateSedimentFluxes()
        osd.update_bottom(2.0)
        gs.updateBottom(2.0)
        _assert_sed(gs, osd)
        res = compare_states(g.state(), o.get_state())
        for k, r in res.items():
            assert r["nan"] == 0 and r["rel"] <= TOL, (it, k, r)
    assert np.max(np.abs(gs.dh_dt)) > 0


def test_given_fluxes_bit_exact():
    """setSedimentFluxesFromResumeFile's arithmetic (:77-103) and the bed update on given
    fluxes: no transcendental functions, so bit for bit; row 0 of the bed never moves.
    Constant drag and viscosity keep the hydro state bit-exact too."""
    ph = S.Physics(drag_model=S.DRAG["constant"], visc_model=S.VISC["constant"], dt=0.25)
    sc = nonuniform_beach(nx=40, ny=33, seed=2, physics=ph)
    o, osd, g, gs = _pair(sc)
    o.run(3)
    g.Run(3)
    rng = np.random.default_rng(11)
    print("[sediment] given fluxes 40x33", flush=True)
    for it in range(3):
        qx = rng.normal(0, 1e-3, (sc.nx, sc.ny))
        qy = rng.normal(0, 1e-3, (sc.nx, sc.ny))
        osd.set_fluxes(qx, qy)
        gs.setSedimentFluxes(qx, qy)
        assert np.array_equal(gs.dh_dt, osd.get("dh_dt"))
        h0 = gs.bathy.h.copy()
        osd.update_bottom(0.7)
        gs.updateBottom(0.7)
        for k in ("dh_dt", "h", "Zb"):
            assert np.array_equal(_sed_state(gs)[k], osd.get(k)), (it, k)
        assert np.array_equal(g.state()["Zs"], o.get_state()["Zs"]), it
        assert np.array_equal(gs.bathy.h[0], h0[0])


def test_running_max_of_dissipation():
    """maxDw is a running maximum over every flux call (:252-262): halving Dw leaves the
    swash coefficient Gamma unchanged where it is non-zero."""
    sc = nonuniform_beach(nx=30, ny=24, seed=5)
    o, osd, g, gs = _pair(sc)
    o.run(4)
    g.Run(4)
    osd.fluxes()
    gs.calculateSedimentFluxes()
    G1 = gs.Gamma.copy()
    half = dict(sc.forcing, Dw=np.ascontiguousarray(0.5 * sc.forcing["Dw"]))
    o.set_forcing(half)
    g.wave.Dw = half["Dw"]
    g.set_forcing()
    print("[sediment] running max 30x24", flush=True)
    osd.fluxes()
    gs.calculateSedimentFluxes()
    _assert_sed(gs, osd)
    assert np.array_equal(gs.Gamma, G1)


def test_parameters_and_dry_shoreline():
    """Non-default coefficients and a beach with a wide dry zone: the swash coefficient's
    dry branch |Zb + Zs_sh| (:241-266) and the shoreline scan (:264-283) over many dry rows."""
    ph = S.Physics(drag_model=S.DRAG["constant"], visc_model=S.VISC["constant"], dt=0.25)
    sc = nonuniform_beach(nx=48, ny=17, seed=9, physics=ph, Hrms=1.5)
    sp = S.SedimentParams(p=0.35, alpha0_max=0.004, alpha0_min=0.0005, D_alpha0_max=2.0, D_alpha0_min=6.0,
                          gamma0=0.8, Gamma0=1.7, A_swash=0.5)
    o, osd, g, gs = _pair(sc, sp)
    o.run(8)
    g.Run(8)
    print("[sediment] params/dry 48x17", flush=True)
    osd.fluxes()
    gs.calculateSedimentFluxes()
    _assert_sed(gs, osd)
    assert np.any(g.bathy.D <= ph.Dmin + 1e-4)


def test_fullsize_budget():
    """At the 2048 x 4096 tide workload: the flux divergence telescopes, so
    sum_ij dx_x[i] dy dh_dt = dy/(1-p) sum_j qx[0][j] (qx row nx-1 is zero, y periodic),
    and the bed update moves h, Zb and Zs by exactly the same increments."""
    import oracle as orc
    from morfo70_amd.hydro import from_scenario
    from morfo70_amd.sediment import Sediment
    sc = S.config2()
    g = from_scenario(sc, solver=2)
    sp = S.SedimentParams()
    gs = Sediment(g, params=sp)
    print("[sediment] fullsize budget 2048x4096", flush=True)
    g.Run(20)
    gs.calculateSedimentFluxes()
    zs0 = g.state()["Zs"].copy()
    h0, zb0 = gs.bathy.h.copy(), gs.bathy.Zb.copy()
    gs.updateBottom(0.1)
    dhdt, qx = gs.dh_dt, gs.qx
    assert np.all(np.isfinite(dhdt)) and np.all(qx[-1] == 0.0)
    dxx = orc.grid_metrics(sc.x)["dx_x"]
    lhs = np.sum(dxx[1:, None] * sc.dy * dhdt[1:])
    rhs = sc.dy / (1 - sp.p) * np.sum(qx[0])
    scale = np.sum(np.abs(dxx[1:, None] * sc.dy * dhdt[1:]))
    assert abs(lhs - rhs) <= 1e-11 * max(scale, 1e-300)
    # first update: the previous dh_dt is zero, so the increment is dt/2 * 3 dh_dt; the
    # differences below carry one rounding of the updated value each
    incr = 0.5 * 0.1 * (3 * dhdt)
    ulp = 4 * np.finfo(np.float64).eps * max(np.max(np.abs(h0)), np.max(np.abs(zb0)), np.max(np.abs(zs0)))
    assert np.array_equal(gs.bathy.h[0], h0[0])
    np.testing.assert_allclose(gs.bathy.h[1:] - h0[1:], incr[1:], rtol=0, atol=ulp)
    np.testing.assert_allclose(gs.bathy.Zb[1:] - zb0[1:], -incr[1:], rtol=0, atol=ulp)
    np.testing.assert_allclose(g.state()["Zs"][1:] - zs0[1:], incr[1:], rtol=0, atol=ulp)
    assert np.max(np.abs(incr)) > 1e3 * ulp
