"""Pins of the wave oracle (oracle/wave_oracle.c, a restatement of class Wave/WaveRow of
src/Wave*.cpp): closed forms, the scheme's own discrete identities and symmetries. The
reference's wave sources cannot be compiled here (Netcdf.h through Grid.h) and its tests hold
no wave outputs, so these are what fixes the restatement."""
from __future__ import annotations

import math

import numpy as np
import pytest
from scipy.optimize import brentq

import oracle
from morfo70_amd import scenarios as S

G = 9.81


def _beach(nx=50, ny=24, dx=2.0, depth=(6.0, 0.4), bar=0.0, seed=0):
    x = np.arange(nx) * dx
    prof = np.linspace(depth[0], depth[1], nx)
    D = np.repeat(prof[:, None], ny, axis=1)
    if bar:
        rng = np.random.default_rng(seed)
        y = np.arange(ny)
        ph = rng.uniform(0, 2 * np.pi)
        D = D + bar * np.exp(-((np.arange(nx) - 0.6 * nx) / (0.1 * nx)) ** 2)[:, None] * \
            np.cos(2 * np.pi * y / ny + ph)[None, :]
    return x, np.ascontiguousarray(D)


def _run(x, D, dy, params, hOff=1.0, theta=0.0, T=8.0, U=None, V=None):
    w = oracle.OracleWave(x, D.shape[1], dy, params)
    w.set_depth(D)
    if U is not None:
        w.set_uv(U, V)
    n = w.calculate(hOff, theta, T)
    return w, n, w.fields()


@pytest.mark.parametrize("D,T", [(0.3, 6.0), (2.0, 8.0), (8.0, 12.0), (40.0, 5.0)])
def test_dispersion_relation(D, T):
    """WaveRow::dispersion (src/WaveParametrizations.cpp:14-60) solves w^2 = g k tanh(kD) with
    no current: k against an independent bracketing root, sigma = w, and the group/phase
    speed ratio n = 1/2 (1 + 2kD/sinh 2kD)."""
    w = 2 * math.pi / T
    r = oracle.wave_dispersion(D, 0.0, w)
    k_ref = brentq(lambda k: G * k * math.tanh(k * D) - w * w, 1e-9, 1e3, xtol=1e-15, rtol=1e-15)
    k = math.exp(r["lK"])
    assert abs(k - k_ref) <= 1e-12 * k_ref
    assert abs(r["sigma"] - w) <= 1e-13 * w
    assert abs(r["C"] - w / k) <= 1e-12 * r["C"]
    n = 0.5 * (1 + 2 * k * D / math.sinh(2 * k * D))
    assert abs(r["Cg"] / r["C"] - n) <= 1e-12


def test_dispersion_doppler():
    """With a current U_w along the ray the intrinsic frequency is Doppler shifted:
    sigma = w - k U_w and sigma^2 = g k tanh(kD)."""
    w = 2 * math.pi / 7.0
    for D, Uw in [(3.0, 0.4), (3.0, -0.4), (10.0, 1.0)]:
        r = oracle.wave_dispersion(D, Uw, w)
        k = math.exp(r["lK"])
        assert abs(r["sigma"] - (w - k * Uw)) <= 1e-12 * w
        assert abs(r["sigma"] ** 2 - G * k * math.tanh(k * D)) <= 1e-12 * r["sigma"] ** 2


@pytest.mark.parametrize("model", ["ThorntonGuza", "RegularWaves"])
def test_breaking_derivative(model):
    """The Newton-Raphson Jacobian uses d(Dw/E)/dH; for Thornton-Guza and the regular-wave
    model (src/WaveParametrizations.cpp:206-302) it is the exact derivative: check against a
    central difference."""
    p = S.WaveParams(breaking_model=S.BREAKING[model], m_mon=10.0, n_mon=10.0)
    for H, D in [(0.5, 1.0), (1.0, 2.5), (0.3, 0.35)]:
        sigma = 0.9
        dw, ddw = oracle.wave_breaking(p, H, sigma, D)
        e = 1e-6 * H
        fd = (oracle.wave_breaking(p, H + e, sigma, D)[0] - oracle.wave_breaking(p, H - e, sigma, D)[0]) / (2 * e)
        assert abs(ddw - fd) <= 1e-6 * max(abs(fd), 1e-12), (H, D, ddw, fd)


def test_breaking_church_thornton_closed_form():
    """Church-Thornton (src/WaveParametrizations.cpp:155-203) at H = gamma_b D, where the
    tanh switch is 0: Dw/E = 3 B^3 sigma H / (4 sqrt(pi) D) (1 - 2^(-5/2))."""
    p = S.WaveParams()
    for D, sigma in [(1.0, 0.8), (3.0, 1.1)]:
        H = p.gamma_b * D
        dw, _ = oracle.wave_breaking(p, H, sigma, D)
        ref = 3 * p.B3 * sigma * H / (4 * math.sqrt(math.pi) * D) * (1 - 2 ** -2.5)
        assert abs(dw - ref) <= 1e-14 * ref


def test_green_law_discrete_identity():
    """Normal incidence on an alongshore-uniform beach with breaking made negligible
    (gamma_b = 1e6): the direction stays 0 and every step is a mesh row. The energy balance
    of src/WaveEnergy.cpp:181-219 is then linear in lE with root
    lE* = lE_(i-1) - 2 (Cg_i - Cg_(i-1)) / (Cg_i + Cg_(i-1)), the scheme's form of E Cg = const
    (Green's law). Its Newton step divides by 2 Gp (`gp += gp + ...`, :97), so each iteration
    halves the distance to the root: after K iterations lE = lE* + (lE_(i-1) - lE*) / 2^K, with
    K the first count whose step is within the tolerance, or the iteration cap."""
    x, D = _beach(nx=40, ny=8, depth=(8.0, 1.0))
    p = S.WaveParams(gamma_b=1e6)
    w, n, f = _run(x, D, 4.0, p, hOff=0.5, theta=0.0, T=9.0)
    assert n == len(x) - 1
    # (a + b) - a - b of the y-derivative of log k (src/WaveDir.cpp:58) leaves ~1e-17
    assert np.max(np.abs(f["Theta"])) < 1e-15
    Cg = f["Cg"][:, 0]
    lE = np.log(f["E"][:, 0])
    for i in range(1, len(x)):
        d = -2 * (Cg[i] - Cg[i - 1]) / (Cg[i] + Cg[i - 1])     # lE* - lE_(i-1)
        K = next((k for k in range(1, p.max_iterations + 1) if abs(d) / 2 ** k <= p.tolerance),
                 p.max_iterations)
        pred = lE[i - 1] + d * (1 - 2.0 ** -K)
        assert abs(lE[i] - pred) <= 1e-12, (i, lE[i], pred)
    assert np.all(f["H"] == f["H"][:, :1])
    flux = f["E"][:, 0] * Cg                                   # E Cg, to the scheme's accuracy
    assert np.max(np.abs(flux / flux[0] - 1)) < 2e-3


def test_snell_discrete_refraction():
    """Oblique incidence on an alongshore-uniform beach: the ray construction of
    src/WaveDir.cpp:46-62 reduces to theta_i = theta_(i-1) - dlK tan(theta_(i-1)) / (1 + dlK/2)
    with dlK = log k_i - log k_(i-1), a consistent form of Snell's law k sin(theta) = const."""
    x, D = _beach(nx=40, ny=8, depth=(8.0, 1.0))
    p = S.WaveParams(gamma_b=1e6)
    th0 = math.radians(25.0)
    w, n, f = _run(x, D, 20.0, p, hOff=0.5, theta=th0, T=9.0)
    assert n == len(x) - 1          # dy / tan(theta) > dx: no intermediate rows
    lK, th = f["K"][:, 0], f["Theta"][:, 0]
    pred = th[:-1] - np.diff(lK) * np.tan(th[:-1]) / (1 + 0.5 * np.diff(lK))
    np.testing.assert_allclose(th[1:], pred, rtol=0, atol=1e-13)
    ksin = np.exp(lK) * np.sin(th)
    assert np.max(np.abs(ksin / ksin[0] - 1)) < 5e-3
    assert np.all(f["Theta"] == f["Theta"][:, :1])                 # uniform along y


def test_mirror_and_shift_symmetry():
    """y -> -y with theta -> -theta mirrors every field (Theta, Sxy change sign); a periodic
    shift of the bathymetry in y shifts every field. Both on a barred beach with oblique
    incidence and intermediate rows."""
    x, D = _beach(nx=40, ny=24, depth=(5.0, 0.3), bar=0.5, seed=7)
    p = S.WaveParams.base_case()
    dy = 1.0                                # dy/tan(theta) < dx: intermediate rows
    w1, n1, f1 = _run(x, D, dy, p, hOff=1.0, theta=0.5, T=8.0)
    assert n1 > len(x) - 1
    Dm = np.ascontiguousarray(np.roll(D[:, ::-1], 1, axis=1))     # j -> -j (mod ny)
    w2, n2, f2 = _run(x, Dm, dy, p, hOff=1.0, theta=-0.5, T=8.0)
    assert n2 == n1
    mir = lambda a: np.roll(a[:, ::-1], 1, axis=1)
    for k, sgn in [("H", 1), ("Theta", -1), ("Sxx", 1), ("Syy", 1), ("Sxy", -1), ("Dw", 1)]:
        np.testing.assert_allclose(f2[k], sgn * mir(f1[k]), rtol=1e-11, atol=1e-11 * np.max(np.abs(f1[k])))
    Ds = np.ascontiguousarray(np.roll(D, 5, axis=1))
    w3, n3, f3 = _run(x, Ds, dy, p, hOff=1.0, theta=0.5, T=8.0)
    for k in ("H", "Theta", "Sxy", "Dw"):
        assert np.array_equal(f3[k], np.roll(f1[k], 5, axis=1)), k


def test_rollers_and_friction_run():
    """Rollers (src/WaveRollers.cpp) and both friction models produce finite fields; the
    roller energy is non-negative and zero where no wave energy has been dissipated yet."""
    x, D = _beach(nx=40, ny=12, depth=(5.0, 0.3), bar=0.4, seed=2)
    for fm in (0, 1):
        p = S.WaveParams(rollers=1, bottom_friction=1, friction_model=fm)
        w, n, f = _run(x, D, 4.0, p, hOff=1.0, theta=0.2, T=8.0)
        for k in ("E_r", "D_r", "Sxx", "Urms"):
            assert np.all(np.isfinite(f[k])), (fm, k)
        assert np.all(f["E_r"] >= -1e-12)
        assert np.all(f["E_r"][0] == 0.0)
