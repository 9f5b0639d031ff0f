"""Shared helpers for the parity tests."""
from __future__ import annotations

import numpy as np

from morfo70_amd import scenarios as S

FIELDS = ("Zs", "D", "Qx", "Qy", "U", "V", "U_c", "V_c", "dQx_dx", "dQy_dy")


def max_rel(a: np.ndarray, b: np.ndarray) -> float:
    """max|a-b| / max|b| (max-norm relative error; dry faces are exact zeros)."""
    den = float(np.max(np.abs(b)))
    num = float(np.max(np.abs(a - b)))
    if den == 0.0:
        return num
    return num / den


def compare_states(got: dict, ref: dict, fields=FIELDS):
    out = {}
    for k in fields:
        g, r = got[k], ref[k]
        assert g.shape == r.shape, (k, g.shape, r.shape)
        out[k] = dict(rel=max_rel(g, r), exact=bool(np.array_equal(g, r)),
                      nan=int(np.isnan(g).sum()), ref_nan=int(np.isnan(r).sum()),
                      mismatch=int(np.count_nonzero(g != r)))
    return out


def nonuniform_beach(nx=24, ny=20, seed=0, steps=5, physics=None, tide=False, Hrms=1.0):
    """Small beach on a stretched cross-shore grid (dx from 6 m offshore to 1.5 m
    onshore), with dry cells and a bar; exercises every non-uniform metric weight."""
    rng = np.random.default_rng(seed)
    t = np.linspace(0.0, 1.0, nx)
    dxs = 6.0 - 4.5 * t[:-1]
    x = np.concatenate([[0.0], np.cumsum(dxs)])
    dy = 3.0
    y = np.arange(ny) * dy
    x_shore = x[-1] - 0.15 * x[-1]
    s = x_shore - x
    Z = np.repeat((0.03 * s)[:, None], ny, axis=1)
    amp = 0.6 + 0.2 * np.sin(2 * np.pi * y / (ny * dy) + rng.uniform(0, 6.28))
    h = np.exp(-((s - 25.0) / 8.0) ** 2)[:, None] * amp[None, :]
    ph = physics or S.Physics(dt=0.25)
    sc = S.Scenario(f"nonuniform_{nx}x{ny}", nx, ny, float(np.mean(np.diff(x))), dy, x,
                    np.ascontiguousarray(Z), np.ascontiguousarray(h), 0.0, ph, {}, steps)
    # unspread closed form: short runs on these small grids stay stable without it, and
    # it keeps the offshore rows quiet (the translation pin depends on that)
    sc.forcing = S.wave_forcing(sc.initial_depth(), x, Hrms, 7.0, 15.0, ph.Dmin, spread_m=0.0)
    if tide:
        sc.f0_series = S.tide_f0_series(steps, ph.dt, x[1] - x[0], amp=0.5, period=600.0)
    return sc
