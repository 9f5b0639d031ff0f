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


@pytest.mark.parametrize("n,slots", [(1, 1), (4, 2), (4, 0)])
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
    """The checkpoint schedule only changes what is recomputed: gradients for a full
    tape (0), n, 3 and 1 buffers agree (to atomics' summation order)."""
    sc = nonuniform_beach(nx=30, ny=20, seed=7)
    st = _lin_point(sc, 4)
    n = 12
    seed = {"Zs": np.random.default_rng(2).normal(size=(sc.nx, sc.ny))}
    ref = None
    for slots in (0, n, 3, 1):
        g, dp, stats = _adjoint(sc, st, n, seed, slots)
        # the reverse sweep tapes every step again unless the forward sweep kept them
        assert stats["taped_steps"] == (0 if slots == 0 else n)
        assert (stats["forward_steps"] == 0) == (slots in (0, n))
        if ref is None:
            ref = (g, dp)
            continue
        for k in STATE:
            scale = max(np.max(np.abs(ref[0][k])), 1e-300)
            assert np.max(np.abs(g[k] - ref[0][k])) <= 1e-12 * scale, (slots, k)
        assert _rel(dp["cD0"], ref[1]["cD0"]) < 1e-12


def _wet_interior_cells(D0, count, rng, depth=1.0, margin=3):
    """Cells at least `depth` deep with no shallower cell within `margin` cells (wet and
    away from the swash zone, so a small perturbation moves no wet/dry mask)."""
    deep = D0 > depth
    ok = deep.copy()
    for di in range(-margin, margin + 1):
        for dj in range(-margin, margin + 1):
            ok &= np.roll(np.roll(deep, di, axis=0), dj, axis=1)
    ok[:margin + 1, :] = False
    ok[-margin - 1:, :] = False
    idx = np.argwhere(ok)
    return idx[rng.choice(len(idx), size=count, replace=False)]


def test_bathymetry_inversion_gradient_c4():
    """512 x 1024 beach, 400 steps, L = 1/2 sum (Zs_N - Zs_obs)^2 with observations from
    a perturbed bar field plus N(0, 0.02 m) noise. The gradient with respect to h (all
    cells) and the Chezy coefficient C, through the torch.autograd wrapper: against
    central differences of the GPU forward on 32 wet non-swash cells (<= 1e-6), and
    against the tangent-linear oracle along a random smooth direction of h and along
    cD0 = g/C^2."""
    from morfo70_amd.adjoint import HydroModel, hydro_run

    sc, dh, rng = S.config4()
    n = sc.steps
    hy = from_scenario(sc, solver=_lib.SOLVER_FUSED)
    hy.set_forcing()
    model = HydroModel(hy, sc.Z, sc.zs0, n, slots=48)
    dev = "cuda"
    C0 = 65.0
    cD0 = S.G / C0 ** 2
    h_true = torch.as_tensor(sc.h + dh, device=dev)
    obs = model.run_plain(h_true, cD0) + torch.as_tensor(rng.normal(0.0, 0.02, size=(sc.nx, sc.ny)), device=dev)
    h0 = torch.as_tensor(np.ascontiguousarray(sc.h), device=dev).requires_grad_(True)
    C = torch.tensor(C0, dtype=torch.float64, device=dev, requires_grad=True)
    eta = hydro_run(h0, S.G / C ** 2, model)
    loss = 0.5 * torch.sum((eta - obs) ** 2)
    loss.backward()
    gh = h0.grad.detach().cpu().numpy()
    gC = float(C.grad)
    assert np.isfinite(gh).all() and np.isfinite(gC)

    def dloss(hp, hm, cp, cm):   # L(+) - L(-) without cancellation
        ep, em = model.run_plain(hp, cp), model.run_plain(hm, cm)
        return float(torch.sum(0.5 * (ep - em) * ((ep - obs) + (em - obs))))

    # Fourth-order central differences (Richardson: (4 D(eps) - D(2 eps)) / 3). The step
    # grows with depth: L changes by ~1e-16 of its size between two runs from rounding
    # alone, so deep cells, whose gradients are 1e-7..1e-9, need steps of centimetres to
    # rise above that floor, while in 1-2 m of water the response turns nonlinear beyond
    # a few tenths of a millimetre.
    hb = torch.as_tensor(np.ascontiguousarray(sc.h), device=dev)
    D0 = sc.initial_depth()
    cells = _wet_interior_cells(D0, 32, np.random.default_rng(44))

    def fd4(step, pert):
        d1 = dloss(*pert(step)) / (2 * step)
        d2 = dloss(*pert(2 * step)) / (4 * step)
        return (4 * d1 - d2) / 3

    for i, j in cells:
        eps = min(1e-4 * max(1.0, D0[i, j] ** 2), 0.2)

        def pert(st, i=i, j=j):
            e = torch.zeros_like(hb)
            e[i, j] = st
            return hb + e, hb - e, cD0, cD0

        fd = fd4(eps, pert)
        rel = abs(fd - gh[i, j]) / max(abs(gh[i, j]), 1e-300)
        assert rel <= 1e-6, (int(i), int(j), float(D0[i, j]), fd, gh[i, j])
    # Tangent-linear oracle: along a smooth random direction of h, and along cD0. (A
    # global friction change moves some wet/dry mask in the swash zone over 400 steps, so
    # L is only piecewise smooth in C and its differences are not a derivative; the TLM
    # differentiates the same branch as the adjoint.)
    obs_h = obs.cpu().numpy()
    v = S.smooth_field(sc.nx, sc.ny, np.random.default_rng(45), 0.01)
    tl = oracle.TangentHydro(sc, dh=v)
    tl.run(n)
    vals, tans = tl.get_state()
    lhs = float(np.sum((vals["Zs"] - obs_h) * tans["Zs"]))
    rhs = float(np.sum(gh * v))
    assert abs(lhs - rhs) <= 1e-8 * max(abs(lhs), abs(rhs)), (lhs, rhs)
    tc = oracle.TangentHydro(sc, dparams={"cD0": 1.0})
    tc.run(n)
    vals, tans = tc.get_state()
    lhs = float(np.sum((vals["Zs"] - obs_h) * tans["Zs"]))
    gcD0 = gC / (-2 * S.G / C0 ** 3)   # dcD0/dC = -2 g / C^3
    assert abs(lhs - gcD0) <= 1e-8 * max(abs(lhs), abs(gcD0)), (lhs, gcD0)
    hy.close()


def test_batched_gradients_match_the_tangent_linear_oracle_at_c5_sizes():
    """The batched inversion's gradients on three of the C5 scenarios at their own sizes
    (nx U[256,1024], ny U[512,2048]; 60 steps instead of 300 to bound the oracle's time):
    the grouped sweeps (one launch per stage over all scenarios) equal each scenario's
    own sweeps, issued concurrently or one after another, to rounding (the reverse adds
    through FP64 atomics), and dL/dh along a random smooth direction and dL/dcD0 agree
    with the tangent-linear oracle (<= 1e-8)."""
    from morfo70_amd.inversion import BatchInversion

    steps = 60
    scs, dhs, rng = S.config5(count=3, steps=steps)
    print("[c5 tlm] " + ", ".join(f"{sc.nx}x{sc.ny}" for sc in scs), flush=True)
    cD0 = S.G / 65.0 ** 2
    inv = BatchInversion(scs, steps, slots=8, cD0=cD0, workers=3)
    truth = inv.final_zs([torch.as_tensor(sc.h + dh, device="cuda") for sc, dh in zip(scs, dhs)])
    inv.set_observations([z + torch.as_tensor(rng.normal(0, 0.02, size=z.shape), device="cuda") for z in truth])
    hs = [torch.as_tensor(np.ascontiguousarray(sc.h), device="cuda") for sc in scs]
    lg, gg = inv.loss_and_grad(hs)                                    # grouped
    dcg = list(inv.last_dcD0)
    lc, gc = inv.loss_and_grad(hs, grouped=False, concurrent=True)
    dcc = list(inv.last_dcD0)
    ls, gs = inv.loss_and_grad(hs, grouped=False, concurrent=False)
    dcs = list(inv.last_dcD0)
    for q in range(len(scs)):
        assert lg[q] == lc[q] == ls[q], q
        for other, dco in ((gc, dcc), (gs, dcs)):
            scale = float(torch.max(torch.abs(other[q])))
            assert float(torch.max(torch.abs(gg[q] - other[q]))) <= 1e-12 * scale, q
            assert abs(dcg[q] - dco[q]) <= 1e-12 * abs(dco[q]), q
    threads = oracle.host_cpus()["threads"]
    for q, sc in enumerate(scs):
        obs = inv.obs[q].cpu().numpy()
        gh = gg[q].cpu().numpy()
        v = S.smooth_field(sc.nx, sc.ny, np.random.default_rng(70 + q), 0.01)
        tl = oracle.TangentHydro(sc, dh=v, threads=threads)
        tl.run(steps)
        vals, tans = tl.get_state()
        lhs, rhs = float(np.sum((vals["Zs"] - obs) * tans["Zs"])), float(np.sum(gh * v))
        print(f"[c5 tlm] {sc.name} h: tlm {lhs:.12e} adjoint {rhs:.12e}", flush=True)
        assert abs(lhs - rhs) <= 1e-8 * max(abs(lhs), abs(rhs)), (sc.name, lhs, rhs)
        tc = oracle.TangentHydro(sc, dparams={"cD0": 1.0}, threads=threads)
        tc.run(steps)
        vals, tans = tc.get_state()
        lhs = float(np.sum((vals["Zs"] - obs) * tans["Zs"]))
        print(f"[c5 tlm] {sc.name} cD0: tlm {lhs:.12e} adjoint {dcg[q]:.12e}", flush=True)
        assert abs(lhs - dcg[q]) <= 1e-8 * max(abs(lhs), abs(dcg[q])), (sc.name, lhs, dcg[q])
    inv.close()


def test_bar_inversion_reaches_the_noise_floor():
    """Recovery of seeded perturbed bars (amplitude x U[0.75,1.25], crest moved by
    U[-12,12] m) from final-surface observations with N(0, 0.02 m) noise, on four small
    scenarios: every scenario's plain-forward misfit falls to that of the true bar or
    below (the noise floor), and the bar error shrinks to at most half its start in at
    least three of the four."""
    from morfo70_amd.inversion import BarInversion

    steps = 120
    scs, dhs, rng = S.config5(count=4, steps=steps, nx_range=(256, 320), ny_range=(512, 640))
    bi = BarInversion(scs, steps, slots=8, cD0=S.G / 65.0 ** 2, workers=4)
    truth = []
    for q, sc in enumerate(scs):
        m = bi.inv.models[q]
        truth.append(m.run_plain(torch.as_tensor(sc.h + dhs[q], device="cuda"), bi.inv.cD0))
        m.stream.synchronize()
    bi.set_observations([z + torch.as_tensor(rng.normal(0, 0.02, size=z.shape), device="cuda") for z in truth])
    z_true = [np.array([sc.meta["true_amp"] / a0 - 1.0, (sc.meta["true_centre"] - c0) / bi.C_SCALE])
              for sc, (a0, c0) in zip(scs, bi.theta0)]
    L_true = bi.plain_losses(range(len(scs)), z_true)
    zs, hist = bi.fit(iters=15)
    L0, LN = hist[0][1], hist[-1][1]
    shrunk = 0
    for q, sc in enumerate(scs):
        a, c = bi.params(q, zs[q])
        e0 = np.sqrt(np.mean(dhs[q] ** 2))
        e1 = np.sqrt(np.mean((S.bar_field(sc, a, c)[0] - (sc.h + dhs[q])) ** 2))
        print(f"[c5 fit] {sc.name} misfit {L0[q]:.6e} -> {LN[q]:.6e} (true bar {L_true[q]:.6e}) "
              f"bar rms {e0:.4f} -> {e1:.4f}", flush=True)
        assert LN[q] <= L_true[q] * (1 + 1e-9), (sc.name, LN[q], L_true[q])
        shrunk += e1 <= 0.5 * e0
    assert shrunk >= 3
    bi.close()
