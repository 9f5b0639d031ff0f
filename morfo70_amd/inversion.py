"""Batched bathymetry inversion: gradients of many scenarios at once, and the recovery of
each scenario's bar.

Each scenario has its own Hydro handle, stream and adjoint buffers. A gradient of
L_s = 1/2 sum (Zs_N - Zs_obs)^2 for every scenario s is one checkpointed forward sweep and
one reverse sweep. By default all scenarios' sweeps run together as grouped launches
(m70h_adjoint_batch_*: one launch per stage of the schedule over every scenario); with
grouped=False each scenario's sweeps are issued from a pool of host threads instead
(ctypes releases the GIL during the library calls), one stream per scenario.
"""
from __future__ import annotations

import ctypes
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import _lib
from .adjoint import HydroModel, _fields, h_adjoint
from .hydro import from_scenario


class BatchInversion:
    def __init__(self, scenarios, nIter: int, slots: int, cD0: float, workers: int = 4,
                 solver: int = _lib.SOLVER_FUSED, grouped: bool = True):
        self.cD0 = float(cD0)
        self.nIter = int(nIter)
        self.slots = int(slots)
        self.models, self.obs = [], []
        for sc in scenarios:
            hy = from_scenario(sc, solver=solver)
            hy.set_forcing()
            self.models.append(HydroModel(hy, sc.Z, sc.zs0, nIter, slots))
        self.workers = max(1, int(workers))
        self.grouped = bool(grouped)
        self._batches = {}                           # scenario subset -> m70h_adjoint_batch
        self.last_dcD0 = [0.0] * len(self.models)   # dL_q/dcD0 of the last gradient

    def _adjoint_batch(self, qs):
        key = tuple(qs)
        if key not in self._batches:
            L = self.models[0].hy._L
            arr = (ctypes.c_void_p * len(qs))(*[self.models[q].hy.handle for q in qs])
            b = ctypes.c_void_p()
            rc = L.m70h_adjoint_batch_create(len(qs), arr, ctypes.byref(b))
            if rc != _lib.M70H_OK:
                raise _lib.M70HError(rc, L.m70h_adjoint_batch_last_error(None).decode())
            self._batches[key] = b
        return self._batches[key]

    def _grouped(self, qs, hs):
        """Losses and h gradients of the scenarios qs from one grouped forward and reverse
        sweep over all of them."""
        import torch
        L = self.models[0].hy._L
        b = self._adjoint_batch(qs)

        def chk(rc):
            if rc != _lib.M70H_OK:
                raise _lib.M70HError(rc, L.m70h_adjoint_batch_last_error(b).decode())
        for q, h in zip(qs, hs):
            m = self.models[q]
            m.set_drag(self.cD0)
            m.load_initial(h)
        torch.cuda.synchronize()
        chk(L.m70h_adjoint_batch_forward(b, self.nIter, self.slots))
        losses, seeds, grads, keep = [], [], [], []
        for q in qs:
            m = self.models[q]
            eta = m.final_zs()
            with torch.cuda.stream(m.stream):
                r = (eta - self.obs[q]).contiguous()
                losses.append(0.5 * torch.sum(r * r))
                g = {"Zs": torch.zeros_like(r), "D": torch.zeros_like(r)}
            keep.append((r, g))
            seeds.append(_fields({"Zs": r}))
            grads.append(_fields(g))
        torch.cuda.synchronize()
        seed_arr = (_lib.Fields * len(qs))(*seeds)
        grad_arr = (_lib.Fields * len(qs))(*grads)
        dp = (ctypes.c_double * (3 * len(qs)))()
        chk(L.m70h_adjoint_batch_reverse(b, ctypes.cast(seed_arr, ctypes.c_void_p),
                                         ctypes.cast(grad_arr, ctypes.c_void_p),
                                         ctypes.cast(dp, ctypes.POINTER(ctypes.c_double))))
        out_l, out_g = [], []
        for k, (q, h) in enumerate(zip(qs, hs)):
            m = self.models[q]
            _, g = keep[k]
            with torch.cuda.stream(m.stream):
                out_g.append(h_adjoint(m.Z, h, m.zs0, m.hy._params.Dmin, g["D"], g["Zs"]))
            self.last_dcD0[q] = dp[3 * k]
            out_l.append(float(losses[k]))
        torch.cuda.synchronize()
        return out_l, out_g

    def final_zs(self, hs):
        """Zs after nIter steps from the initial state of each h (forward only)."""
        out = []
        for m, h in zip(self.models, hs):
            out.append(m.run_plain(h, self.cD0))
            m.stream.synchronize()
        return out

    def set_observations(self, obs):
        import torch
        self.obs = []
        for m, ob in zip(self.models, obs):
            with torch.cuda.stream(m.stream):
                self.obs.append(torch.as_tensor(ob, device="cuda").clone())
            m.stream.synchronize()

    def _one(self, q: int, h):
        import torch
        m = self.models[q]
        m.set_drag(self.cD0)
        m.load_initial(h)
        m.adj.forward(self.nIter, m.slots)
        eta = m.final_zs()
        with torch.cuda.stream(m.stream):
            r = eta - self.obs[q]
            loss = 0.5 * torch.sum(r * r)
        grads, dp = m.adj.reverse({"Zs": r}, want=("Zs", "D"))
        with torch.cuda.stream(m.stream):
            gh = h_adjoint(m.Z, h, m.zs0, m.hy._params.Dmin, grads["D"], grads["Zs"])
        m.stream.synchronize()
        self.last_dcD0[q] = dp["cD0"]
        return float(loss), gh

    def loss_and_grad(self, hs, concurrent: bool = True, which=None, grouped: bool | None = None):
        """hs: one device tensor per scenario (ready on the device). Returns the losses
        and the gradients dL_s/dh_s (dL_s/dcD0 in last_dcD0). `which`: the scenarios to
        evaluate (default all; hs then holds one tensor per listed scenario). grouped
        (default: the instance's choice) runs all of them as grouped launches; otherwise
        concurrent issues each scenario's sweeps from its own host thread and stream."""
        import torch
        qs = list(range(len(self.models))) if which is None else list(which)
        torch.cuda.synchronize()
        if self.grouped if grouped is None else grouped:
            return self._grouped(qs, hs)
        if concurrent and self.workers > 1 and len(qs) > 1:
            with ThreadPoolExecutor(self.workers) as ex:
                res = list(ex.map(lambda k: self._one(qs[k], hs[k]), range(len(qs))))
        else:
            res = [self._one(q, h) for q, h in zip(qs, hs)]
        return [r[0] for r in res], [r[1] for r in res]

    def descend(self, hs, iters: int, max_step: float = 0.02, concurrent: bool = True):
        """Gradient descent, h <- h - a g with a = max_step / max|g| (the largest change
        of h per iteration is max_step metres). Returns the final h and a history of
        (iteration, losses, seconds)."""
        import torch
        hist = []
        hs = [h.clone() for h in hs]
        for it in range(iters):
            t0 = time.perf_counter()
            losses, grads = self.loss_and_grad(hs, concurrent)
            for q, g in enumerate(grads):
                gmax = float(torch.max(torch.abs(g)))
                if gmax > 0:
                    hs[q] = hs[q] - (max_step / gmax) * g
            torch.cuda.synchronize()
            hist.append((it, losses, time.perf_counter() - t0))
        final, _ = self.loss_and_grad(hs, concurrent)
        hist.append((iters, final, 0.0))
        return hs, hist

    def close(self):
        if self._batches:
            L = self.models[0].hy._L
            for b in self._batches.values():
                L.m70h_adjoint_batch_destroy(b)
            self._batches = {}
        for m in self.models:
            m.adj.release()
            m.hy.close()


class BarInversion:
    """Recovers each scenario's bar from its final-surface observations.

    The control of scenario q is its bar, amplitude a and crest position c (metres seaward
    of the shoreline; scenarios.bar_field), in scaled variables z = (a/a0 - 1, (c - c0)/10 m)
    around the first guess (a0, c0). The gradient of L_q = 1/2 sum (Zs_N - Zs_obs)^2 with
    respect to z is the adjoint's gradient with respect to h at every cell, contracted with
    dh/da and dh/dc. Each scenario runs its own BFGS iteration (2 x 2 inverse Hessian) with
    a backtracking Armijo line search; trial points are judged by the plain forward Run
    (no tape), as is the misfit reported at the end."""

    C_SCALE = 10.0   # metres of crest position per unit of z[1]

    def __init__(self, scenarios, nIter: int, slots: int, cD0: float, workers: int = 8):
        from .scenarios import bar_field
        self._bar = bar_field
        self.scs = list(scenarios)
        self.inv = BatchInversion(self.scs, nIter, slots, cD0, workers=workers)
        self.theta0 = [(sc.meta["bar_amp"], sc.meta["bar_centre"]) for sc in self.scs]

    def params(self, q: int, z):
        a0, c0 = self.theta0[q]
        return a0 * (1.0 + z[0]), c0 + self.C_SCALE * z[1]

    def h_of(self, q: int, z):
        import torch
        a, c = self.params(q, z)
        h, da, dc = self._bar(self.scs[q], a, c)
        return (torch.as_tensor(h, device="cuda"), torch.as_tensor(da, device="cuda"),
                torch.as_tensor(dc, device="cuda"))

    def set_observations(self, obs):
        self.inv.set_observations(obs)

    def plain_losses(self, qs, zs):
        """L_q at z for the scenarios qs, from the plain forward Run."""
        import torch
        out = []
        for q, z in zip(qs, zs):
            h, _, _ = self.h_of(q, z)
            m = self.inv.models[q]
            eta = m.run_plain(h, self.inv.cD0)
            with torch.cuda.stream(m.stream):
                r = eta - self.inv.obs[q]
                v = 0.5 * torch.sum(r * r)
            m.stream.synchronize()
            out.append(float(v))
        return out

    def gradients(self, zs, which=None):
        """dL_q/dz for the scenarios `which` (default all): one checkpointed forward and
        reverse sweep each, the scenarios' sweeps concurrent."""
        import torch
        qs = list(range(len(self.scs))) if which is None else list(which)
        hs, jac = [], []
        for q in qs:
            h, da, dc = self.h_of(q, zs[q])
            hs.append(h)
            a0, _ = self.theta0[q]
            jac.append((da * a0, dc * self.C_SCALE))
        torch.cuda.synchronize()
        _, gh = self.inv.loss_and_grad(hs, which=qs)
        return {q: np.array([float(torch.sum(g * ja)), float(torch.sum(g * jc))]) for q, g, (ja, jc) in zip(qs, gh, jac)}

    def fit(self, iters: int = 20, first_step: float = 0.05, max_halvings: int = 8, rtol: float = 1e-9, log=None):
        """BFGS with Armijo backtracking on every scenario; returns (z per scenario, history
        of (iteration, plain losses, seconds, scenarios still moving)). A scenario stops
        once its line search finds no decrease, or one below rtol of its misfit."""
        n = len(self.scs)
        zs = [np.zeros(2) for _ in range(n)]
        f = self.plain_losses(range(n), zs)
        g = self.gradients(zs)
        H = {q: np.eye(2) * (first_step / max(np.max(np.abs(g[q])), 1e-300)) for q in range(n)}
        active = list(range(n))
        hist = [(0, list(f), 0.0, n)]
        for it in range(1, iters + 1):
            if not active:
                break
            t0 = time.perf_counter()
            p = {q: -(H[q] @ g[q]) for q in active}
            t = {q: 1.0 for q in active}
            accepted = {q: False for q in active}
            for _ in range(max_halvings + 1):
                todo = [q for q in active if not accepted[q]]
                if not todo:
                    break
                trial = self.plain_losses(todo, [zs[q] + t[q] * p[q] for q in todo])
                for q, fq in zip(todo, trial):
                    if fq <= f[q] + 1e-4 * t[q] * float(g[q] @ p[q]):
                        accepted[q] = True
                        if f[q] - fq <= rtol * f[q]:
                            active = [a for a in active if a != q]   # converged
                        zs[q] = zs[q] + t[q] * p[q]
                        f[q] = fq
                    else:
                        t[q] *= 0.5
            active = [q for q in active if accepted[q]]
            g_new = self.gradients(zs, which=active) if active else {}
            for q in active:
                s_, y_ = t[q] * p[q], g_new[q] - g[q]
                sy = float(s_ @ y_)
                if sy > 1e-12 * np.linalg.norm(s_) * np.linalg.norm(y_):
                    rho = 1.0 / sy
                    I = np.eye(2)
                    H[q] = (I - rho * np.outer(s_, y_)) @ H[q] @ (I - rho * np.outer(y_, s_)) + rho * np.outer(s_, s_)
                g[q] = g_new[q]
            hist.append((it, list(f), time.perf_counter() - t0, len(active)))
            if log:
                log(f"iteration {it}: mean misfit {np.mean(f):.9e}, {len(active)} scenarios moving, "
                    f"{hist[-1][2]:.2f} s")
        return zs, hist

    def close(self):
        self.inv.close()
