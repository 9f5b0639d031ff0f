"""Batched bathymetry inversion: gradients of many scenarios at once, and gradient
descent on the bed perturbation h.

Each scenario has its own Hydro handle, HIP stream and adjoint buffers. A gradient of
L_s = 1/2 sum (Zs_N - Zs_obs)^2 for every scenario s is one checkpointed forward sweep and
one reverse sweep per scenario (morfo70_amd.adjoint); the scenarios' sweeps are issued
from a pool of host threads (ctypes releases the GIL during the library calls), so the
GPU runs several scenarios' kernels at the same time.
"""
from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import _lib
from .adjoint import HydroModel, h_adjoint
from .hydro import from_scenario


class BatchInversion:
    def __init__(self, scenarios, nIter: int, slots: int, cD0: float, workers: int = 4,
                 solver: int = _lib.SOLVER_FUSED):
        self.cD0 = float(cD0)
        self.nIter = int(nIter)
        self.models, self.obs = [], []
        for sc in scenarios:
            hy = from_scenario(sc, solver=solver)
            hy.set_forcing()
            self.models.append(HydroModel(hy, sc.Z, sc.zs0, nIter, slots))
        self.workers = max(1, int(workers))

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
        grads, _ = m.adj.reverse({"Zs": r}, want=("Zs", "D"))
        with torch.cuda.stream(m.stream):
            gh = h_adjoint(m.Z, h, m.zs0, m.hy._params.Dmin, grads["D"], grads["Zs"])
        m.stream.synchronize()
        return float(loss), gh

    def loss_and_grad(self, hs, concurrent: bool = True):
        """hs: one device tensor per scenario (ready on the device). Returns the losses
        and the gradients dL_s/dh_s."""
        import torch
        torch.cuda.synchronize()
        if concurrent and self.workers > 1:
            with ThreadPoolExecutor(self.workers) as ex:
                res = list(ex.map(lambda q: self._one(q, hs[q]), range(len(self.models))))
        else:
            res = [self._one(q, hs[q]) for q in range(len(self.models))]
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
        for m in self.models:
            m.adj.release()
            m.hy.close()
