"""Adjoint of Hydro::Run on the GPU, and its torch.autograd wrapper.

The library (include/m70h.h, m70h_adjoint_*) runs a Run with binomial checkpoints and
then sweeps backwards through hand-written adjoint kernels: given the adjoint of the
final state it returns the adjoint of the initial state and of the parameters cD0, nu0
and M. This module maps that onto the quantities a bathymetry or friction fit needs:

* the initial state is built from the bed perturbation h as the reference builds it,
  D = max(Z - h + zs0, Dmin) (src/Bathymetry.cpp:33-39) and Zs = D - Z + h
  (src/Hydro.cpp:35-37), all other fields zero;
* cD0 = g / C^2 for a Chezy coefficient C (the constant drag model).

``HydroRun.apply(h, cD0, model)`` is a torch.autograd.Function: forward returns the
final free surface Zs (a device tensor), backward returns dL/dh and dL/dcD0.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .hydro import Hydro

G = 9.81


def _fields(tensors: dict | None):
    f = _lib.Fields()
    if tensors:
        for k in _lib.STATE_FIELDS:
            t = tensors.get(k)
            if t is not None:
                assert t.is_cuda and t.dtype.is_floating_point and t.is_contiguous()
                setattr(f, k, t.data_ptr())
    return f


class RunAdjoint:
    """Adjoint of one Run on a Hydro handle (morfo70_amd.hydro.Hydro)."""

    def __init__(self, hydro: Hydro):
        self.hy = hydro
        self._L = hydro._L

    def forward(self, nIter: int, slots: int):
        """Run(nIter) keeping checkpoints; the handle ends at the final state."""
        self.nIter = int(nIter)
        self.hy._chk(self._L.m70h_adjoint_forward(self.hy.handle, self.nIter, int(slots)))

    def reverse(self, seed: dict, want=("Zs", "D")):
        """seed: adjoints of the final state (device float64 tensors, by field name),
        ready on the handle's stream. Returns (grads, dparams): adjoints of the initial
        state for the fields in ``want`` (device tensors) and a dict of the parameter
        adjoints. Returns after the sweep has finished."""
        import torch
        shapes = {k: ((self.hy.nx + 1, self.hy.ny) if k in ("Qx", "Qy", "U", "V") else (self.hy.nx, self.hy.ny))
                  for k in _lib.STATE_FIELDS}
        dev = next(iter(seed.values())).device
        grads = {k: torch.zeros(shapes[k], dtype=torch.float64, device=dev) for k in want}
        seedc = {k: v.contiguous() for k, v in seed.items()}
        sf, gf = _fields(seedc), _fields(grads)
        dp = (ctypes.c_double * 3)()
        torch.cuda.current_stream().synchronize()   # seeds and zeroed outputs are ready
        self.hy._chk(self._L.m70h_adjoint_reverse(self.hy.handle, ctypes.byref(sf), ctypes.byref(gf),
                                                  ctypes.cast(dp, ctypes.POINTER(ctypes.c_double))))
        return grads, {"cD0": dp[0], "nu0": dp[1], "M": dp[2]}

    def stats(self):
        f, t, b = ctypes.c_long(), ctypes.c_long(), ctypes.c_int()
        self.hy._chk(self._L.m70h_adjoint_stats(self.hy.handle, ctypes.byref(f), ctypes.byref(t), ctypes.byref(b)))
        return {"forward_steps": f.value, "taped_steps": t.value, "buffers": b.value}

    def release(self):
        self.hy._chk(self._L.m70h_adjoint_release(self.hy.handle))


def initial_state(Z, h, zs0: float, Dmin: float):
    """D and Zs of the reference's constructors from the bed perturbation h (torch)."""
    import torch
    d = (Z - h) + zs0
    D = torch.where(d < Dmin, torch.full_like(d, Dmin), d)
    Zs = (D - Z) + h
    return D, Zs


def h_adjoint(Z, h, zs0: float, Dmin: float, gD, gZs):
    """Chain rule through initial_state: dL/dh from the adjoints of D and Zs."""
    import torch
    wet = ((Z - h) + zs0) >= Dmin
    return torch.where(wet, -gD, gZs)


class HydroModel:
    """A Hydro handle plus what the autograd wrapper needs: the fixed Z, zs0, the run
    length, the checkpoint budget (0 = tape every step), and the state to restart from.

    The handle and the model's torch work share one stream (``self.stream``), so calls
    are ordered without device-wide synchronisation and several models can run at once
    from different threads."""

    def __init__(self, hydro: Hydro, Z, zs0: float, nIter: int, slots: int, stream=None):
        import torch
        self.hy = hydro
        self.adj = RunAdjoint(hydro)
        self.stream = stream if stream is not None else torch.cuda.Stream()
        hydro.set_stream(self.stream.cuda_stream)
        with torch.cuda.stream(self.stream):
            self.Z = torch.as_tensor(np.ascontiguousarray(Z), dtype=torch.float64, device="cuda")
        self.stream.synchronize()
        self.zs0 = float(zs0)
        self.nIter = int(nIter)
        self.slots = int(slots)

    def set_drag(self, cD0: float):
        p = self.hy._params
        p.cD0 = float(cD0)
        self.hy._chk(self.hy._L.m70h_set_params(self.hy.handle, ctypes.cast(ctypes.pointer(p), ctypes.c_void_p)))

    def load_initial(self, h):
        """Sets the handle's state to the reference's initial state for h (h must be
        ready on this model's stream)."""
        import torch
        with torch.cuda.stream(self.stream):
            D, Zs = initial_state(self.Z, h, self.zs0, self.hy._params.Dmin)
            z = torch.zeros_like(D)
            zp = torch.zeros((self.hy.nx + 1, self.hy.ny), dtype=torch.float64, device=D.device)
            keep = [Zs.contiguous(), D.contiguous(), zp, zp.clone(), zp.clone(), zp.clone(), z, z.clone(), z.clone(),
                    z.clone()]
        # the copy runs on the same stream, after the kernels that produce its sources
        self.hy._chk(self.hy._L.m70h_set_state_dev(self.hy.handle, *[ctypes.c_void_p(t.data_ptr()) for t in keep]))

    def run_plain(self, h, cD0: float):
        """Forward only (no checkpoints): Zs after nIter steps from the state of h."""
        self.set_drag(cD0)
        self.load_initial(h)
        self.hy._chk(self.hy._L.m70h_run(self.hy.handle, self.nIter))
        return self.final_zs()

    def final_zs(self):
        import torch
        with torch.cuda.stream(self.stream):
            out = torch.empty((self.hy.nx, self.hy.ny), dtype=torch.float64, device="cuda")
        self.hy._chk(self.hy._L.m70h_get_state_dev(self.hy.handle, ctypes.c_void_p(out.data_ptr()),
                                                   *([None] * 9)))
        return out


def _hydro_run_function():
    import torch

    class HydroRun(torch.autograd.Function):
        """Zs after nIter steps as a function of the bed perturbation h and cD0."""

        @staticmethod
        def forward(ctx, h, cD0, model: HydroModel):
            model.set_drag(float(cD0))
            cur = torch.cuda.current_stream()
            model.stream.wait_stream(cur)
            model.load_initial(h.detach())
            model.adj.forward(model.nIter, model.slots)
            ctx.model = model
            ctx.save_for_backward(h.detach())
            out = model.final_zs()
            cur.wait_stream(model.stream)
            return out

        @staticmethod
        def backward(ctx, gout):
            model = ctx.model
            (h,) = ctx.saved_tensors
            cur = torch.cuda.current_stream()
            model.stream.wait_stream(cur)
            with torch.cuda.stream(model.stream):
                seed = gout.contiguous().to(torch.float64)
            grads, dp = model.adj.reverse({"Zs": seed}, want=("Zs", "D"))
            with torch.cuda.stream(model.stream):
                gh = h_adjoint(model.Z, h, model.zs0, model.hy._params.Dmin, grads["D"], grads["Zs"])
            cur.wait_stream(model.stream)
            gc = torch.tensor(dp["cD0"], dtype=torch.float64, device=h.device)
            return gh, gc, None

    return HydroRun


_HR = None


def hydro_run(h, cD0, model: HydroModel):
    """Zs after model.nIter steps; differentiable in h and cD0 (torch tensors)."""
    global _HR
    if _HR is None:
        _HR = _hydro_run_function()
    return _HR.apply(h, cD0, model)
