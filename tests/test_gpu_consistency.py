"""The HIP step (FUSED and EXACT families, through the C ABI) against the continuous
equations it discretises: the checks of test_oracle_consistency.py run on the GPU path
directly. Second-order convergence of the one-step tendency to the flux-form
shallow-water right-hand sides, and the exact semi-implicit gravity identities of each
half-step (see that module's docstring)."""
from __future__ import annotations

import pytest

from morfo70_amd import _lib
from tests import test_oracle_consistency as C

pytestmark = pytest.mark.gpu


def gpu_step(solver):
    def step(sc, st):
        from morfo70_amd.hydro import from_scenario
        g = from_scenario(sc, solver=solver)
        for k, v in st.items():
            if k == "D":
                g.bathy.D[...] = v
            else:
                setattr(g, k, v.copy())
        g.Run(1)
        out = g.state()
        g.close()
        return out
    return step


@pytest.mark.parametrize("solver", [_lib.SOLVER_FUSED, _lib.SOLVER_EXACT])
def test_gpu_step_is_second_order_consistent(solver):
    print(f"[consistency] solver {solver}", flush=True)
    step = gpu_step(solver)
    e128, e256 = C.consistency_errors(128, step=step), C.consistency_errors(256, step=step)
    print(e128, e256, flush=True)
    for k in e256:
        assert e128[k] / e256[k] >= C.ORDER_MIN and e256[k] <= C.CONSISTENCY_TOL, (k, e128[k], e256[k])


@pytest.mark.parametrize("solver", [_lib.SOLVER_FUSED, _lib.SOLVER_EXACT])
def test_gpu_half_steps_are_semi_implicit_gravity_steps(solver):
    res = C.identity_residuals(step=gpu_step(solver))
    print(res, flush=True)
    for k, v in res.items():
        assert v <= C.IDENTITY_TOL[k], (k, v)
