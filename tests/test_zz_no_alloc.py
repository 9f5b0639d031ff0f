"""m70h_tridiag_solve_noalloc (static device scratch, no allocation) against
m70h_tridiag_solve with a caller-allocated workspace: bitwise equal results for both
solver families and both line layouts, over two calls in a row where the second is
smaller (the static array keeps the first call's contents), and a clear rejection of a
solve whose scratch exceeds the static maximum."""
from __future__ import annotations

import numpy as np
import pytest

from morfo70_amd import _lib

pytestmark = pytest.mark.gpu

MAX_SCRATCH = 4194304   # M70H_NOALLOC_SCRATCH_DOUBLES


def _system(rng, n, nl, periodic):
    a = rng.uniform(2.5, 5.0, (nl, n))
    b = rng.uniform(-1.0, 1.0, (nl, n))
    c = rng.uniform(-1.0, 1.0, (nl, n))
    f = rng.normal(size=(nl, n))
    if not periodic:
        b[:, 0] = 0.0
        c[:, -1] = 0.0
    return a, b, c, f


def _solve(noalloc, solver, rows, periodic, a, b, c, f):
    import torch
    L = _lib.load()
    dev = torch.device("cuda:0")
    nl, n = a.shape
    if rows:
        arrs, ld = [np.ascontiguousarray(v) for v in (a, b, c, f)], n
    else:
        arrs, ld = [np.ascontiguousarray(v.T) for v in (a, b, c, f)], nl
    T = [torch.from_numpy(v).to(dev) for v in arrs]
    Y = torch.full_like(T[0], float("nan"))
    torch.cuda.synchronize()
    ptrs = [t.data_ptr() for t in T]
    if noalloc:
        rc = L.m70h_tridiag_solve_noalloc(None, solver, int(rows), int(periodic), n, nl, ld, *ptrs, Y.data_ptr(), ld, 0)
    else:
        scratch = torch.zeros(3 * n * (((nl + 63) // 64) * 64) + n * ld + 64, dtype=torch.float64, device=dev)
        rc = L.m70h_tridiag_solve(None, solver, int(rows), int(periodic), n, nl, ld, *ptrs, Y.data_ptr(), ld,
                                  scratch.data_ptr(), scratch.numel(), 0)
    assert rc == 0, L.m70h_last_error(None)
    torch.cuda.synchronize()
    y = Y.cpu().numpy()
    return y if rows else np.ascontiguousarray(y.T)


@pytest.mark.parametrize("solver,rows,periodic", [(_lib.SOLVER_EXACT, True, True), (_lib.SOLVER_EXACT, False, False),
                                                  (_lib.SOLVER_PARTITION, True, True),
                                                  (_lib.SOLVER_PARTITION, False, False)])
def test_noalloc_equals_workspace_entry_point(solver, rows, periodic):
    rng = np.random.default_rng(11 + solver + 2 * rows)
    print(f"[noalloc] solver={solver} rows={rows} periodic={periodic}", flush=True)
    # two calls in a row on the shared static array, the second smaller
    for n, nl in ((1500, 700), (37, 90)):
        a, b, c, f = _system(rng, n, nl, periodic)
        want = _solve(False, solver, rows, periodic, a, b, c, f)
        got = _solve(True, solver, rows, periodic, a, b, c, f)
        assert np.isfinite(want).all()
        assert np.array_equal(got, want), (n, nl)


def test_noalloc_rejects_scratch_above_the_static_maximum():
    L = _lib.load()
    n, nl = 4096, 400     # EXACT periodic rows: 3 * 4096 * 448 doubles > the maximum
    assert 3 * n * 448 > MAX_SCRATCH
    rc = L.m70h_tridiag_solve_noalloc(None, _lib.SOLVER_EXACT, 1, 1, n, nl, n, 1, 1, 1, 1, 1, n, 0)
    assert rc == _lib.M70H_EINVAL
    assert b"static maximum" in L.m70h_last_error(None)
