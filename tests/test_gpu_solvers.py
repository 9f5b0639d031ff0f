"""GPU line solvers through m70h_tridiag_solve against dense solves and the oracle.

EXACT must reproduce the reference's Thomas / El-Mikkawy recurrences bit for bit (the
oracle's solvers, themselves pinned to the reference's Math.cpp in test_oracle_pins).
PARTITION solves the same systems with a different elimination order; on the diagonally
dominant systems of the Hydro step it must agree to 1e-13 relative (a few ulps times the
condition number).
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from morfo70_amd import _lib

pytestmark = pytest.mark.gpu


def dominant_system(rng, n, nlines, periodic, dominance=2.5):
    a = rng.uniform(1.0, 2.0, (nlines, n)) * dominance
    b = rng.uniform(-1.0, 1.0, (nlines, n))
    c = rng.uniform(-1.0, 1.0, (nlines, n))
    f = rng.normal(size=(nlines, n))
    if not periodic:
        b[:, 0] = 0.0
        c[:, -1] = 0.0
    return a, b, c, f


def dense_solve(a, b, c, f, periodic):
    nl, n = a.shape
    out = np.zeros_like(f)
    for L in range(nl):
        M = np.diag(a[L]) + np.diag(b[L, 1:], -1) + np.diag(c[L, :-1], 1)
        if periodic:
            M[0, n - 1] += b[L, 0]
            M[n - 1, 0] += c[L, n - 1]
        out[L] = np.linalg.solve(M, f[L])
    return out


def gpu_solve(solver, rows, periodic, a, b, c, f, tile=0):
    """a,b,c,f: (nlines, n) line-major host arrays."""
    import torch
    L = _lib.load()
    dev = torch.device("cuda:0")
    nl, n = a.shape
    if rows:
        arrs = [np.ascontiguousarray(v) for v in (a, b, c, f)]
        ld = n
    else:
        arrs = [np.ascontiguousarray(v.T) for v in (a, b, c, f)]   # element (k, line) at k*nl + line
        ld = nl
    T = [torch.from_numpy(v).to(dev) for v in arrs]
    Y = torch.zeros_like(T[0])
    scratch = torch.zeros(3 * n * (((nl + 63) // 64) * 64) + n * ld + 64, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    rc = L.m70h_tridiag_solve(None, solver, 1 if rows else 0, 1 if periodic else 0, n, nl, ld,
                              *[t.data_ptr() for t in T], Y.data_ptr(), ld, scratch.data_ptr(),
                              scratch.numel(), tile)
    assert rc == 0, L.m70h_last_error(None)
    torch.cuda.synchronize()
    y = Y.cpu().numpy()
    return y if rows else np.ascontiguousarray(y.T)


@pytest.mark.parametrize("n", [3, 4, 9, 17, 64, 100, 257, 512])
@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("rows", [True, False])
def test_partition_vs_dense(n, periodic, rows):
    """Row lines up to 512 unknowns (at most 64 segments of 8) run their reduced system's
    PCR in one wavefront by cross-lane permutes (1, 2, 16 and 64 segments here: n = 3 and
    4, 9, 100, 257 and 512); longer ones through LDS."""
    if periodic and not rows:
        pytest.skip("column lines are never periodic in the Hydro step")
    rng = np.random.default_rng(n + 7 * periodic + 3 * rows)
    a, b, c, f = dominant_system(rng, n, 37, periodic)
    ref = dense_solve(a, b, c, f, periodic)
    for tile in (0, 8, 16):
        y = gpu_solve(_lib.SOLVER_PARTITION, rows, periodic, a, b, c, f, tile)
        err = np.max(np.abs(y - ref)) / np.max(np.abs(ref))
        assert err < 1e-13, (tile, err)


@pytest.mark.parametrize("n,nl", [(3, 5), (31, 70), (4096, 9), (2047, 130)])
def test_exact_matches_oracle_solvers_bitwise(n, nl):
    import oracle
    rng = np.random.default_rng(n)
    a, b, c, f = dominant_system(rng, n, nl, periodic=True)
    # periodic rows (y-sweeps): El-Mikkawy per line
    y = gpu_solve(_lib.SOLVER_EXACT, True, True, a, b, c, f)
    ref = np.stack([oracle.tridiag_periodic(a[L], b[L], c[L], f[L]) for L in range(nl)])
    assert np.array_equal(y, ref)
    # column lines (x-sweeps): batched Thomas, element (k, line)
    a2, b2, c2, f2 = dominant_system(rng, n, nl, periodic=False)
    y2 = gpu_solve(_lib.SOLVER_EXACT, False, False, a2, b2, c2, f2)
    ref2 = oracle.tridiag(a2.T, b2.T, c2.T, f2.T).T
    assert np.array_equal(y2, ref2)


@pytest.mark.parametrize("n,nl,rows,periodic", [(4096, 33, True, True), (2047, 300, False, False),
                                                (1000, 50, True, False), (3000, 40, False, False),
                                                (6000, 9, True, True), (8192, 7, True, False),
                                                (8192, 70, False, False)])
def test_partition_long_lines_vs_exact(n, nl, rows, periodic):
    """Up to 8192 unknowns: rows above 4096 take the 1024-thread LDS row kernel, columns
    above 2048 the partition kernel with 2 columns (up to 4096) or 1 (up to 8192) per
    1024-thread workgroup."""
    rng = np.random.default_rng(n + nl)
    a, b, c, f = dominant_system(rng, n, nl, periodic, dominance=2.05)
    import oracle
    if rows and periodic:
        ref = np.stack([oracle.tridiag_periodic(a[L], b[L], c[L], f[L]) for L in range(nl)])
    else:
        ref = oracle.tridiag(a.T, b.T, c.T, f.T).T
    for tile in (0, 8, 16):
        y = gpu_solve(_lib.SOLVER_PARTITION, rows, periodic, a, b, c, f, tile)
        err = np.max(np.abs(y - ref)) / np.max(np.abs(ref))
        assert err < 1e-13, (tile, err)
