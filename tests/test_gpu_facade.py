"""The C++ drop-in Hydro (morfo70_amd/host/Hydro.h) driven like the reference's main
loop (Grid, Bathymetry, Wave, Hydro(grid,bathy,wave), Run(n) several times), against
the oracle on the same inputs."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from morfo70_amd import scenarios as S
from morfo70_amd.build import FACADE
from tests.helpers import FIELDS, compare_states, nonuniform_beach

pytestmark = pytest.mark.gpu


def write_input(path, sc, runs, steps, solver, sediment=0, bed_dt=0.1, resume_runs=0):
    ph = sc.physics
    with open(path, "wb") as f:
        np.array([sc.nx, sc.ny, runs, steps, ph.drag_model, ph.visc_model, ph.offshore_bc, solver, sediment,
                  resume_runs], dtype=np.int32).tofile(f)
        np.array([sc.dy, sc.zs0, ph.dt, ph.Dmin, ph.alpha_f, ph.cD0, ph.nu0, ph.M, ph.ka, ph.z0, ph.alpha_r,
                  bed_dt], dtype=np.float64).tofile(f)
        for a in [sc.x, sc.Z, sc.h] + [sc.forcing[k] for k in ("Sxx", "Sxy", "Syy", "Urms", "Dw", "D_r", "H")]:
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)


SED_FIELDS = ("Alpha", "gamma", "Gamma", "qx", "qy", "dh_dt", "h", "Zb")


def read_output(path, nx, ny, sediment=False, resume=False):
    raw = np.fromfile(path, dtype=np.float64)
    out, off = {}, 0
    names = FIELDS + (SED_FIELDS if sediment else ())
    if resume:
        names = tuple(f"cont_{k}" for k in FIELDS) + tuple(f"res_{k}" for k in FIELDS)
        if sediment:
            names += ("cont_qx", "cont_qy", "cont_dh_dt", "res_qx", "res_qy", "res_dh_dt")
    for k in names:
        rows = nx + 1 if k.split("_")[-1] in ("Qx", "Qy", "U", "V") and not k.endswith("_dx") else nx
        out[k] = raw[off:off + rows * ny].reshape(rows, ny)
        off += rows * ny
    assert off == raw.size
    return out


@pytest.mark.parametrize("solver,exact", [(0, True), (2, False)])
def test_facade_matches_oracle(tmp_path, solver, exact):
    import oracle
    assert os.path.exists(FACADE), "build first: python -m morfo70_amd.build"
    ph = S.Physics(drag_model=S.DRAG["constant"], visc_model=S.VISC["constant"], dt=0.25)
    sc = nonuniform_beach(nx=50, ny=44, physics=ph, seed=21)
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(inp, sc, runs=3, steps=7, solver=solver)
    print(f"[facade] solver={solver} 50x44 3x7 steps", flush=True)
    subprocess.run([FACADE, str(inp), str(outp)], check=True, timeout=120)
    got = read_output(outp, sc.nx, sc.ny)
    o = oracle.OracleHydro(sc)
    for _ in range(3):
        o.run(7)
    res = compare_states(got, o.get_state())
    for k, r in res.items():
        assert r["nan"] == 0, k
        assert (r["exact"] if exact else r["rel"] <= 1e-10), (k, r)


def test_facade_sediment_outer_loop(tmp_path):
    """The C++ Sediment (morfo70_amd/host/Sediment.h) in the outer loop: Run, fluxes, bed
    update, three times; the bed update's Zs reaches the next Run through the host mirror.
    Bar: max-norm relative error <= 1e-10 (the coefficients' tanh)."""
    import oracle
    ph = S.Physics(drag_model=S.DRAG["constant"], visc_model=S.VISC["constant"], dt=0.25)
    sc = nonuniform_beach(nx=50, ny=44, physics=ph, seed=22)
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(inp, sc, runs=3, steps=5, solver=2, sediment=1, bed_dt=2.0)
    print("[facade] sediment outer loop 50x44 3x(5 steps + bed)", flush=True)
    subprocess.run([FACADE, str(inp), str(outp)], check=True, timeout=120)
    got = read_output(outp, sc.nx, sc.ny, sediment=True)
    o = oracle.OracleHydro(sc)
    sd = oracle.OracleSediment(o, S.SedimentParams())
    for _ in range(3):
        o.run(5)
        sd.fluxes()
        sd.update_bottom(2.0)
    ref = dict(o.get_state(), **{k: sd.get(k) for k in SED_FIELDS})
    res = compare_states(got, ref, fields=FIELDS + SED_FIELDS)
    for k, r in res.items():
        assert r["nan"] == 0 and r["rel"] <= 1e-10, (k, r)
    assert np.max(np.abs(got["dh_dt"])) > 0


def test_facade_resume_continues_bit_for_bit(tmp_path):
    """Hydro::setHydroFromResumeFile and Sediment::setSedimentFluxesFromResumeFile on the
    facades (include/Hydro.h:117, include/Sediment.h:60): after three Runs with sediment,
    the state is written as a resume file, a fresh Hydro and Sediment read it, and two
    more Runs of the original and of the resumed Hydro agree bit for bit; the resumed
    Sediment holds the original's qx, qy and the dh_dt they imply."""
    ph = S.Physics(drag_model=S.DRAG["constant"], visc_model=S.VISC["constant"], dt=0.25)
    sc = nonuniform_beach(nx=50, ny=44, physics=ph, seed=23)
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(inp, sc, runs=3, steps=5, solver=2, sediment=1, bed_dt=2.0, resume_runs=2)
    print("[facade] resume 50x44 3x(5 steps + bed), then 2x5 steps", flush=True)
    subprocess.run([FACADE, str(inp), str(outp)], check=True, timeout=120)
    got = read_output(outp, sc.nx, sc.ny, sediment=True, resume=True)
    for k in FIELDS:
        assert np.array_equal(got[f"res_{k}"], got[f"cont_{k}"]), k
    assert np.max(np.abs(got["cont_Qx"])) > 0
    for k in ("qx", "qy", "dh_dt"):
        assert np.array_equal(got[f"res_{k}"][1:], got[f"cont_{k}"][1:]), k
    assert np.max(np.abs(got["cont_qx"])) > 0
