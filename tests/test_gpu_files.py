"""Snapshot files and resume (m70o_save_snapshot, m70_resume; Morfo70::saveTimeStep and
::resume, src/Morfo70.cpp:225-326) on the GPU, and the hydro step on the reference's own
bathymetry input file."""
from __future__ import annotations

import os

import numpy as np
import pytest
from scipy.io import netcdf_file

from morfo70_amd import scenarios as S
from tests.helpers import compare_states, nonuniform_beach

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SAVED = ["x", "y", "Z", "h", "D", "H", "E", "Cg", "lK", "dQx_dx", "dQy_dy", "Theta", "Dw", "E_r", "U", "U0",
         "Qx0", "U_x", "Qx_x", "V", "V0", "Qy0", "V_y", "Qy_y", "Zs", "qx", "qy", "Gamma_Swash", "Alpha", "gamma"]


def _handles(sc, wp, sp):
    from morfo70_amd.hydro import from_scenario
    from morfo70_amd.sediment import Sediment
    from morfo70_amd.wave import Wave
    h = from_scenario(sc)
    w = Wave(h.grid, h.bathy, wp)
    w.couple(h)
    return h, w, Sediment(h, params=sp)


def test_snapshot_file_and_resume(tmp_path):
    """A snapshot written to an m70_*.nc file holds saveTimeStep's variables in its order
    with the snapshot's values; handles resumed from it continue exactly as the run that
    wrote it (sediment transport off, since the reference does not save the running
    maximum of Dw that the fluxes use)."""
    from morfo70_amd.morfo import Morfo70Run, OuterParams, resume, save_snapshot
    ph = S.Physics(drag_model=S.DRAG["constant"], visc_model=S.VISC["constant"], dt=0.25)
    sc = nonuniform_beach(nx=40, ny=24, physics=ph, seed=21)
    wp, sp = S.WaveParams.base_case(), S.SedimentParams()
    op = OuterParams(H_off=0.8, T_off=8.0, Theta_off=0.1, dt=1.0, saveIntervalHydro=1.0, runIterations=2,
                     calculate_sediment=0)
    print("[files] snapshot + resume 40x24", flush=True)
    hA, wA, sA = _handles(sc, wp, sp)
    runA = Morfo70Run(hA, wA, sA, op, max_snapshots=4)
    runA.run(2.0)
    i = runA.snapshot_count() - 1
    y = np.arange(sc.ny) * sc.dy
    path = str(tmp_path / "m70_000000002.nc")
    save_snapshot(runA, i, path, sc.x, y, sc.Z)
    f = netcdf_file(path, "r", mmap=False)
    assert list(f.variables) == SAVED
    assert f._attributes["time"] == 2.0
    for k, name in [("Zs", "Zs"), ("H", "H"), ("U", "U"), ("lK", "lK")]:
        assert np.array_equal(f.variables[k].data, runA.snapshot(i, name)[1]), k
    Qx = runA.snapshot(i, "Qx_full")[1]
    assert np.array_equal(f.variables["Qx0"].data, Qx[0]) and np.array_equal(f.variables["Qx_x"].data, Qx[1:])
    f.close()
    runA.run(5.0)
    recA = runA.records()
    hB, wB, sB = _handles(sc, wp, sp)
    t0 = resume(path, hB, wB, sB)
    assert t0 == 2.0
    runB = Morfo70Run(hB, wB, sB, op, time0=t0, max_snapshots=4)
    runB.run(5.0)
    recB = runB.records()
    assert np.array_equal(recB, recA[2:])
    for k, r in compare_states(hB.state(), hA.state()).items():
        assert r["exact"], (k, r)


def test_hydro_on_reference_bathymetry():
    """The reference's own input beach (test/input/beach_cb_nowall.nc, 247 x 401), read by
    this library's NetCDF reader, forced by this library's wave driver (base case settings):
    the FUSED step against the oracle after 20 steps, max-norm relative error <= 1e-10."""
    import oracle
    from morfo70_amd import netcdf as nc
    from morfo70_amd.hydro import Hydro, Wave as Forcing
    from morfo70_amd.wave import Wave
    path = os.path.join(GOLD, "beach_cb_nowall.nc")
    ph = S.Physics(dt=0.2)
    grid, bathy = nc.read_bathymetry(path, Dmin=ph.Dmin)
    w = Wave(grid, bathy, S.WaveParams.base_case())
    w.calculateWaveField(1.0, 0.1, 8.0)
    forcing = w.forcing()
    assert all(np.all(np.isfinite(v)) for v in forcing.values())
    print("[files] beach_cb_nowall 247x401, 20 steps", flush=True)
    sc = S.Scenario("beach_cb_nowall", grid.nx, grid.ny, float(np.mean(np.diff(grid.x))), grid.dy, grid.x,
                    bathy.Z, bathy.h, 0.0, ph, forcing, 20)
    o = oracle.OracleHydro(sc)
    o.run(20)
    g = Hydro(grid, bathy, Forcing(grid, forcing), ph, solver=2)
    g.Run(20)
    res = compare_states(g.state(), o.get_state())
    for k, r in res.items():
        assert r["nan"] == 0 and r["rel"] <= 1e-10, (k, r)


def test_python_mirror_resume_members(tmp_path):
    """Hydro.setHydroFromResumeFile and Sediment.setSedimentFluxesFromResumeFile of the
    Python mirror (include/Hydro.h:117, include/Sediment.h:60): a fresh Hydro read from a
    file of the running one's state continues bit for bit; the fresh Sediment holds the
    file's fluxes and the dh_dt they imply, equal to the running Sediment's."""
    from morfo70_amd import _lib, netcdf
    from morfo70_amd.hydro import from_scenario
    from morfo70_amd.sediment import Sediment
    from tests.helpers import nonuniform_beach
    sc = nonuniform_beach(nx=40, ny=36, steps=10, seed=31)
    a = from_scenario(sc, solver=_lib.SOLVER_FUSED)
    sa = Sediment(a)
    a.Run(4)
    sa.calculateSedimentFluxes()
    sa.updateBottom(2.0)
    st = a.state()
    path = str(tmp_path / "m70_resume.nc")
    y = np.arange(sc.ny) * sc.dy
    netcdf.write(path, sc.x, y, {"U": st["U_c"], "U0": st["U"][0], "U_x": st["U"][1:], "Qx0": st["Qx"][0],
                                 "Qx_x": st["Qx"][1:], "V": st["V_c"], "V0": st["V"][0], "V_y": st["V"][1:],
                                 "Qy0": st["Qy"][0], "Qy_y": st["Qy"][1:], "Zs": st["Zs"], "D": st["D"],
                                 "dQx_dx": st["dQx_dx"], "dQy_dy": st["dQy_dy"], "qx": sa.qx, "qy": sa.qy},
                 {"time": 0.0})
    b = from_scenario(sc, solver=_lib.SOLVER_FUSED)
    b.setHydroFromResumeFile(path)
    sb = Sediment(b)
    sb.setSedimentFluxesFromResumeFile(path)
    for k in ("qx", "qy"):
        assert np.array_equal(getattr(sb, k), getattr(sa, k)), k
    assert np.array_equal(sb.dh_dt[1:], sa.dh_dt[1:])
    sb.close()
    sa.close()
    a.Run(5)
    b.Run(5)
    sa_, sb_ = a.state(), b.state()
    for k in sa_:
        assert np.array_equal(sb_[k], sa_[k]), k
    a.close()
    b.close()
