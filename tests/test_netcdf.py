"""NetCDF classic I/O (m70nc_*, in place of MorfoIO::Netcdf, src/Netcdf.cpp) against scipy's
independent reader on the reference's own input files (tests/golden/*.nc, copied from the
reference's test/input) and on files this library writes."""
from __future__ import annotations

import os

import numpy as np
import pytest
from scipy.io import netcdf_file

from morfo70_amd import netcdf as nc

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name", ["beach_cb_nowall.nc", "beach_cb_wall.nc"])
def test_reads_reference_bathymetry(name):
    path = os.path.join(GOLD, name)
    ref = netcdf_file(path, "r", mmap=False)
    nx, ny = nc.dims(path)
    assert (nx, ny) == (ref.dimensions["x"], ref.dimensions["y"])
    for v, shape in [("x", (nx,)), ("y", (ny,)), ("Z", (nx, ny)), ("h", (nx, ny))]:
        assert np.array_equal(nc.read(path, v, shape), np.asarray(ref.variables[v].data, dtype=np.float64)), v
    ref.close()
    grid, bathy = nc.read_bathymetry(path, Dmin=0.01)
    assert grid.nx == nx and grid.ny == ny
    np.testing.assert_array_equal(bathy.D, np.maximum(bathy.Z - bathy.h, 0.01))


def test_write_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    nx, ny = 7, 5
    x, y = np.cumsum(rng.uniform(1, 2, nx)), np.arange(ny) * 2.0
    vars_ = {"Z": rng.standard_normal((nx, ny)), "U0": rng.standard_normal(ny), "Qx_x": rng.standard_normal((nx, ny))}
    p = str(tmp_path / "m70_000000010.nc")
    nc.write(p, x, y, vars_, {"time": 10.5})
    ref = netcdf_file(p, "r", mmap=False)
    assert ref.version_byte == 1
    assert list(ref.variables) == ["x", "y", "Z", "U0", "Qx_x"]
    for k, v in vars_.items():
        assert np.array_equal(ref.variables[k].data, v), k
        assert np.array_equal(nc.read(p, k, v.shape), v), k
    assert ref._attributes["time"] == 10.5 and nc.read_att(p, "time") == 10.5
    ref.close()


def test_reader_types_and_errors(tmp_path):
    """Other numeric classic types are read as doubles; missing variables and wrong sizes
    are errors, not silent garbage."""
    p = str(tmp_path / "types.nc")
    f = netcdf_file(p, "w", version=2)
    f.createDimension("x", 3)
    f.createDimension("y", 2)
    for k, t in [("f", "f4"), ("i", "i4"), ("s", "i2"), ("b", "i1")]:
        v = f.createVariable(k, t, ("x", "y"))
        v[:] = np.arange(6).reshape(3, 2) - 2
    f.scale = 2.5
    f.close()
    assert nc.dims(p) == (3, 2)
    for k in ("f", "i", "s", "b"):
        np.testing.assert_array_equal(nc.read(p, k, (3, 2)), np.arange(6).reshape(3, 2) - 2.0)
    assert nc.read_att(p, "scale") == 2.5
    with pytest.raises(RuntimeError):
        nc.read(p, "nope", (3, 2))
    with pytest.raises(RuntimeError):
        nc.read(p, "f", (4, 2))
