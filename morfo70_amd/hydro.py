"""Python mirror of the reference's Hydro call surface over the C ABI (include/m70h.h).

The reference (include/Hydro.h:24-478) is driven as
    Grid* grid = new Grid(batFile); Bathymetry* bathy = new Bathymetry(batFile, zs0, grid);
    Wave* wave = new Wave(grid, bathy); Hydro* hydro = new Hydro(grid, bathy, wave);
    ... wave fills Sxx/Sxy/Syy/Urms/Dw/D_r/H ...; hydro->Run(nIter); read hydro->Zs ...
(src/main.cpp:17-21, src/Morfo70.cpp:141-153). The classes below keep those names,
members and semantics; the state lives on the GPU and the public arrays are host
mirrors refreshed after every Run, as the reference's public pointers are.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "float64 C-contiguous arrays only"
    return a.ctypes.data_as(_dp)


class Grid:
    """Grid metric holder (include/Grid.h:21-54): x, dx, dy, nx, ny, size."""

    def __init__(self, x, ny: int, dy: float):
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.nx = int(self.x.shape[0])
        self.ny = int(ny)
        self.dy = float(dy)
        self.size = self.nx * self.ny
        self.dx = np.diff(self.x)                                     # src/Grid.cpp:32-34


class Bathymetry:
    """Bathymetry (include/Bathymetry.h:22-53): Z, h, Zb = Z - h and
    D = max(Zb + zs0, Dmin) (src/Bathymetry.cpp:33-39)."""

    def __init__(self, grid: Grid, Z, h, zs0: float, Dmin: float):
        shape = (grid.nx, grid.ny)
        self.Z = np.ascontiguousarray(np.asarray(Z, dtype=np.float64).reshape(shape))
        self.h = np.ascontiguousarray(np.asarray(h, dtype=np.float64).reshape(shape))
        self.Zb = self.Z - self.h
        d = self.Zb + zs0
        self.D = np.ascontiguousarray(np.where(d < Dmin, Dmin, d))


class Wave:
    """The Wave members Hydro reads (include/Wave.h:36-53)."""

    FIELDS = ("Sxx", "Sxy", "Syy", "Urms", "Dw", "D_r", "H")

    def __init__(self, grid: Grid, forcing: dict | None = None):
        for k in self.FIELDS:
            setattr(self, k, np.zeros((grid.nx, grid.ny)))
        if forcing:
            for k, v in forcing.items():
                setattr(self, k, np.ascontiguousarray(v, dtype=np.float64))


STATE = ("Zs", "D", "Qx", "Qy", "U", "V", "U_c", "V_c", "dQx_dx", "dQy_dy")
PLUS = ("Qx", "Qy", "U", "V")


class Hydro:
    """Drop-in for the reference Hydro: Hydro(grid, bathy, wave, params), Run(nIter),
    public arrays U, U_c, V, V_c, Zs, Qx, Qy, dQx_dx, dQy_dy, and bathy.D updated in place.
    """

    def __init__(self, grid: Grid, bathy: Bathymetry, wave: Wave, params, solver: int = 0):
        self._L = _lib.load()
        self.grid, self.bathy, self.wave = grid, bathy, wave
        self._params = _lib.Params(*params.as_tuple()) if hasattr(params, "as_tuple") else params
        h = ctypes.c_void_p()
        rc = self._L.m70h_create(grid.nx, grid.ny, _ptr(grid.x), grid.dy,
                                 ctypes.cast(ctypes.pointer(self._params), ctypes.c_void_p),
                                 ctypes.byref(h))
        _lib.check(self._L, None, rc)
        self._h = h
        self.nx, self.ny = grid.nx, grid.ny
        self._chk(self._L.m70h_set_solver(self._h, solver))
        # src/Hydro.cpp:25-37: zero state, Zs = D - Z + h
        self._chk(self._L.m70h_init_from_bathymetry(self._h, _ptr(bathy.D), _ptr(bathy.Z),
                                                     _ptr(bathy.h)))
        for k in STATE:
            if k != "D":
                setattr(self, k, np.zeros(self._shape(k)))
        self.pull()

    # -- plumbing -----------------------------------------------------------------
    def _chk(self, rc):
        return _lib.check(self._L, self._h, rc)

    def _shape(self, k):
        return (self.nx + 1, self.ny) if k in PLUS else (self.nx, self.ny)

    def close(self):
        if getattr(self, "_h", None):
            self._L.m70h_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def push(self):
        """Host mirrors -> device (the reference lets callers write hydro->Zs and
        bathy->D between Runs, e.g. src/Sediment.cpp:128,137,148)."""
        arrs = [np.ascontiguousarray(self.bathy.D if k == "D" else getattr(self, k)) for k in STATE]
        self._chk(self._L.m70h_set_state(self._h, *[_ptr(a) for a in arrs]))

    def pull(self):
        """Device -> host mirrors."""
        out = {k: np.zeros(self._shape(k)) for k in STATE}
        self._chk(self._L.m70h_get_state(self._h, *[_ptr(out[k]) for k in STATE]))
        for k in STATE:
            if k == "D":
                self.bathy.D[...] = out[k]
            else:
                setattr(self, k, out[k])

    def set_forcing(self, wave: Wave | None = None):
        wave = wave or self.wave
        arrs = [np.ascontiguousarray(getattr(wave, k), dtype=np.float64) for k in Wave.FIELDS]
        self._chk(self._L.m70h_set_forcing(self._h, *[_ptr(a) for a in arrs]))

    def set_offshore_f0(self, f0: float):
        self._chk(self._L.m70h_set_offshore_f0(self._h, float(f0)))

    def set_offshore_f0_series(self, series):
        s = np.ascontiguousarray(series, dtype=np.float64)
        self._chk(self._L.m70h_set_offshore_f0_series(self._h, _ptr(s), s.size))

    def bc_constants(self):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._chk(self._L.m70h_bc_constants(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    # -- the reference call surface ------------------------------------------------
    def Run(self, nIter: int, push: bool = True, pull: bool = True):
        """Hydro::Run (include/Hydro.h:124): forcing from self.wave, nIter steps."""
        if push:
            self.push()
        self.set_forcing()
        rc = self._L.m70h_run(self._h, int(nIter))
        # a non-finite state is still the Run's result: mirror it, then report it
        if pull and rc in (_lib.M70H_OK, _lib.M70H_ENONFINITE):
            self.pull()
        self._chk(rc)

    def setHydroFromResumeFile(self, resumeFile: str):
        """Hydro::setHydroFromResumeFile (include/Hydro.h:117; src/Hydro.cpp:204-227): the
        state of an m70_*.nc snapshot into the public arrays and bathy.D; the next Run
        uploads them."""
        from . import netcdf
        nx, ny = self.nx, self.ny
        self.U_c = netcdf.read(resumeFile, "U", (nx, ny))
        self.V_c = netcdf.read(resumeFile, "V", (nx, ny))
        for full, row0, rest in (("U", "U0", "U_x"), ("Qx", "Qx0", "Qx_x"), ("V", "V0", "V_y"),
                                 ("Qy", "Qy0", "Qy_y")):
            setattr(self, full, np.vstack([netcdf.read(resumeFile, row0, (1, ny)),
                                           netcdf.read(resumeFile, rest, (nx, ny))]))
        for k in ("Zs", "dQx_dx", "dQy_dy"):
            setattr(self, k, netcdf.read(resumeFile, k, (nx, ny)))
        self.bathy.D[...] = netcdf.read(resumeFile, "D", (nx, ny))

    # -- device-resident use (benchmarks, batched drivers) --------------------------
    def run_async(self, nIter: int):
        self._chk(self._L.m70h_run_async(self._h, int(nIter)))

    def synchronize(self):
        self._chk(self._L.m70h_synchronize(self._h))

    def check_finite(self) -> bool:
        ok = ctypes.c_int(0)
        rc = self._L.m70h_check_finite(self._h, ctypes.byref(ok))
        if rc not in (_lib.M70H_OK, _lib.M70H_ENONFINITE):
            self._chk(rc)
        return bool(ok.value)

    def set_stream(self, stream_ptr: int | None):
        self._chk(self._L.m70h_set_stream(self._h, ctypes.c_void_p(stream_ptr or 0)))
        self._stream = stream_ptr or 0

    def _stream_ptr(self) -> int:
        return getattr(self, "_stream", 0)

    def set_profiling(self, on: bool, only: str | None = None):
        """Per-kernel HIP-event timing; `only` restricts it to the launches of one kernel."""
        if on and only is not None:
            self._chk(self._L.m70h_set_profiling_only(self._h, only.encode()))
        else:
            self._chk(self._L.m70h_set_profiling(self._h, 1 if on else 0))

    def profile(self) -> dict:
        out = {}
        k = 0
        while True:
            name = ctypes.c_char_p()
            ms = ctypes.c_double()
            n = ctypes.c_long()
            rc = self._L.m70h_profile(self._h, k, ctypes.byref(name), ctypes.byref(ms), ctypes.byref(n))
            if rc == _lib.M70H_EINVAL:
                break
            self._chk(rc)
            out[name.value.decode()] = (ms.value, n.value)
            k += 1
        return out

    def state_ptr(self, name: str) -> int:
        p = ctypes.c_void_p()
        self._chk(self._L.m70h_state_ptr(self._h, name.encode(), ctypes.byref(p)))
        return p.value

    def state(self) -> dict:
        self.pull()
        d = {k: getattr(self, k) for k in STATE if k != "D"}
        d["D"] = self.bathy.D.copy()
        return d


def from_scenario(sc, solver: int = 0) -> Hydro:
    """Builds Grid/Bathymetry/Wave/Hydro for a scenarios.Scenario."""
    grid = Grid(sc.x, sc.ny, sc.dy)
    bathy = Bathymetry(grid, sc.Z, sc.h, sc.zs0, sc.physics.Dmin)
    wave = Wave(grid, sc.forcing)
    hy = Hydro(grid, bathy, wave, sc.physics, solver=solver)
    if sc.f0_series is not None:
        hy.set_offshore_f0_series(sc.f0_series)
    return hy
