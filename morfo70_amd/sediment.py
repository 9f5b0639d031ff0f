"""Python mirror of the reference's Sediment class over the C ABI (include/m70h.h, m70s_*).

The reference (include/Sediment.h:17-115, src/Sediment.cpp) is driven from the outer loop
of src/Morfo70.cpp:141-163 as
    hydro->Run(n); sediment->calculateSedimentFluxes(); if (bed) sediment->updateBottom();
It reads the hydro's Zs, D, U, V and the wave's Dw, owns the sediment arrays and mutates
the bathymetry's h and Zb and the hydro's Zs. Here all of that lives on the GPU beside the
hydro handle's state; the public arrays are host mirrors refreshed by pull().
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .hydro import Hydro, _ptr

FIELDS = ("Alpha", "gamma", "Gamma", "qx", "qy", "dh_dt")


class _SedParams(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in
                ("p", "alpha0_max", "alpha0_min", "D_alpha0_max", "D_alpha0_min", "gamma0", "Gamma0", "A_swash")]


class Sediment:
    """Drop-in for the reference Sediment: Sediment(hydro, wave, bathy, grid, params),
    calculateSedimentFluxes(), updateBottom(dt), setSedimentFluxes(qx, qy); public arrays
    Alpha, gamma, Gamma, qx, qy, dh_dt; bathy.h, bathy.Zb and hydro.Zs updated in place.

    dt is Parameters::Run::dt (0.1 s by default, src/OptionsParametersConstants.cpp).
    """

    def __init__(self, hydro: Hydro, wave=None, bathy=None, grid=None, params=None):
        from .scenarios import SedimentParams
        self._L = _lib.load()
        self.hydro = hydro
        self.wave = wave or hydro.wave
        self.bathy = bathy or hydro.bathy
        self.grid = grid or hydro.grid
        self.params = params or SedimentParams()
        self._p = _SedParams(*self.params.as_tuple())
        s = ctypes.c_void_p()
        h = np.ascontiguousarray(self.bathy.h, dtype=np.float64)
        zb = np.ascontiguousarray(self.bathy.Zb, dtype=np.float64)
        rc = self._L.m70s_create(hydro.handle, _ptr(h), _ptr(zb), ctypes.cast(ctypes.pointer(self._p), ctypes.c_void_p),
                                 ctypes.byref(s))
        if rc != _lib.M70H_OK:
            raise RuntimeError(f"m70s_create: {self._L.m70s_last_error(None).decode()} (rc={rc})")
        self._s = s
        shape = (self.grid.nx, self.grid.ny)
        for k in FIELDS:
            setattr(self, k, np.zeros(shape))

    # -- plumbing -----------------------------------------------------------------
    def _chk(self, rc):
        if rc != _lib.M70H_OK:
            raise RuntimeError(f"m70s: {self._L.m70s_last_error(self._s).decode()} (rc={rc})")

    def close(self):
        if getattr(self, "_s", None):
            self._L.m70s_destroy(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pull(self):
        """Device -> host mirrors of the sediment arrays and the bed (bathy.h, bathy.Zb)."""
        shape = (self.grid.nx, self.grid.ny)
        out = [np.zeros(shape) for _ in range(8)]
        self._chk(self._L.m70s_get(self._s, *[_ptr(a) for a in out]))
        for k, a in zip(FIELDS, out[:6]):
            setattr(self, k, a)
        self.bathy.h = out[6]
        self.bathy.Zb = out[7]

    def push_bed(self):
        """bathy.h / bathy.Zb host mirrors -> device (callers' edits between calls)."""
        h = np.ascontiguousarray(self.bathy.h, dtype=np.float64)
        zb = np.ascontiguousarray(self.bathy.Zb, dtype=np.float64)
        self._chk(self._L.m70s_set_bed(self._s, _ptr(h), _ptr(zb)))

    # -- the reference call surface ------------------------------------------------
    def calculateSedimentFluxes(self, pull: bool = True):
        """Sediment::calculateSedimentFluxes (src/Sediment.cpp:47-65): Dw from the hydro's
        forcing, Zs/D/U/V from its state."""
        self._chk(self._L.m70s_calculate_fluxes(self._s))
        if pull:
            self.pull()

    def updateBottom(self, dt: float = 0.1, pull: bool = True):
        """Sediment::updateBottom (src/Sediment.cpp:67-75,105-151). Moves the hydro's Zs with
        the bed, so the hydro's host mirrors are refreshed too."""
        self._chk(self._L.m70s_update_bottom(self._s, float(dt)))
        if pull:
            self.pull()
            self.hydro.pull()

    def setSedimentFluxesFromResumeFile(self, resumeFile: str, pull: bool = True):
        """Sediment::setSedimentFluxesFromResumeFile (include/Sediment.h:60;
        src/Sediment.cpp:77-103): qx, qy of an m70_*.nc snapshot and the dh_dt they imply."""
        from . import netcdf
        shape = (self.hydro.nx, self.hydro.ny)
        self.setSedimentFluxes(netcdf.read(resumeFile, "qx", shape), netcdf.read(resumeFile, "qy", shape), pull)

    def setSedimentFluxes(self, qx, qy, pull: bool = True):
        """The arithmetic of Sediment::setSedimentFluxesFromResumeFile (src/Sediment.cpp:77-103)
        on given qx, qy (the NetCDF read stays with the caller)."""
        qx = np.ascontiguousarray(qx, dtype=np.float64)
        qy = np.ascontiguousarray(qy, dtype=np.float64)
        self._chk(self._L.m70s_set_fluxes(self._s, _ptr(qx), _ptr(qy)))
        if pull:
            self.pull()
