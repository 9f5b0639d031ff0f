"""Python mirror of the reference's Wave class over the C ABI (include/m70h.h, m70w_*).

The reference (include/Wave.h:16-128, src/Wave.cpp) is driven from the outer loop of
src/Morfo70.cpp:141-154 as
    wave->calculateWaveField(H_off_t, Theta_off, T_off); hydro->Run(n); wave->setUV(U_c, V_c)
and Hydro reads its public arrays Sxx, Sxy, Syy, Urms, Dw, D_r, H. The field is computed on
the GPU; the public arrays are host mirrors refreshed by pull(). Coupled to a hydro handle
(couple()), the field is computed on the hydro's device depth and pushed into its forcing
without a host round trip.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .hydro import _ptr

FIELDS = ("K", "Theta", "Cg", "H", "E", "Urms", "Sxx_w", "Sxy_w", "Syy_w", "Dw", "E_r", "D_r", "Sxx_r", "Sxy_r",
          "Syy_r", "Sxx", "Sxy", "Syy", "CosTheta", "SinTheta", "U_w")
FORCING = ("Sxx", "Sxy", "Syy", "Urms", "Dw", "D_r", "H")


class _WaveParams(ctypes.Structure):
    _fields_ = [("breaking_model", ctypes.c_int), ("friction_model", ctypes.c_int), ("bottom_friction", ctypes.c_int),
                ("rollers", ctypes.c_int), ("max_iterations", ctypes.c_int)] + \
               [(n, ctypes.c_double) for n in ("tolerance", "B3", "gamma_b", "m_mon", "n_mon", "theta_max", "fw0",
                                               "z0", "sinBeta_r", "alpha_r")]


class Wave:
    """Drop-in for the reference Wave: Wave(grid, bathy, params), calculateWaveField(hOff,
    thetaOff, T), setUV(U, V), and the public arrays (Sxx, Sxy, Syy, Urms, Dw, D_r, H, E, Theta,
    ...). grid needs x, ny, dy; bathy needs D (read at every calculation)."""

    def __init__(self, grid, bathy=None, params=None):
        from .scenarios import WaveParams
        self._L = _lib.load()
        self.grid, self.bathy = grid, bathy
        self.nx, self.ny = int(grid.nx), int(grid.ny)
        self.params = params or WaveParams()
        p = self.params.as_tuple()
        self._p = _WaveParams(*[int(v) for v in p[:5]], *[float(v) for v in p[5:]])
        w = ctypes.c_void_p()
        x = np.ascontiguousarray(grid.x, dtype=np.float64)
        rc = self._L.m70w_create(self.nx, self.ny, _ptr(x), float(grid.dy),
                                 ctypes.cast(ctypes.pointer(self._p), ctypes.c_void_p), ctypes.byref(w))
        if rc != _lib.M70H_OK:
            raise RuntimeError(f"m70w_create: {self._L.m70w_last_error(None).decode()} (rc={rc})")
        self._w = w
        self.hydro = None
        for k in FIELDS:
            setattr(self, k, np.zeros((self.nx, self.ny)))

    def _chk(self, rc):
        if rc != _lib.M70H_OK:
            raise RuntimeError(f"m70w: {self._L.m70w_last_error(self._w).decode()} (rc={rc})")

    def close(self):
        if getattr(self, "_w", None):
            self._L.m70w_destroy(self._w)
            self._w = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._w

    def pull(self, names=FIELDS):
        for k in names:
            out = np.zeros((self.nx, self.ny))
            self._chk(self._L.m70w_get(self._w, k.encode(), _ptr(out)))
            setattr(self, k, out)

    def steps(self) -> int:
        n = ctypes.c_long()
        self._chk(self._L.m70w_steps(self._w, ctypes.byref(n)))
        return n.value

    def set_stream(self, stream_ptr):
        self._chk(self._L.m70w_set_stream(self._w, ctypes.c_void_p(stream_ptr or 0)))

    # -- the reference call surface ------------------------------------------------
    def calculateWaveField(self, hOff: float, thetaOff: float, T: float, pull: bool = True):
        """Wave::calculateWaveField (include/Wave.h:76). Depth from bathy.D, or from the
        coupled hydro's device state."""
        if self.hydro is None:
            D = np.ascontiguousarray(self.bathy.D, dtype=np.float64)
            self._chk(self._L.m70w_set_depth(self._w, _ptr(D)))
        self._chk(self._L.m70w_calculate(self._w, float(hOff), float(thetaOff), float(T)))
        if pull:
            self.pull()

    def setUV(self, U, V):
        """Wave::setUV (include/Wave.h:84)."""
        U = np.ascontiguousarray(U, dtype=np.float64)
        V = np.ascontiguousarray(V, dtype=np.float64)
        self._chk(self._L.m70w_set_uv(self._w, _ptr(U), _ptr(V)))

    # -- device coupling -----------------------------------------------------------------
    def couple(self, hydro):
        """Compute on hydro's device depth and stream; push_forcing() then hands the field to it."""
        self._chk(self._L.m70w_couple_hydro(self._w, hydro.handle))
        self.hydro = hydro

    def setUV_from_hydro(self):
        self._chk(self._L.m70w_set_uv_from_hydro(self._w))

    def push_forcing(self):
        self._chk(self._L.m70w_push_forcing(self._w))

    def forcing(self) -> dict:
        """The arrays Hydro reads, as a dict (for Hydro's host Wave container)."""
        self.pull(FORCING)
        return {k: getattr(self, k) for k in FORCING}
