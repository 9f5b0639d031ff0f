"""The morphodynamic outer loop on the device (Morfo70::Run, src/Morfo70.cpp:90-218) over the
C ABI (m70o_*): wave field -> hydro Run -> current feedback -> sediment fluxes -> bed update
per outer step, with the reference's printed change maxima kept as per-step records and its
snapshots (saveTimeStep) copied to pinned host memory asynchronously.
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

from . import _lib

RECORD_FIELDS = ("time", "H_off", "dE", "dR", "dU", "dV", "dZs", "dh")


@dataclasses.dataclass
class OuterParams:
    """Morfo70::H_off/T_off/Theta_off (radians), Parameters::Run::*, Parameters::Hydro::run*
    and the Options switches the loop reads (include/OptionsParametersConstants.h). Defaults:
    src/OptionsParametersConstants.cpp (runIterations 1, dt 0.1, bed evolution off)."""

    H_off: float = 1.0
    T_off: float = 8.0
    Theta_off: float = 0.0
    warmTime: float = 0.0
    bedStartTime: float = 0.0
    bedStopTime: float = 1e30
    dt: float = 0.1
    saveIntervalHydro: float = 0.0
    saveIntervalMorfo: float = 0.0
    runIterations: int = 1
    runIterationsInBedEvolution: int = 1
    calculate_hydro: int = 1
    wave_current_interaction: int = 0
    calculate_sediment: int = 1
    calculate_bed_evolution: int = 0


class _OuterParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("H_off", "T_off", "Theta_off", "warmTime", "bedStartTime",
                                               "bedStopTime", "dt", "saveIntervalHydro", "saveIntervalMorfo")] + \
               [(n, ctypes.c_int) for n in ("runIterations", "runIterationsInBedEvolution", "calculate_hydro",
                                            "wave_current_interaction", "calculate_sediment",
                                            "calculate_bed_evolution")]


class Morfo70Run:
    """Morfo70::Run on device handles: hydro (morfo70_amd.hydro.Hydro), wave
    (morfo70_amd.wave.Wave, coupled to the hydro here), sediment (morfo70_amd.sediment.Sediment
    or None)."""

    def __init__(self, hydro, wave, sediment=None, params: OuterParams | None = None, time0: float = 0.0,
                 max_snapshots: int = 2):
        self._L = _lib.load()
        self.hydro, self.wave, self.sediment = hydro, wave, sediment
        self.params = params or OuterParams()
        if wave.hydro is not hydro:
            wave.couple(hydro)
        self._p = _OuterParams(**dataclasses.asdict(self.params))
        o = ctypes.c_void_p()
        rc = self._L.m70o_create(hydro.handle, wave.handle, sediment._s if sediment is not None else None,
                                 ctypes.cast(ctypes.pointer(self._p), ctypes.c_void_p), float(time0),
                                 int(max_snapshots), ctypes.byref(o))
        if rc != _lib.M70H_OK:
            raise RuntimeError(f"m70o_create: {self._L.m70o_last_error(None).decode()} (rc={rc})")
        self._o = o
        self.nx, self.ny = hydro.nx, hydro.ny

    def _chk(self, rc):
        if rc != _lib.M70H_OK:
            raise RuntimeError(f"m70o: {self._L.m70o_last_error(self._o).decode()} (rc={rc})")

    def close(self):
        if getattr(self, "_o", None):
            self._L.m70o_destroy(self._o)
            self._o = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, stopTime: float) -> int:
        n = ctypes.c_long()
        self._chk(self._L.m70o_run(self._o, float(stopTime), ctypes.byref(n)))
        return n.value

    @property
    def time(self) -> float:
        t = ctypes.c_double()
        self._chk(self._L.m70o_time(self._o, ctypes.byref(t)))
        return t.value

    def records(self) -> np.ndarray:
        """(steps, 8): time, offshore height, dE, dR, dU, dV, dZs, dh (src/Morfo70.cpp:194)."""
        n = ctypes.c_long()
        self._chk(self._L.m70o_records(self._o, ctypes.byref(n), None))
        out = np.zeros((n.value, 8))
        if n.value:
            self._chk(self._L.m70o_records(self._o, ctypes.byref(n), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def snapshot_count(self) -> int:
        n = ctypes.c_long()
        self._chk(self._L.m70o_snapshot_count(self._o, ctypes.byref(n)))
        return n.value

    def snapshot(self, i: int, field: str):
        rows = self.nx + 1 if field in ("U_full", "V_full", "Qx_full", "Qy_full") else self.nx
        out = np.zeros((rows, self.ny))
        t = ctypes.c_double()
        self._chk(self._L.m70o_snapshot(self._o, int(i), ctypes.byref(t), field.encode(),
                                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return t.value, out


def save_snapshot(run: Morfo70Run, i: int, path: str, x, y, Z):
    """Morfo70::saveTimeStep (src/Morfo70.cpp:225-276) of snapshot i to an m70_*.nc file."""
    L = run._L
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, Z)]
    rc = L.m70o_save_snapshot(run._o, int(i), str(path).encode(),
                              *[a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for a in arrs])
    if rc != _lib.M70H_OK:
        raise RuntimeError(f"m70o_save_snapshot: {L.m70nc_last_error().decode()} (rc={rc})")


def resume(path: str, hydro, wave=None, sediment=None, wave_current_interaction: bool = False) -> float:
    """Morfo70::resume's reads (src/Morfo70.cpp:298-312) into the device handles; returns
    the file's time. The hydro's host mirrors are refreshed."""
    L = _lib.load()
    t = ctypes.c_double()
    rc = L.m70_resume(str(path).encode(), hydro.handle, wave.handle if wave is not None else None,
                      sediment._s if sediment is not None else None, int(bool(wave_current_interaction)),
                      ctypes.byref(t))
    if rc != _lib.M70H_OK:
        raise RuntimeError(f"m70_resume: {L.m70nc_last_error().decode()} (rc={rc})")
    hydro.pull()
    return t.value
