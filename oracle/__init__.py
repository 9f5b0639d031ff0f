"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
package; the product (morfo70_amd) never does. See m70_oracle.h for what the oracle
restates.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")
REF_LIB_PATH = os.path.join(_HERE, "_ref", "libm70ref.so")

_dp = ctypes.POINTER(ctypes.c_double)
_lib = None


class _Params(ctypes.Structure):
    _fields_ = [("drag_model", ctypes.c_int), ("visc_model", ctypes.c_int),
                ("offshore_bc", ctypes.c_int), ("dt", ctypes.c_double), ("Dmin", ctypes.c_double),
                ("alpha_f", ctypes.c_double), ("cD0", ctypes.c_double), ("nu0", ctypes.c_double),
                ("M", ctypes.c_double), ("ka", ctypes.c_double), ("z0", ctypes.c_double),
                ("alpha_r", ctypes.c_double)]


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _bind(L):
    """Sets the argument and result types of the oracle's API on a loaded build."""
    L.orc_create.restype = ctypes.c_void_p
    L.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double,
                             ctypes.POINTER(_Params), _dp, _dp, ctypes.c_double]
    L.orc_destroy.argtypes = [ctypes.c_void_p]
    L.orc_set_threads.argtypes = [ctypes.c_int]
    L.orc_set_forcing.argtypes = [ctypes.c_void_p] + [_dp] * 7
    L.orc_set_f0.argtypes = [ctypes.c_void_p, ctypes.c_double]
    L.orc_set_f0_series.argtypes = [ctypes.c_void_p, _dp, ctypes.c_long]
    L.orc_bc_constants.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
    L.orc_get_state.argtypes = [ctypes.c_void_p] + [_dp] * 10
    L.orc_set_state.argtypes = [ctypes.c_void_p] + [_dp] * 10
    L.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.orc_run.restype = ctypes.c_int
    L.orc_get_field.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _dp]
    L.orc_get_field.restype = ctypes.c_long
    L.orc_grid_metrics.argtypes = [ctypes.c_int, _dp, _dp]
    L.orc_tridiag.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int]
    L.orc_tridiag_periodic.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int, _dp]
    L.orc_wet_rule.argtypes = [ctypes.c_double] * 6 + [ctypes.c_double]
    L.orc_wet_rule.restype = ctypes.c_int
    return L


def load(path: str):
    """A build of the oracle at `path` (the default one, an older copy, or a mutant)."""
    return _bind(ctypes.CDLL(path))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = load(LIB_PATH)
    return _lib


MUTANT_COUNT = 22
MUTANT_DIR = os.path.join(_HERE, "build", "mutants")


def mutant_lib(k: int):
    """The oracle built with -DORC_MUTANT=k: term k of the assembly scaled by 1.1 (see the
    mutation hooks in m70_oracle.c). Built on first use by `make mutants`."""
    path = os.path.join(MUTANT_DIR, f"liboracle_mut{k}.so")
    if not os.path.exists(path):
        import subprocess
        subprocess.run(["make", "-s", "-C", _HERE, "mutants"], check=True)
    return load(path)


def ref_lib():
    """The reference's own Math.cpp/MeshOperations.h built by oracle/Makefile, or None."""
    if not os.path.exists(REF_LIB_PATH):
        return None
    L = ctypes.CDLL(REF_LIB_PATH)
    L.ref_tridiag.argtypes = [_dp] * 6 + [ctypes.c_int, ctypes.c_int]
    L.ref_tridiag_periodic.argtypes = [_dp] * 4 + [ctypes.c_int, _dp]
    L.ref_tridiag_periodic_batched.argtypes = [_dp] * 5 + [ctypes.c_int, ctypes.c_int]
    L.ref_c2x.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int]
    L.ref_x2c.argtypes = [_dp] * 4 + [ctypes.c_int, ctypes.c_int]
    return L


def ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


STATE_FIELDS = ("Zs", "D", "Qx", "Qy", "U", "V", "U_c", "V_c", "dQx_dx", "dQy_dy")
PLUS_FIELDS = ("Qx", "Qy", "U", "V")
FORCING_FIELDS = ("Sxx", "Sxy", "Syy", "Urms", "Dw", "D_r", "H")


def params_struct(physics) -> _Params:
    return _Params(*physics.as_tuple())


def grid_metrics(x: np.ndarray) -> dict:
    nx = x.shape[0]
    out = np.zeros(20 * nx, dtype=np.float64)
    lib().orc_grid_metrics(nx, ptr(np.ascontiguousarray(x, dtype=np.float64)), ptr(out))
    names = [("dx", nx - 1), ("dx_x", nx)] + [(n, nx - 1) for n in (
        "difx_x_up", "difx_x_d", "difx_x_low", "dif2x_x_up", "dif2x_x_d", "dif2x_x_low")] + \
        [(n, nx) for n in ("difx_c_up", "difx_c_d", "difx_c_low", "dif2x_c_up", "dif2x_c_d",
                           "dif2x_c_low", "x2c_up", "x2c_d")]
    res, off = {}, 0
    for n, ln in names:
        res[n] = out[off:off + ln].copy()
        off += ln
    return res


class OracleHydro:
    """The reference Hydro object restated on the CPU (see m70_oracle.h)."""

    def __init__(self, sc, threads: int = 0, library=None):
        L = self._L = library if library is not None else lib()
        if threads:
            L.orc_set_threads(threads)
        self.nx, self.ny = sc.nx, sc.ny
        self._p = params_struct(sc.physics)
        x = np.ascontiguousarray(sc.x, dtype=np.float64)
        Z = np.ascontiguousarray(sc.Z, dtype=np.float64)
        h = np.ascontiguousarray(sc.h, dtype=np.float64)
        self.h = L.orc_create(sc.nx, sc.ny, ptr(x), sc.dy, ctypes.byref(self._p), ptr(Z), ptr(h),
                              sc.zs0)
        if not self.h:
            raise ValueError("orc_create failed")
        self.set_forcing(sc.forcing)
        if sc.f0_series is not None:
            self.set_f0_series(sc.f0_series)

    def __del__(self):
        if getattr(self, "h", None):
            self._L.orc_destroy(self.h)
            self.h = None

    def set_forcing(self, f: dict):
        arrs = [np.ascontiguousarray(f[k], dtype=np.float64) for k in FORCING_FIELDS]
        self._L.orc_set_forcing(self.h, *[ptr(a) for a in arrs])

    def set_f0(self, f0: float):
        self._L.orc_set_f0(self.h, f0)

    def set_f0_series(self, s):
        s = np.ascontiguousarray(s, dtype=np.float64)
        self._f0 = s
        self._L.orc_set_f0_series(self.h, ptr(s), s.size)

    def bc_constants(self):
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._L.orc_bc_constants(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value

    def run(self, n: int):
        self._L.orc_run(self.h, int(n))

    def _shape(self, name):
        return (self.nx + 1, self.ny) if name in PLUS_FIELDS else (self.nx, self.ny)

    def get_state(self) -> dict:
        out = {k: np.zeros(self._shape(k)) for k in STATE_FIELDS}
        self._L.orc_get_state(self.h, *[ptr(out[k]) for k in STATE_FIELDS])
        return out

    def set_state(self, st: dict):
        arrs = []
        for k in STATE_FIELDS:
            a = st.get(k)
            arrs.append(None if a is None else np.ascontiguousarray(a, dtype=np.float64))
        self._L.orc_set_state(self.h, *[ptr(a) if a is not None else None for a in arrs])

    def field(self, name: str) -> np.ndarray:
        n = self._L.orc_get_field(self.h, name.encode(), None)
        if n < 0:
            raise KeyError(name)
        out = np.zeros(n)
        self._L.orc_get_field(self.h, name.encode(), ptr(out))
        return out.reshape(-1, self.ny)


def tridiag(a, b, c, f):
    """Batched Thomas (n, m) arrays as src/Math.cpp:15-65: a=diag, b=low, c=up."""
    a, b, c, f = (np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, c, f))
    n, m = a.shape
    Y = np.zeros((n, m))
    V = np.zeros((n, m))
    lib().orc_tridiag(ptr(Y), ptr(V), ptr(a), ptr(b), ptr(c), ptr(f), n, m)
    return Y


def tridiag_periodic(a, b, c, f):
    a, b, c, f = (np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, c, f))
    y = np.zeros(a.shape[0])
    lib().orc_tridiag_periodic(ptr(a), ptr(b), ptr(c), ptr(f), a.shape[0], ptr(y))
    return y


def wet_rule(d, dup, zs, zsup, s, sup, Dmin):
    return lib().orc_wet_rule(d, dup, zs, zsup, s, sup, Dmin)


# ------------------------------------------------------------------ tangent-linear model
TLM_LIB_PATH = os.path.join(_HERE, "build", "liboracle_tlm.so")
_tlm = None


class _Dual(ctypes.Structure):
    _fields_ = [("v", ctypes.c_double), ("d", ctypes.c_double)]


class _DualParams(ctypes.Structure):
    _fields_ = [("drag_model", ctypes.c_int), ("visc_model", ctypes.c_int),
                ("offshore_bc", ctypes.c_int)] + [(n, _Dual) for n in (
                    "dt", "Dmin", "alpha_f", "cD0", "nu0", "M", "ka", "z0", "alpha_r")]


def tlm_lib():
    """The tangent-linear build (tlm.cpp): the oracle's API over (value, tangent) pairs."""
    global _tlm
    if _tlm is None:
        if not os.path.exists(TLM_LIB_PATH):
            build()
        L = ctypes.CDLL(TLM_LIB_PATH)
        vp = ctypes.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, vp, _Dual, ctypes.POINTER(_DualParams), vp, vp, _Dual]
        L.orc_destroy.argtypes = [vp]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_forcing.argtypes = [vp] + [vp] * 7
        L.orc_set_f0_series.argtypes = [vp, vp, ctypes.c_long]
        L.orc_get_state.argtypes = [vp] + [vp] * 10
        L.orc_set_state.argtypes = [vp] + [vp] * 10
        L.orc_run.argtypes = [vp, ctypes.c_int]
        L.orc_run.restype = ctypes.c_int
        _tlm = L
    return _tlm


def _pair(v, d=None) -> np.ndarray:
    v = np.asarray(v, dtype=np.float64)
    out = np.zeros(v.shape + (2,))
    out[..., 0] = v
    if d is not None:
        out[..., 1] = d
    return np.ascontiguousarray(out)


def _vp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"] and a.shape[-1] == 2
    return ctypes.c_void_p(a.ctypes.data)


class TangentHydro:
    """Tangent-linear model of the oracle: runs the step on (value, tangent) pairs.

    ``dh`` is the tangent of the bar field h (bathymetry perturbation, Z - h = bed
    level) and ``dparams`` maps parameter names (cD0, nu0, M, ...) to their tangents.
    The initial depth D = max(Z - h + zs0, Dmin) and surface Zs = D - Z + h are formed
    inside the model from h, so their tangents follow from dh. After run(n),
    get_state() returns (values, tangents) per state field."""

    def __init__(self, sc, dh=None, dparams: dict | None = None, threads: int = 0):
        L = tlm_lib()
        if threads:
            L.orc_set_threads(threads)
        self.nx, self.ny = sc.nx, sc.ny
        ph = sc.physics
        dparams = dparams or {}
        vals = dict(zip(("dt", "Dmin", "alpha_f", "cD0", "nu0", "M", "ka", "z0", "alpha_r"),
                        ph.as_tuple()[3:]))
        self._p = _DualParams(ph.drag_model, ph.visc_model, ph.offshore_bc,
                              *[_Dual(vals[n], dparams.get(n, 0.0)) for n in vals])
        self._keep = [_pair(sc.x), _pair(sc.Z), _pair(sc.h, dh)]
        self.h = L.orc_create(sc.nx, sc.ny, _vp(self._keep[0]), _Dual(sc.dy, 0.0), ctypes.byref(self._p),
                              _vp(self._keep[1]), _vp(self._keep[2]), _Dual(sc.zs0, 0.0))
        if not self.h:
            raise ValueError("orc_create failed")
        f = [_pair(sc.forcing[k]) for k in FORCING_FIELDS]
        L.orc_set_forcing(self.h, *[_vp(a) for a in f])
        if sc.f0_series is not None:
            self._f0 = _pair(sc.f0_series)
            L.orc_set_f0_series(self.h, _vp(self._f0), self._f0.shape[0])

    def __del__(self):
        if getattr(self, "h", None):
            tlm_lib().orc_destroy(self.h)
            self.h = None

    def run(self, n: int):
        tlm_lib().orc_run(self.h, int(n))

    def set_state(self, values: dict, tangents: dict | None = None):
        """Sets the state fields given in ``values`` with the tangents in ``tangents``
        (missing tangents are zero); fields absent from ``values`` keep theirs."""
        tangents = tangents or {}
        arrs = []
        for k in STATE_FIELDS:
            if k in values:
                arrs.append(_pair(values[k], tangents.get(k)))
            else:
                arrs.append(None)
        self._st = arrs
        tlm_lib().orc_set_state(self.h, *[(_vp(a) if a is not None else None) for a in arrs])

    def get_state(self):
        """(values, tangents): two dicts of the state fields."""
        shp = {k: ((self.nx + 1, self.ny) if k in PLUS_FIELDS else (self.nx, self.ny)) for k in STATE_FIELDS}
        out = {k: np.zeros(shp[k] + (2,)) for k in STATE_FIELDS}
        tlm_lib().orc_get_state(self.h, *[_vp(out[k]) for k in STATE_FIELDS])
        return {k: out[k][..., 0].copy() for k in out}, {k: out[k][..., 1].copy() for k in out}


# ------------------------------------------------------------------------ sediment
class _SedParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("p", "alpha0_max", "alpha0_min", "D_alpha0_max", "D_alpha0_min",
                                               "gamma0", "Gamma0", "A_swash")]


SED_FIELDS = ("Alpha", "gamma", "Gamma", "qx", "qy", "dh_dt", "h", "Zb")


class OracleSediment:
    """Sediment transport and bed update of the oracle (Sediment, src/Sediment.cpp) on an
    OracleHydro: reads its Zs, D, U, V, Dw; update_bottom mutates its Zs."""

    def __init__(self, hydro: OracleHydro, params):
        L = lib()
        if not hasattr(L, "_sed_ready"):
            L.orc_sed_create.restype = ctypes.c_void_p
            L.orc_sed_create.argtypes = [ctypes.c_void_p, ctypes.POINTER(_SedParams)]
            L.orc_sed_destroy.argtypes = [ctypes.c_void_p]
            L.orc_sed_fluxes.argtypes = [ctypes.c_void_p]
            L.orc_sed_update_bottom.argtypes = [ctypes.c_void_p, ctypes.c_double]
            L.orc_sed_set_fluxes.argtypes = [ctypes.c_void_p, _dp, _dp]
            L.orc_sed_get.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _dp]
            L.orc_sed_get.restype = ctypes.c_long
            L._sed_ready = True
        self.hy = hydro
        self._p = _SedParams(*params.as_tuple())
        self.s = L.orc_sed_create(hydro.h, ctypes.byref(self._p))
        if not self.s:
            raise ValueError("orc_sed_create failed")

    def __del__(self):
        if getattr(self, "s", None):
            lib().orc_sed_destroy(self.s)
            self.s = None

    def fluxes(self):
        lib().orc_sed_fluxes(self.s)

    def update_bottom(self, dt: float):
        lib().orc_sed_update_bottom(self.s, float(dt))

    def set_fluxes(self, qx, qy):
        qx, qy = (np.ascontiguousarray(a, dtype=np.float64) for a in (qx, qy))
        lib().orc_sed_set_fluxes(self.s, ptr(qx), ptr(qy))

    def get(self, name: str) -> np.ndarray:
        n = lib().orc_sed_get(self.s, name.encode(), None)
        if n < 0:
            raise KeyError(name)
        out = np.zeros(n)
        lib().orc_sed_get(self.s, name.encode(), ptr(out))
        return out.reshape(-1, self.hy.ny) if n != self.hy.ny else out


# ----------------------------------------------------------------------------- wave
class _WaveParams(ctypes.Structure):
    _fields_ = [("breaking_model", ctypes.c_int), ("friction_model", ctypes.c_int), ("bottom_friction", ctypes.c_int),
                ("rollers", ctypes.c_int), ("max_iterations", ctypes.c_int), ("tolerance", ctypes.c_double)] + \
               [(n, ctypes.c_double) for n in ("B3", "gamma_b", "m_mon", "n_mon", "theta_max", "fw0", "z0",
                                               "sinBeta_r", "alpha_r")]


WAVE_FIELDS = ("K", "Theta", "Cg", "H", "E", "Urms", "Sxx_w", "Sxy_w", "Syy_w", "Dw", "E_r", "D_r", "Sxx_r", "Sxy_r",
               "Syy_r", "Sxx", "Sxy", "Syy", "CosTheta", "SinTheta", "U_w")


def _wave_lib():
    L = lib()
    if not hasattr(L, "_wave_ready"):
        L.orc_wave_create.restype = ctypes.c_void_p
        L.orc_wave_create.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double, ctypes.POINTER(_WaveParams)]
        L.orc_wave_destroy.argtypes = [ctypes.c_void_p]
        L.orc_wave_set_depth.argtypes = [ctypes.c_void_p, _dp]
        L.orc_wave_set_uv.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_wave_calculate.restype = ctypes.c_long
        L.orc_wave_calculate.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_wave_get.restype = ctypes.c_long
        L.orc_wave_get.argtypes = [ctypes.c_void_p, ctypes.c_char_p, _dp]
        L.orc_wave_dispersion.argtypes = [ctypes.c_double] * 4 + [_dp] * 5
        L.orc_wave_breaking.argtypes = [ctypes.POINTER(_WaveParams), ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, _dp, _dp]
        L._wave_ready = True
    return L


class OracleWave:
    """Wave and rollers driver of the oracle (class Wave, src/Wave.cpp; class WaveRow)."""

    def __init__(self, x, ny: int, dy: float, params):
        L = _wave_lib()
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.nx, self.ny = int(self.x.shape[0]), int(ny)
        self._p = _WaveParams(*params.as_tuple())
        self.w = L.orc_wave_create(self.nx, self.ny, ptr(self.x), float(dy), ctypes.byref(self._p))
        if not self.w:
            raise ValueError("orc_wave_create failed")

    def __del__(self):
        if getattr(self, "w", None):
            _wave_lib().orc_wave_destroy(self.w)
            self.w = None

    def set_depth(self, D):
        D = np.ascontiguousarray(D, dtype=np.float64).reshape(self.nx, self.ny)
        _wave_lib().orc_wave_set_depth(self.w, ptr(D))

    def set_uv(self, U, V):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.nx, self.ny)
        V = np.ascontiguousarray(V, dtype=np.float64).reshape(self.nx, self.ny)
        _wave_lib().orc_wave_set_uv(self.w, ptr(U), ptr(V))

    def calculate(self, hOff: float, thetaOff: float, T: float) -> int:
        n = _wave_lib().orc_wave_calculate(self.w, float(hOff), float(thetaOff), float(T))
        if n < 0:
            raise ValueError(f"orc_wave_calculate: {n}")
        return int(n)

    def get(self, name: str) -> np.ndarray:
        out = np.zeros((self.nx, self.ny))
        if _wave_lib().orc_wave_get(self.w, name.encode(), ptr(out)) < 0:
            raise KeyError(name)
        return out

    def fields(self) -> dict:
        return {k: self.get(k) for k in WAVE_FIELDS}


def wave_dispersion(D: float, U_w: float, w: float, z0: float = 0.001) -> dict:
    out = [ctypes.c_double() for _ in range(5)]
    _wave_lib().orc_wave_dispersion(D, U_w, w, z0, *[ctypes.cast(ctypes.byref(o), _dp) for o in out])
    return dict(zip(("lK", "sigma", "C", "Cg", "Urms_H"), (o.value for o in out)))


def wave_breaking(params, H: float, sigma: float, D: float):
    p = _WaveParams(*params.as_tuple())
    dw, ddw = ctypes.c_double(), ctypes.c_double()
    _wave_lib().orc_wave_breaking(ctypes.byref(p), H, sigma, D, ctypes.cast(ctypes.byref(dw), _dp),
                                  ctypes.cast(ctypes.byref(ddw), _dp))
    return dw.value, ddw.value


def host_cpus() -> dict:
    """The host cores this process may run on: the affinity mask, bounded by a cgroup CPU
    quota when one is set (cgroup v2 cpu.max), and the CPU model. `threads` is the
    OpenMP thread count that uses all of them."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
            if q != "max":
                quota = float(q) / float(period)
    except (OSError, ValueError):
        pass
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    threads = aff if quota is None else max(1, min(aff, int(quota)))
    return {"threads": threads, "affinity": aff, "cgroup_quota": quota, "model": model}


def set_omp_binding() -> None:
    """OMP_PROC_BIND=close and OMP_PLACES=cores for the oracle's OpenMP runtime; takes
    effect when the oracle library is first loaded."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
