"""Seeded beach scenarios: grid, bathymetry, wave forcing and tide for the Hydro step.

The reference reads its bathymetry from NetCDF (src/Grid.cpp:16-29,
src/Bathymetry.cpp:13-31) and gets its forcing from the row-marching wave driver
(src/Wave.cpp:80-172). Neither is on the Hydro hot path, so this module defines the
inputs in closed form instead. The same arrays feed the HIP path and the CPU oracle.

Conventions (those of the reference):
  * x index 0 is offshore; x increases onshore. ``Z`` is the still-water depth of the
    alongshore-uniform profile (Z > 0 in water, Z < 0 on land), ``h`` is the bed
    perturbation (positive raises the bed), ``Zb = Z - h`` (include/Bathymetry.h:29-32).
  * Arrays are (nx, ny) float64, C order: index i*ny + j, j alongshore and periodic.

Physics choices for inputs the reference does not define:
  * Chezy friction C maps to the constant drag model with cD0 = g / C**2
    (include/OptionsParametersConstants.h:50-54 has no Chezy model).
  * The tide enters through the offshore flux hook Qx0 = f1*Qx1 + f0
    (include/Hydro.h:89-90) with f1 = 1 (ConstantZs) and
    f0(t) = dx_x[0] * d(eta0)/dt, the long-wave boundary condition of the model's
    equations document (eq. 2.93).
  * Wave forcing: linear dispersion, Snell refraction and energy-flux shoaling from the
    offshore row, depth-limited saturation H <= gamma_b * D (gamma_b = 0.475), radiation
    stresses E*((1+cos^2)n - 1/2), E*n*sin*cos, E*((2-cos^2)n - 1/2) (the per-unit-energy
    forms of src/WaveRow.cpp:200-205), breaking dissipation Dw = max(0, -d(E cg cos)/dx),
    rollers off (D_r = 0) and a floor on Urms so that the drag term stays finite.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

G = 9.81          # include/OptionsParametersConstants.h:14
RHO = 1025.0      # include/OptionsParametersConstants.h:16
GAMMA_B = 0.475   # src/OptionsParametersConstants.cpp:48
URMS_FLOOR = 1e-2  # m/s
TIDE_PERIOD = 12.42 * 3600.0

DRAG = {"constant": 0, "logdepth": 1, "manning": 2}
VISC = {"constant": 0, "rollers": 1, "depth": 2, "hrms": 3}
OFFSHORE_BC = {"sponge": 0, "constant_zs": 1}


@dataclasses.dataclass
class Physics:
    """Hydro parameters (Parameters::Hydro::*, Parameters::Bottom::z0,
    Parameters::Rollers::alpha_r and Options::Hydro::* of
    include/OptionsParametersConstants.h:49-113)."""

    drag_model: int = DRAG["constant"]
    visc_model: int = VISC["hrms"]
    offshore_bc: int = OFFSHORE_BC["sponge"]
    dt: float = 0.5
    Dmin: float = 0.01
    alpha_f: float = 1.16
    cD0: float = G / 65.0 ** 2
    nu0: float = 1e-4
    M: float = 1.0
    ka: float = 0.0219
    z0: float = 0.001
    alpha_r: float = 1.0

    def as_tuple(self):
        return (self.drag_model, self.visc_model, self.offshore_bc, self.dt, self.Dmin,
                self.alpha_f, self.cD0, self.nu0, self.M, self.ka, self.z0, self.alpha_r)


@dataclasses.dataclass
class SedimentParams:
    """Parameters::Bottom::p and Parameters::Sediment::CWS / ::Swash
    (include/OptionsParametersConstants.h:84-138), defaults of
    src/OptionsParametersConstants.cpp:44,76-85."""

    p: float = 0.4
    alpha0_max: float = 0.002
    alpha0_min: float = 0.0
    D_alpha0_max: float = 4.0
    D_alpha0_min: float = 10.0
    gamma0: float = 0.5
    Gamma0: float = 1.0
    A_swash: float = 0.3

    def as_tuple(self):
        return (self.p, self.alpha0_max, self.alpha0_min, self.D_alpha0_max, self.D_alpha0_min, self.gamma0,
                self.Gamma0, self.A_swash)


BREAKING = {"ChurchThornton": 1, "ThorntonGuza": 2, "RegularWaves": 3}
FRICTION = {"Constant": 0, "Soulsby97": 1}


@dataclasses.dataclass
class WaveParams:
    """Options::Wave::*, Options::Rollers::CalculateRollers, Parameters::Wave::*,
    Parameters::Rollers::* and Parameters::Bottom::z0 (include/OptionsParametersConstants.h),
    with the compiled-in defaults of src/OptionsParametersConstants.cpp:14-58. theta_max is
    in radians as the reference stores it (its .ini reader converts degrees,
    src/MorfoIniFile.cpp:125-127; the compiled-in 85 is taken as radians)."""

    breaking_model: int = 1
    friction_model: int = 0
    bottom_friction: int = 0
    rollers: int = 0
    max_iterations: int = 10
    tolerance: float = 1e-10
    B3: float = 2.0
    gamma_b: float = 0.475
    m_mon: float = 20.0
    n_mon: float = 30.0
    theta_max: float = 85.0
    fw0: float = 0.01
    z0: float = 0.001
    sinBeta_r: float = 0.05
    alpha_r: float = 1.0

    @classmethod
    def base_case(cls) -> "WaveParams":
        """The wave and roller settings of the reference's test/input/base_case.ini."""
        return cls(breaking_model=BREAKING["ThorntonGuza"], B3=1.0, m_mon=10.0, n_mon=10.0,
                   theta_max=86.0 * math.pi / 180.0, z0=0.01)

    def as_tuple(self):
        return (self.breaking_model, self.friction_model, self.bottom_friction, self.rollers, self.max_iterations,
                self.tolerance, self.B3, self.gamma_b, self.m_mon, self.n_mon, self.theta_max, self.fw0, self.z0,
                self.sinBeta_r, self.alpha_r)


@dataclasses.dataclass
class Scenario:
    name: str
    nx: int
    ny: int
    dx: float
    dy: float
    x: np.ndarray          # (nx,) cross-shore coordinates
    Z: np.ndarray          # (nx, ny) profile depth
    h: np.ndarray          # (nx, ny) bed perturbation
    zs0: float             # initial still-water level
    physics: Physics
    forcing: dict          # Sxx, Sxy, Syy, Urms, Dw, D_r, H -> (nx, ny)
    steps: int
    f0_series: np.ndarray | None = None   # per-step offshore flux offset (tide)
    meta: dict = dataclasses.field(default_factory=dict)

    @property
    def cells(self) -> int:
        return self.nx * self.ny

    def initial_depth(self) -> np.ndarray:
        """D = max(Z - h + zs0, Dmin), as src/Bathymetry.cpp:33-39."""
        d = (self.Z - self.h) + self.zs0
        return np.where(d < self.physics.Dmin, self.physics.Dmin, d)


def _dispersion_k(omega: float, depth: np.ndarray) -> np.ndarray:
    """Wavenumber from omega^2 = g k tanh(k h) by Newton iteration (h > 0)."""
    depth = np.maximum(depth, 1e-3)
    k0 = omega * omega / G
    k = k0 / np.sqrt(np.tanh(k0 * depth))
    for _ in range(12):
        th = np.tanh(k * depth)
        f = G * k * th - omega * omega
        df = G * th + G * k * depth * (1.0 - th * th)
        k = k - f / df
    return k


def _smooth_x(a: np.ndarray, sigma_cells: float) -> np.ndarray:
    """Gaussian smoothing along x (axis 0), edge values held."""
    if sigma_cells <= 0:
        return a
    r = int(math.ceil(3.0 * sigma_cells))
    t = np.arange(-r, r + 1, dtype=np.float64)
    w = np.exp(-0.5 * (t / sigma_cells) ** 2)
    w /= w.sum()
    p = np.concatenate([np.repeat(a[:1], r, axis=0), a, np.repeat(a[-1:], r, axis=0)], axis=0)
    out = np.zeros_like(a)
    for k in range(2 * r + 1):
        out += w[k] * p[k:k + a.shape[0]]
    return out


def wave_forcing(D: np.ndarray, x: np.ndarray, Hrms: float, T: float, theta_deg: float,
                 Dmin: float, spread_m: float = 8.0) -> dict:
    """Closed-form wave forcing on the water depth field D (nx, ny).

    The radiation stresses and the dissipation are spread along x by a Gaussian of
    ``spread_m`` metres, as a surface roller spreads the break-point forcing: the
    unspread forcing has a kink at the break point that, with Hrms above ~1.5 m on
    steep beaches, drives the step unstable within a few hundred steps."""
    omega = 2.0 * math.pi / T
    wet = D > Dmin + 1e-4
    depth = np.where(wet, D, 1e-3)
    k = _dispersion_k(omega, depth)
    kh = k * depth
    c = omega / k
    n = 0.5 * (1.0 + 2.0 * kh / np.sinh(2.0 * kh))
    cg = n * c
    c0, cg0 = c[0:1, :], cg[0:1, :]
    th0 = math.radians(theta_deg)
    sin_t = np.clip(math.sin(th0) * c / c0, -0.999, 0.999)
    cos_t = np.sqrt(1.0 - sin_t * sin_t)
    Hs = Hrms * np.sqrt((cg0 * math.cos(th0)) / (cg * cos_t))
    H = np.minimum(Hs, GAMMA_B * depth)
    H = np.where(wet, H, 0.0)
    E = 0.125 * RHO * G * H * H
    Sxx = E * ((1.0 + cos_t * cos_t) * n - 0.5)
    Syy = E * ((2.0 - cos_t * cos_t) * n - 0.5)
    Sxy = E * (sin_t * cos_t * n)
    Urms = math.pi * H / (T * np.sinh(kh)) / math.sqrt(2.0)
    Urms = np.maximum(np.where(wet, Urms, 0.0), URMS_FLOOR)
    Ef = E * cg * cos_t
    dEf = np.gradient(Ef, x, axis=0)
    Dw = np.where(wet & (Hs > GAMMA_B * depth), np.maximum(-dEf, 0.0), 0.0)
    D_r = np.zeros_like(D)
    sig = spread_m / float(np.mean(np.diff(x))) if x.size > 1 else 0.0
    Sxx, Sxy, Syy, Dw = (np.where(wet, _smooth_x(a, sig), 0.0) for a in (Sxx, Sxy, Syy, Dw))
    f = {"Sxx": Sxx, "Sxy": Sxy, "Syy": Syy, "Urms": Urms, "Dw": Dw, "D_r": D_r, "H": H}
    return {kk: np.ascontiguousarray(v, dtype=np.float64) for kk, v in f.items()}


def _bar_beach(nx, ny, dx, dy, slope, bar_amp, bar_mod, bar_centre, bar_width, wavelength,
               dry_fraction=0.08, phase=0.0):
    x = np.arange(nx, dtype=np.float64) * dx
    y = np.arange(ny, dtype=np.float64) * dy
    x_shore = x[-1] - dry_fraction * x[-1]
    s = x_shore - x                       # distance seaward of the shoreline
    Z = np.repeat((slope * s)[:, None], ny, axis=1)
    amp = bar_amp + bar_mod * np.sin(2.0 * math.pi * y / wavelength + phase)
    h = np.exp(-((s - bar_centre) / bar_width) ** 2)[:, None] * amp[None, :]
    return x, np.ascontiguousarray(Z), np.ascontiguousarray(h)


def _periodic_wavelength(Ly: float, lam: float) -> float:
    m = max(1, int(round(Ly / lam)))
    return Ly / m


def tide_f0_series(steps: int, dt: float, dx0: float, amp: float = 1.0,
                   period: float = TIDE_PERIOD, t0: float = 0.0) -> np.ndarray:
    """f0 for step k: dx_x[0] * d/dt [amp*sin(2 pi t/period)] at t = t0 + k*dt."""
    w = 2.0 * math.pi / period
    t = t0 + np.arange(steps, dtype=np.float64) * dt
    return dx0 * amp * w * np.cos(w * t)


def config1(steps: int = 1000) -> Scenario:
    """Plain Hydro ADI, 120 x 200 cells, dx=dy=5 m: planar beach slope 0.02 plus a
    Gaussian bar (1.0 m, 150 m from shore, width 30 m) modulated alongshore by a 0.3 m
    sinusoid of 250 m wavelength; Hrms 1 m, T 8 s, 10 deg; Chezy 65; dt 0.5 s; sponge."""
    nx, ny, dx, dy = 120, 200, 5.0, 5.0
    x, Z, h = _bar_beach(nx, ny, dx, dy, 0.02, 1.0, 0.3, 150.0, 30.0, 250.0)
    ph = Physics(offshore_bc=OFFSHORE_BC["sponge"], dt=0.5)
    sc = Scenario("beach_120x200", nx, ny, dx, dy, x, Z, h, 0.0, ph, {}, steps)
    sc.forcing = wave_forcing(sc.initial_depth(), x, 1.0, 8.0, 10.0, ph.Dmin)
    sc.meta = dict(slope=0.02, bar_amp=1.0, wavelength=250.0, Hrms=1.0, T=8.0, theta=10.0)
    return sc


def seeded_beach(nx: int, ny: int, dx: float, dy: float, seed: int, steps: int,
                 tide: bool = False, dt: float = 0.5, random_waves: bool = False,
                 name: str | None = None) -> Scenario:
    """Bar-trough beach drawn from ``seed``: slope U[0.01,0.04], bar amplitude
    U[0.5,1.5] m, alongshore wavelength U[100,400] m (rounded to divide the domain).
    With ``random_waves`` the forcing is drawn too: Hrms U[0.5,3] m, T U[6,14] s,
    angle U[-30,30] deg; otherwise Hrms 1 m, T 8 s, 10 deg. With ``tide`` the offshore
    boundary is ConstantZs and carries a 1 m, 12.42 h tide through f0."""
    rng = np.random.default_rng(seed)
    slope = rng.uniform(0.01, 0.04)
    amp = rng.uniform(0.5, 1.5)
    lam = _periodic_wavelength(ny * dy, rng.uniform(100.0, 400.0))
    phase = rng.uniform(0.0, 2.0 * math.pi)
    # bar crest in water about 2.5x the largest bar height deep
    centre = 2.5 * (amp * 1.3) / slope
    x, Z, h = _bar_beach(nx, ny, dx, dy, slope, amp, 0.3 * amp, centre, 30.0, lam, phase=phase)
    if random_waves:
        Hrms, T, theta = rng.uniform(0.5, 3.0), rng.uniform(6.0, 14.0), rng.uniform(-30.0, 30.0)
    else:
        Hrms, T, theta = 1.0, 8.0, 10.0
    bc = OFFSHORE_BC["constant_zs"] if tide else OFFSHORE_BC["sponge"]
    ph = Physics(offshore_bc=bc, dt=dt)
    sc = Scenario(name or f"beach_{nx}x{ny}_s{seed}", nx, ny, dx, dy, x, Z, h, 0.0, ph, {}, steps)
    sc.forcing = wave_forcing(sc.initial_depth(), x, Hrms, T, theta, ph.Dmin)
    if tide:
        sc.f0_series = tide_f0_series(steps, dt, dx)  # uniform grid: dx_x[0] = dx
    sc.meta = dict(seed=seed, slope=slope, bar_amp=amp, wavelength=lam, Hrms=Hrms, T=T,
                   theta=theta, tide=tide, bar_centre=centre, bar_width=30.0, bar_mod=0.3, phase=phase,
                   x_shore=x[-1] - 0.08 * x[-1])
    return sc


def bar_field(sc: Scenario, amp: float, centre: float):
    """The bar of a seeded beach with amplitude ``amp`` (m) and crest ``centre`` metres
    seaward of the shoreline: amp (1 + mod sin(2 pi y / wavelength + phase))
    exp(-((s - centre)/width)^2), s the distance seaward of the shoreline (the bar of
    _bar_beach). Returns (h, dh/d amp, dh/d centre), each (nx, ny)."""
    m = sc.meta
    s = (m["x_shore"] - sc.x)[:, None]
    y = (np.arange(sc.ny) * sc.dy)[None, :]
    shape = (1.0 + m["bar_mod"] * np.sin(2.0 * math.pi * y / m["wavelength"] + m["phase"]))
    g = np.exp(-((s - centre) / m["bar_width"]) ** 2)
    h = amp * shape * g
    return (np.ascontiguousarray(h), np.ascontiguousarray(shape * g),
            np.ascontiguousarray(h * 2.0 * (s - centre) / m["bar_width"] ** 2))


def config2(steps: int = 2000, seed: int = 2, nx: int = 2048, ny: int = 4096) -> Scenario:
    """Plain Hydro ADI with tide: 2048 x 4096 cells, dx=dy=2 m, seeded bathymetry."""
    return seeded_beach(nx, ny, 2.0, 2.0, seed, steps, tide=True,
                        name=f"beach_tide_{nx}x{ny}_s{seed}")


def smooth_field(nx: int, ny: int, rng, amp: float, corr_cells: float = 12.0) -> np.ndarray:
    """Smooth random field, periodic along y: white noise filtered by a Gaussian of
    ``corr_cells`` cells (x by convolution with held edges, y spectrally), scaled so its
    largest magnitude is ``amp``."""
    w = rng.normal(size=(nx, ny))
    w = _smooth_x(w, corr_cells)
    ky = np.fft.rfftfreq(ny)
    w = np.fft.irfft(np.fft.rfft(w, axis=1) * np.exp(-0.5 * (2 * np.pi * ky * corr_cells) ** 2)[None, :],
                     n=ny, axis=1)
    return np.ascontiguousarray(amp * w / np.max(np.abs(w)))


def config4(seed: int = 4, steps: int = 400, nx: int = 512, ny: int = 1024, dx: float = 2.0,
            noise: float = 0.02):
    """Bathymetry inversion set-up: a seeded beach (Hrms 1 m, T 8 s, 10 deg, dt 0.5 s),
    the true bar field h_true = h + dh with dh a smooth random field of 0.2 m amplitude,
    and the observation noise level (0.02 m). Returns (scenario, dh, noise_rng)."""
    sc = seeded_beach(nx, ny, dx, dx, seed, steps, name=f"inversion_{nx}x{ny}_s{seed}")
    rng = np.random.default_rng(seed + 1000)
    dh = smooth_field(nx, ny, rng, 0.2)
    return sc, dh, rng


def config5(count: int = 16, seed: int = 5, steps: int = 300, nx_range=(256, 1024), ny_range=(512, 2048),
            dx: float = 2.0, noise: float = 0.02):
    """Batched inversion set-up: ``count`` seeded beaches with nx ~ U[nx_range],
    ny ~ U[ny_range] (Hrms U[0.5, 1.5] m, T U[6, 12] s, angle U[-20, 20] deg). The
    scenario's own bar is the first guess; the truth is that bar perturbed by a seed, its
    amplitude scaled by U[0.75, 1.25] and its crest moved by U[-12, 12] m (sc.meta
    true_amp, true_centre). Observations carry N(0, noise) noise. Returns (scenarios, dhs,
    rng) with dhs the true bar minus the first guess."""
    rng = np.random.default_rng(seed)
    scs, dhs = [], []
    for k in range(count):
        nx = int(rng.integers(nx_range[0], nx_range[1] + 1))
        ny = int(rng.integers(ny_range[0], ny_range[1] + 1))
        sc = seeded_beach(nx, ny, dx, dx, seed * 1000 + k, steps, name=f"inv{k}_{nx}x{ny}")
        Hrms, T, theta = rng.uniform(0.5, 1.5), rng.uniform(6.0, 12.0), rng.uniform(-20.0, 20.0)
        sc.forcing = wave_forcing(sc.initial_depth(), sc.x, Hrms, T, theta, sc.physics.Dmin)
        true_amp = sc.meta["bar_amp"] * rng.uniform(0.75, 1.25)
        true_centre = sc.meta["bar_centre"] + rng.uniform(-12.0, 12.0)
        sc.meta.update(Hrms=Hrms, T=T, theta=theta, true_amp=true_amp, true_centre=true_centre, noise=noise)
        scs.append(sc)
        dhs.append(np.ascontiguousarray(bar_field(sc, true_amp, true_centre)[0] - sc.h))
    return scs, dhs, rng


def batch_scenarios(count: int = 64, seed: int = 3, steps: int = 500,
                    nx_range=(128, 1024), ny_range=(256, 2048), dx: float = 2.0):
    """Uneven scenarios for the batched form: nx ~ U[nx_range], ny ~ U[ny_range]."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        nx = int(rng.integers(nx_range[0], nx_range[1] + 1))
        ny = int(rng.integers(ny_range[0], ny_range[1] + 1))
        out.append(seeded_beach(nx, ny, dx, dx, seed * 1000 + k, steps, random_waves=True,
                                name=f"batch{k}_{nx}x{ny}"))
    return out


def lake_at_rest(nx=24, ny=16, dx=5.0, dy=5.0, slope=0.02) -> Scenario:
    """Still water on a bar beach with zero forcing: an exact fixed point of the step."""
    x, Z, h = _bar_beach(nx, ny, dx, dy, slope, 0.5, 0.1, 40.0, 15.0, ny * dy / 2)
    ph = Physics(dt=0.5)
    sc = Scenario("lake", nx, ny, dx, dy, x, Z, h, 0.0, ph, {}, 10)
    D = sc.initial_depth()
    z = np.zeros_like(D)
    sc.forcing = {"Sxx": z.copy(), "Sxy": z.copy(), "Syy": z.copy(),
                  "Urms": np.full_like(D, 0.1), "Dw": z.copy(), "D_r": z.copy(), "H": z.copy()}
    return sc
