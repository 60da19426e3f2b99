"""The reference's run API over the HIP step: FixedDomain, FlowConditions, the forcings, the
schemes, ResolvedFields, Tracers, Heun and Simulation with advance/run.

Names, argument meanings and error classes follow the reference (an ``ArgumentError`` there is a
``ValueError`` here, an ``ErrorException`` a ``RuntimeError``).  Everything that moves a field runs
in libhdns.so on the GPU; this module keeps the host-side bookkeeping the reference keeps on the
host: the clock, the cache epochs, the CFL arithmetic and the blow-up guards.
"""
from __future__ import annotations

import ctypes
import math
import warnings
from ctypes import c_double, c_int32, c_int64, c_uint8, c_void_p
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .stochastic import OUProcess, StochasticOU


# --- frame ---------------------------------------------------------------------------------

class TwoThirds:
    """The sharp 2/3 rule (src/SimulationDomain.jl:61-74): keep the closed ball |k| <= (2/3) k_Ny."""
    code = 0
    retained_fraction = 2.0 / 3.0     # src/SimulationDomain.jl:181
    resolved_fraction = 2.0 / 3.0     # src/SimulationDomain.jl:192

    def __repr__(self) -> str:
        return "TwoThirds()"


class HouLi:
    """The smooth Hou-Li filter (src/SimulationDomain.jl:76-138): every mode of the nonlinear terms
    is multiplied by exp(-strength (|k|/k_Ny)^order).  ``strength=None`` is the element type's
    -log(eps): 36.04 for FP64, 15.94 for FP32.  ``materialized`` is accepted for the reference's
    signature; the library always evaluates the filter once per |m|^2 into a table, which is the
    materialized form (the two advance a run bit-identically in the reference)."""
    code = 1
    retained_fraction = 1.0           # src/SimulationDomain.jl:182
    resolved_fraction = 4.0 / 5.0     # src/SimulationDomain.jl:193

    def __init__(self, order: int = 36, strength: float | None = None, materialized: bool = False):
        if int(order) != order or int(order) < 1:
            raise ValueError(f"the HouLi filter's `order` must be a positive integer, got {order}")
        if strength is not None and not math.isfinite(float(strength)):
            raise ValueError(f"the HouLi filter's `strength` must be finite, got {strength}")
        self.order = int(order)
        self.strength = None if strength is None else float(strength)
        self.materialized = bool(materialized)

    def strength_for(self, element_type) -> float:
        """The strength a run of `element_type` uses (src/SimulationDomain.jl:367)."""
        return -math.log(float(np.finfo(element_type).eps)) if self.strength is None else self.strength

    def __repr__(self) -> str:
        return f"HouLi(order={self.order}, strength={self.strength})"


class PhaseShift:
    """Phase-shift dealiasing (src/SimulationDomain.jl:140-176): the nonlinear terms are the average
    of the products on the grid and on the grid displaced by half a cell along every axis, truncated
    to the closed ball |k| <= (sqrt 2 / 3) N (2 pi/L) = (2 sqrt 2 / 3) k_Ny."""
    code = 2
    retained_fraction = 2.0 * math.sqrt(2.0) / 3.0   # src/SimulationDomain.jl:184
    resolved_fraction = 2.0 * math.sqrt(2.0) / 3.0   # src/SimulationDomain.jl:194

    def __repr__(self) -> str:
        return "PhaseShift()"


def _ball9(n: int, dealiasing) -> float:
    """The bound of 9 |m|^2 that a strategy's closed ball keeps (inf under HouLi)."""
    if isinstance(dealiasing, HouLi):
        return math.inf
    return 2 * n * n if isinstance(dealiasing, PhaseShift) else n * n


# the transform lengths with a radix plan in the library: 2^k and 3 * 2^k
SUPPORTED_GRIDS = (16, 32, 48, 64, 96, 128, 192, 256, 512, 1024)
# the grids the HouLi and PhaseShift pipelines are built for
WIDE_GRIDS = (16, 32, 48, 64, 128, 256)


class FixedDomain:
    """The periodic box a run lives on (src/SimulationDomain.jl:199-328).

    Every argument is a keyword, as in the reference; ``element_type`` is ``np.float32``
    (FP32-bulk: FP32 transforms and pointwise arithmetic, FP64 exponentials, twiddles and
    reductions) or ``np.float64`` (plain FP64).  ``device`` is the HIP ordinal.
    """

    def __init__(self, *, grid_points_per_axis: int, side_length: float, dealiasing, element_type,
                 device: int = 0):
        n = int(grid_points_per_axis)
        if n not in SUPPORTED_GRIDS:
            raise NotImplementedError(f"grid_points_per_axis must be a power of two in 16..1024, or 48, 96 or 192; "
                                      f"got {n}")
        if not (side_length > 0 and math.isfinite(side_length)):
            raise ValueError("side_length must be positive and finite")
        if not isinstance(dealiasing, (TwoThirds, HouLi, PhaseShift)):
            raise ValueError(f"`dealiasing` must be TwoThirds(), HouLi(...) or PhaseShift(), got {dealiasing!r}")
        if not isinstance(dealiasing, TwoThirds) and n not in WIDE_GRIDS:
            raise NotImplementedError(f"HouLi and PhaseShift dealiasing are built for grid_points_per_axis in "
                                      f"{WIDE_GRIDS}; got {n}")
        if element_type not in (np.float32, np.float64):
            raise ValueError("element_type must be np.float32 or np.float64")
        self.grid_points_per_axis = n
        self.side_length = float(side_length)
        self.dealiasing = dealiasing
        self.element_type = element_type
        self.device = int(device)

    @property
    def precision(self) -> int:
        return L.FP32_BULK if self.element_type is np.float32 else L.FP64

    @property
    def grid_spacing(self) -> float:
        return self.side_length / self.grid_points_per_axis

    def _cutoff(self, fraction: float) -> float:
        # src/SimulationDomain.jl:306-311: T(fraction) * T(N/2) * T(2 pi/L), in the element type
        t = self.element_type
        return float(t(fraction) * t(self.grid_points_per_axis / 2) * t(2 * math.pi / self.side_length))

    def max_retained_wavenumber(self) -> float:
        """f_ret (N/2)(2 pi/L): the band the strategy keeps energetic (src/SimulationDomain.jl:391-418)."""
        return self._cutoff(self.dealiasing.retained_fraction)

    def max_resolved_wavenumber(self) -> float:
        """f_res (N/2)(2 pi/L): the band whose dynamics is Navier-Stokes, the resolution criterion's
        kappa_max (src/SimulationDomain.jl:420-457); 4/5 of Nyquist under HouLi."""
        return self._cutoff(self.dealiasing.resolved_fraction)

    def __repr__(self) -> str:
        n = self.grid_points_per_axis
        return f"FixedDomain({n}x{n}x{n}, L = {self.side_length}, {self.dealiasing}, {self.element_type.__name__})"


def truncate_to_band(f: np.ndarray, domain: FixedDomain) -> np.ndarray:
    """Apply the dealias kernel to a spectral field, in place (src/SimulationDomain.jl:520-537): zero
    every coefficient outside the closed ball (TwoThirds, PhaseShift), or multiply each mode by the
    HouLi filter exp(-strength (|k|/k_Ny)^order)."""
    n = domain.grid_points_per_axis
    i = np.arange(n)
    w = np.where(i <= n // 2, i, i - n)
    m2 = np.arange(n // 2 + 1)[None, None, :] ** 2 + w[None, :, None] ** 2 + w[:, None, None] ** 2
    dl = domain.dealiasing
    if isinstance(dl, HouLi):
        s = dl.strength_for(domain.element_type)
        f *= np.exp(-s * (np.sqrt(m2.astype(np.float64)) / (n / 2)) ** dl.order)
        return f
    f[..., 9 * m2 > _ball9(n, dl)] = 0
    return f


# --- physics -------------------------------------------------------------------------------

class NoForcing:
    """The unforced case (src/Forcing.jl:456-464)."""

    def __repr__(self) -> str:
        return "NoForcing()"


class ConstantPower:
    """Hold the injection rate on shells lowest..highest (src/Forcing.jl:714-719,1069-1117).

    A = eps / (2 k_band) is re-measured at every stage's own field on the device, with the
    relative empty-band clamp (src/Forcing.jl:110-111).
    """

    def __init__(self, domain: FixedDomain, *, injection_rate: float, lowest_shell: int = 1, highest_shell=2):
        # src/Forcing.jl:388-409
        if lowest_shell < 1:
            raise ValueError(f"`lowest_shell` must be at least 1, got {lowest_shell}: shell 0 is the mean "
                             "(kappa = 0) mode, which a band forcing cannot legitimately drive")
        if not (highest_shell == math.inf or (math.isfinite(highest_shell) and float(highest_shell).is_integer())):
            raise ValueError(f"`highest_shell` must be an integer shell or `Inf`, got {highest_shell}")
        if highest_shell < lowest_shell:
            raise ValueError(f"`highest_shell` must be at least `lowest_shell`, got {highest_shell} < {lowest_shell}")
        self.injection_rate = float(domain.element_type(injection_rate))
        self.lowest_shell = int(lowest_shell)
        self.highest_shell = math.inf if highest_shell == math.inf else int(highest_shell)

    def __repr__(self) -> str:
        return f"ConstantPower(injection_rate {self.injection_rate}, shells {self.lowest_shell}-{self.highest_shell})"


class Linear:
    """Linear (negative-damping) forcing f = A u on shells lowest..highest (src/Forcing.jl:600-642):
    the rate is folded into the integrating factor, so band modes grow at A - nu |k|^2, with no
    per-stage reduction.  highest_shell = Inf forces every mode but the mean (Lundgren's forcing).
    The rate must be given (the reference's deferred form, which takes it from a predicted
    stationary state, is not built)."""

    def __init__(self, domain: FixedDomain, *, rate: float, lowest_shell: int = 1, highest_shell=2):
        band = ConstantPower(domain, injection_rate=0.0, lowest_shell=lowest_shell, highest_shell=highest_shell)
        rate = float(domain.element_type(rate))
        if not math.isfinite(rate):
            raise ValueError(f"the Linear `rate` must be finite, got {rate}")
        self.rate = rate
        self.lowest_shell = band.lowest_shell
        self.highest_shell = band.highest_shell

    def __repr__(self) -> str:
        return f"Linear(rate {self.rate}, shells {self.lowest_shell}-{self.highest_shell})"


# (1/highest_shell, k~): the stationary kinetic energy in box units k / (eps L)^(2/3) that the
# reference's placeholder model assigns to a band starting at shell 1, piecewise linear in
# 1/highest_shell (src/Forcing.jl:887, :945-953)
_PLACEHOLDER_KINETIC_ENERGY_IN_BOX_UNITS = ((0.0, 0.47), (0.25, 0.56), (0.5, 0.72))


def _placeholder_kinetic_energy_in_box_units(highest_shell) -> float:
    x = 0.0 if highest_shell == math.inf else 1.0 / highest_shell
    table = _PLACEHOLDER_KINETIC_ENERGY_IN_BOX_UNITS
    if x >= table[-1][0]:
        return table[-1][1]
    for (x0, y0), (x1, y1) in zip(table, table[1:]):
        if x0 <= x <= x1:
            return y0 + (y1 - y0) * (x - x0) / (x1 - x0)
    raise AssertionError("unreachable")


class VarianceController:
    """Hold the whole field's kinetic energy at `target` by feedback (src/Forcing.jl:722-871,
    :1041-1065): A = (eps - G (k - k_inf) / t_l) / (2 k_band) on shells lowest..highest, re-measured
    at every stage's own field on the device, so that dk/dt = -G (k - k_inf) / t_l.  gain = 0 is the
    bare marginal law.  turnover_time defaults to 2 k_inf / (3 eps_inf) with eps_inf =
    (k_inf / k~)^(3/2) / L from the placeholder model of the band (bands from shell 1 to 2 or more)."""

    def __init__(self, domain: FixedDomain, *, target: float, gain: float = 1.0, turnover_time=None,
                 lowest_shell: int = 1, highest_shell=2):
        band = ConstantPower(domain, injection_rate=0.0, lowest_shell=lowest_shell, highest_shell=highest_shell)
        cast = domain.element_type
        target, gain = float(cast(target)), float(cast(gain))
        if not (math.isfinite(target) and target > 0):
            raise ValueError(f"`target` must be finite and positive, got {target}: it is the variance the "
                             "controller holds, and the rate divides by the band's own variance")
        if not (math.isfinite(gain) and gain >= 0):
            raise ValueError(f"`gain` must be finite and non-negative, got {gain}: a negative gain makes the "
                             "restoring term push the variance AWAY from the target. The value `gain = 0` is "
                             "the bare marginal law, no restoring at all, and is allowed.")
        if turnover_time is None:
            if band.lowest_shell != 1:
                raise ValueError("`turnover_time` has no default here: the model covers bands that start at "
                                 f"shell 1, but this VarianceController starts at shell {band.lowest_shell}. "
                                 "Pass `turnover_time`.")
            if not (band.highest_shell == math.inf or band.highest_shell >= 2):
                raise ValueError("`turnover_time` has no default here: the model has no value for a band ending "
                                 f"at shell {band.highest_shell}. Pass `turnover_time`.")
            kt = float(cast(_placeholder_kinetic_energy_in_box_units(band.highest_shell)))
            eps = float(cast((target / kt) ** 1.5 / float(cast(domain.side_length))))
            tl = float(cast(2 * target / (3 * eps)))
        else:
            tl = float(cast(turnover_time))
        if not (math.isfinite(tl) and tl > 0):
            raise ValueError(f"`turnover_time` must be finite and positive, got {tl}")
        self.target, self.gain, self.turnover_time = target, gain, tl
        self.turnover_time_derived = turnover_time is None
        self.lowest_shell = band.lowest_shell
        self.highest_shell = band.highest_shell

    def __repr__(self) -> str:
        return (f"VarianceController(target {self.target}, gain {self.gain}, turnover_time {self.turnover_time}, "
                f"shells {self.lowest_shell}-{self.highest_shell})")


class VarianceRescaling:
    """Sullivan-Mahalingam-Kerr rescaling (src/Forcing.jl:467-597): after each step the band's modes
    are rescaled by one factor that holds the band's kinetic energy (over = "band") or the whole
    field's (over = "total") at `target`; an empty band is left alone.  First order in time by
    construction (a post-step projection)."""

    def __init__(self, domain: FixedDomain, *, target: float, over: str, lowest_shell: int = 1, highest_shell=2):
        if over not in ("band", "total"):
            raise ValueError(f"VarianceRescaling `over` must be 'band' or 'total'; got {over!r}")
        band = ConstantPower(domain, injection_rate=0.0, lowest_shell=lowest_shell, highest_shell=highest_shell)
        target = float(domain.element_type(target))
        if not (math.isfinite(target) and target >= 0):
            raise ValueError(f"VarianceRescaling `target` must be finite and non-negative, got {target}: it is "
                             "the kinetic energy or scalar variance the forcing holds, and the rescale takes "
                             "its square root")
        self.target = target
        self.over = over
        self.lowest_shell = band.lowest_shell
        self.highest_shell = band.highest_shell

    def __repr__(self) -> str:
        return f"VarianceRescaling(target {self.target}, over {self.over}, shells {self.lowest_shell}-{self.highest_shell})"


class FlowConditions:
    """nu, Sc and the two forcings (src/Generator.jl:81-147), with the reference's guards."""

    def __init__(self, *, viscosity: float, schmidt: float, velocity_forcing=None, scalar_forcing=None):
        if not viscosity > 0:
            raise ValueError(f"`viscosity` must be positive, got {viscosity}: it sets the Reynolds number "
                             "through Re = 1/nu, and a non-positive value has no dissipation scale")
        if not schmidt > 0:
            raise ValueError(f"`schmidt` must be positive, got {schmidt}: the scalar diffusivity is nu/Sc, "
                             "so a non-positive Schmidt number makes it infinite or negative")
        self.viscosity = float(viscosity)
        self.schmidt = float(schmidt)
        self.velocity_forcing = velocity_forcing if velocity_forcing is not None else NoForcing()
        self.scalar_forcing = scalar_forcing if scalar_forcing is not None else NoForcing()
        if not isinstance(self.scalar_forcing,
                          (NoForcing, ConstantPower, Linear, VarianceController, VarianceRescaling, StochasticOU)):
            raise NotImplementedError("the scalar forcings built are NoForcing, ConstantPower, Linear, "
                                      "VarianceController, VarianceRescaling and StochasticOU")
        if velocity_forcing is not None and velocity_forcing is scalar_forcing and not isinstance(velocity_forcing, NoForcing):
            raise ValueError("one forcing instance is bound to both `velocity_forcing` and `scalar_forcing`; "
                             "build two (src/Generator.jl:111-115)")
        sf = self.scalar_forcing
        if isinstance(sf, VarianceController) and sf.turnover_time_derived and sf.gain > 0:
            # src/Generator.jl:120-127
            raise ValueError("this `VarianceController` is bound as `scalar_forcing` but its turnover_time was "
                             "derived from the kinetic-energy model, which is meaningless for a scalar variance; "
                             "pass the flow's turnover time 2k/(3 eps) explicitly")
        if not isinstance(self.velocity_forcing,
                          (NoForcing, ConstantPower, Linear, VarianceController, VarianceRescaling, StochasticOU)):
            raise NotImplementedError("the velocity forcings built are NoForcing, ConstantPower, Linear, "
                                      "VarianceController, VarianceRescaling and StochasticOU")

    @property
    def scalar_diffusivity(self) -> float:
        return self.viscosity / self.schmidt


def predicted_viscosity(domain: FixedDomain, dissipation: float, kmax_eta: float = 2.0) -> float:
    """nu = eps^{1/3} eta^{4/3} with eta = kmax_eta / kappa_max (src/SimulationRun.jl:647-655); the
    resolution criterion defaults to 2 as the reference's does (src/SimulationRun.jl:648).  A
    non-positive or non-finite kmax_eta or dissipation is refused by name (src/SimulationRun.jl:648-649)."""
    kmax_eta, dissipation = float(kmax_eta), float(dissipation)
    if not (math.isfinite(kmax_eta) and kmax_eta > 0):
        raise ValueError(f"`kmax_eta` must be finite and positive, got {kmax_eta}: it is the resolution criterion "
                         "kappa_max eta, from which the viscosity is solved")
    if not (math.isfinite(dissipation) and dissipation > 0):
        raise ValueError(f"`dissipation` must be finite and positive, got {dissipation}: it is the dissipation "
                         "rate eps the viscosity is solved for")
    eta = kmax_eta / domain.max_resolved_wavenumber()
    return dissipation ** (1.0 / 3.0) * eta ** (4.0 / 3.0)


# --- schemes -------------------------------------------------------------------------------

class IFLowStorageRK3:
    """Williamson 2N third-order IF scheme (src/Propagator.jl:220-274)."""
    code = L.RK3
    base_cfl = 0.5


class IFLowStorageRK4:
    """Carpenter-Kennedy 5-stage 2N fourth-order IF scheme (src/Propagator.jl:277-320)."""
    code = L.RK4
    base_cfl = 0.96


class IFSSPRK3:
    """Shu-Osher three-stage SSP IF scheme (src/Propagator.jl:547-691); its three-stage third-order
    stability function shares IFLowStorageRK3's ceiling (:911)."""
    code = L.SSPRK3
    base_cfl = 0.5


class IFHeun:
    """Two-stage second-order IF Heun (src/Propagator.jl:420-544); unstable on the imaginary axis
    without viscous damping, hence its lower ceiling (:906)."""
    code = L.HEUN
    base_cfl = 0.25


def recommended_cfl(scheme, domain: FixedDomain) -> float:
    """The advective CFL ceiling (src/Propagator.jl:906-919): the scheme's base ceiling times
    (2/3)/f_ret, so TwoThirds leaves it unscaled and HouLi, which keeps modes to Nyquist, takes 2/3 of it."""
    if isinstance(domain.dealiasing, TwoThirds):
        return scheme.base_cfl   # the factor is exactly 1
    return scheme.base_cfl * (2.0 / 3.0) / domain.dealiasing.retained_fraction


# --- derived scales of (k, eps, nu) (src/Observables.jl:611-760) ---------------------------------

def _pos(name: str, v: float, allow_zero: bool = False) -> float:
    v = float(v)
    if not (math.isfinite(v) and (v >= 0 if allow_zero else v > 0)):
        raise ValueError(f"`{name}` must be finite and {'non-negative' if allow_zero else 'positive'}, got {v}")
    return v


def reynolds_box(side_length: float, dissipation: float, viscosity: float) -> float:
    """Re_box = eps^(1/3) L^(4/3) / nu (src/Observables.jl:611-617)."""
    L, e, nu = _pos("side_length", side_length), _pos("dissipation", dissipation, True), _pos("viscosity", viscosity)
    return e ** (1 / 3) * L ** (4 / 3) / nu


def reynolds_taylor(kinetic_energy: float, dissipation: float, viscosity: float) -> float:
    """Re_lambda = sqrt(20/3) k / sqrt(eps nu) (src/Observables.jl:647-653)."""
    k, e, nu = _pos("kinetic_energy", kinetic_energy, True), _pos("dissipation", dissipation), _pos("viscosity", viscosity)
    return math.sqrt(20 / 3) * k / math.sqrt(e * nu)


def taylor_length(kinetic_energy: float, dissipation: float, viscosity: float) -> float:
    """lambda = sqrt(10 nu k / eps) (src/Observables.jl:682-688)."""
    k, e, nu = _pos("kinetic_energy", kinetic_energy, True), _pos("dissipation", dissipation), _pos("viscosity", viscosity)
    return math.sqrt(10 * nu * k / e)


def kolmogorov_length(dissipation: float, viscosity: float) -> float:
    """eta = (nu^3 / eps)^(1/4) (src/Observables.jl:704-709)."""
    e, nu = _pos("dissipation", dissipation), _pos("viscosity", viscosity)
    return math.sqrt(math.sqrt(nu ** 3 / e))


def kolmogorov_time(dissipation: float, viscosity: float) -> float:
    """tau_eta = sqrt(nu / eps) (src/Observables.jl:724-729)."""
    e, nu = _pos("dissipation", dissipation), _pos("viscosity", viscosity)
    return math.sqrt(nu / e)


def kolmogorov_velocity(dissipation: float, viscosity: float) -> float:
    """u_eta = (eps nu)^(1/4) (src/Observables.jl:745-750)."""
    e, nu = _pos("dissipation", dissipation), _pos("viscosity", viscosity)
    return math.sqrt(math.sqrt(e * nu))


def integral_length(spectrum) -> float:
    """L11 = (3 pi / 4) sum_{kappa > 0} E / kappa / sum_{kappa > 0} E of a measured spectrum
    (kappa, E) (src/Observables.jl:567-581); no energy off the kappa = 0 shell is a ValueError."""
    kappa, energy = (np.asarray(a, dtype=np.float64) for a in spectrum)
    live = kappa > 0
    total = float(np.sum(energy[live]))
    if not total > 0:
        raise ValueError("this velocity field carries no energy, so it has no integral scale")
    return 3 * math.pi / 4 * float(np.sum(energy[live] / kappa[live])) / total


def cfl_bound(speed: float, dx: float, cfl: float, diffusivity: float, side_length: float,
              max_dt: float = math.inf, step: int = 0) -> float:
    """The adaptive step, the smallest of three bounds (src/SimulationRun.jl:1051-1096):
    advective cfl dx / max|u| (cfl dx at rest), resolution 1 / (D kappa_min^2) with D the largest
    diffusivity and kappa_min = 2 pi / L, and the caller's max_dt.  The fourth, the start-up bound
    of a multistep tracer scheme, is ``startup_bound``.  A non-finite speed is the blow-up error."""
    if not math.isfinite(speed):
        raise RuntimeError(f"non-finite velocity at step {step}: the run appears to have blown up")
    advective = cfl * dx / speed if speed > 0 else cfl * dx
    kmin = 2 * math.pi / side_length
    resolution = 1 / (diffusivity * kmin ** 2)
    return min(advective, resolution, max_dt)


def startup_growth(order: int) -> float:
    """How much the step may grow per iteration during the start-up (src/SimulationRun.jl:1116)."""
    return 2.0 if order <= 5 else (1.5 if order <= 7 else 1.25)


def startup_bound(scheme, trail_times, dt: float, viscosity: float, dissipation: float) -> float:
    """The start-up bound of shortened_first_steps (src/SimulationRun.jl:1099-1129): inf for a
    scheme on one sample; with one sample in the trail tau (dt / tau)^(p/3), tau = sqrt(nu / eps)
    the Kolmogorov time of the current state; after that the growth factor times the last step.
    ``dt`` is the smallest of the other bounds; ``trail_times`` newest first."""
    if scheme is None or scheme.steps <= 1:
        return math.inf
    if len(trail_times) > 1:
        return startup_growth(scheme.order) * (trail_times[0] - trail_times[1])
    tau = math.sqrt(viscosity / dissipation) if dissipation > 0 else math.inf
    return tau * (dt / tau) ** (scheme.order / 3) if math.isfinite(tau) and tau > 0 else math.inf


class ABM:
    """The k-step Adams-Bashforth-Moulton predictor-corrector for the tracers, of order k + 1,
    k = 1..7 (src/ParticlePropagator.jl:58-117): an AB step over the trail's k newest samples, the
    velocity gathered at the predicted positions, and the Adams-Moulton corrector; the weights come
    from the samples' actual times.  Its trail keeps k + 1 samples.  Until it holds k, the steps are
    ABM(1), ABM(2), ..."""
    code = L.P_ABM

    def __init__(self, k: int):
        if not (isinstance(k, int) and 1 <= k <= 7):
            raise ValueError(f"`ABM(k)` takes a number of steps `k` from 1 to 7, got {k}: the four-point "
                             "Gauss-Legendre rule that integrates the weights is exact only through the "
                             "degree `k = 7` needs, so a longer scheme would be given silently wrong coefficients")
        self.steps = k
        self.corrects = True
        self.order = k + 1
        self.trail_capacity = k + 1

    def __repr__(self) -> str:
        return f"ABM({self.steps})"


class Heun(ABM):
    """ABM(1), the default tracer scheme (src/ParticlePropagator.jl:118-165)."""
    code = L.P_HEUN

    def __init__(self):
        super().__init__(1)

    def __repr__(self) -> str:
        return "Heun()"


class AB:
    """Variable-step Adams-Bashforth of order k for the tracers, k = 1..8, TurTLE's scheme at k = 4
    (src/ParticlePropagator.jl:168-234): one gather per step (the closing sample).  Its trail keeps
    k samples.  Until it holds k, the steps are ABM(1), ABM(2), ..."""
    code = L.P_AB

    def __init__(self, k: int):
        if not (isinstance(k, int) and 1 <= k <= 8):
            raise ValueError(f"`AB(k)` takes a number of steps `k` from 1 to 8, got {k}: the four-point "
                             "Gauss-Legendre rule that integrates the weights is exact only through degree 7, "
                             "so a higher order would be given silently wrong coefficients")
        self.steps = k
        self.corrects = False
        self.order = k
        self.trail_capacity = k

    def __repr__(self) -> str:
        return f"AB({self.steps})"


# --- state and tracers ---------------------------------------------------------------------

@dataclass
class ResolvedFields:
    """(u_hat, phi_hat) in the reference layout: complex [3, N, N, N/2+1] and [N, N, N/2+1] (src/State.jl:119-137)."""
    velocity_spectral: np.ndarray
    scalar_spectral: np.ndarray | None = None


def unwrapped_positions(run) -> np.ndarray:
    """unwrapped_positions (src/Particles.jl:317-323) of a Simulation's or an EnsembleBox's tracers:
    the trajectory with its winding kept, [3, M]; raises when the set was built without
    ``unwrapped=True``."""
    return run.unwrapped_positions()


class Tracers:
    """A tracer set (src/Particles.jl:135-176): positions [3, M] in [0, L).  With ``unwrapped=True``
    the run also tracks the unwrapped positions, which start equal to the positions and take every
    push's displacement without the periodic reduction (src/Particles.jl:131-138,171-175)."""

    def __init__(self, domain: FixedDomain, *, positions: np.ndarray | None = None,
                 number_of_particles: int | None = None, seed: int = 2028, unwrapped: bool = False):
        if (positions is None) == (number_of_particles is None):
            raise ValueError("give exactly one of `positions` (a [3, M] array) or `number_of_particles`")
        if positions is None:
            from .seeding import uniform_positions
            positions = uniform_positions(int(number_of_particles), domain.side_length, seed, domain.element_type)
        positions = np.ascontiguousarray(positions, dtype=np.float64)
        if positions.ndim != 2 or positions.shape[0] != 3:
            raise ValueError("positions must have shape [3, M]")
        self.positions = positions
        self.interpolation_scheme = "CatmullRom"
        self.unwrapped = bool(unwrapped)

    @property
    def count(self) -> int:
        return self.positions.shape[1]


def _check_reality(f: np.ndarray, name: str, dtype) -> None:
    """src/SimulationRun.jl:1236-1251: the kx = 0 plane must be Hermitian to sqrt(eps(T)) of scale."""
    n = f.shape[-2]
    neg = (-np.arange(n)) % n
    s = f[..., 0]
    res = float(np.max(np.abs(s - np.conj(s[..., neg, :][..., :, neg])))) if s.size else 0.0
    scale = float(np.max(np.abs(f))) if f.size else 0.0
    if not res <= math.sqrt(np.finfo(dtype).eps) * scale:   # NaN refused, as `residual <= ...` refuses it
        raise ValueError(f"Simulation: the given `state`'s {name} coefficients are not the spectrum of a real "
                         f"field -- the conjugate-symmetry residual is {res} against a field scale of {scale}. "
                         "Repair with `reality_projection!` or fix the conversion that produced them.")


# --- the run -------------------------------------------------------------------------------

def _state_arrays(state: ResolvedFields, domain: FixedDomain):
    """The state's spectra as contiguous complex128 arrays, shape- and reality-checked
    (src/SimulationRun.jl:786-787,1236-1248)."""
    n = domain.grid_points_per_axis
    u = np.ascontiguousarray(state.velocity_spectral, dtype=np.complex128)
    if u.shape != (3, n, n, n // 2 + 1):
        raise ValueError(f"velocity_spectral must have shape {(3, n, n, n // 2 + 1)}, got {u.shape}")
    _check_reality(u, "velocity", domain.element_type)
    p = None
    if state.scalar_spectral is not None:
        p = np.ascontiguousarray(state.scalar_spectral, dtype=np.complex128)
        if p.shape != (n, n, n // 2 + 1):
            raise ValueError(f"scalar_spectral must have shape {(n, n, n // 2 + 1)}, got {p.shape}")
        _check_reality(p, "scalar", domain.element_type)
    return u, p


def _config(domain: FixedDomain, conditions: FlowConditions, scheme, carries_scalar: bool, mean_gradient: float,
            tracers, tracer_capacity, keep_physical: bool, keep_vorticity: bool,
            particle_scheme=None) -> L.HdnsConfig:
    """The C-ABI configuration of one box."""
    cfg = L.HdnsConfig()
    cfg.device = domain.device
    cfg.grid_points = domain.grid_points_per_axis
    cfg.side_length = domain.side_length
    cfg.precision = domain.precision
    cfg.scheme = scheme.code
    dl = domain.dealiasing
    cfg.dealiasing = dl.code
    if isinstance(dl, HouLi):
        cfg.filter_order = dl.order
        cfg.filter_strength = math.nan if dl.strength is None else dl.strength
    cfg.has_scalar = int(carries_scalar)
    cfg.viscosity = conditions.viscosity
    cfg.schmidt = conditions.schmidt
    f = conditions.velocity_forcing
    if isinstance(f, (ConstantPower, Linear, VarianceController, VarianceRescaling, StochasticOU)):
        cfg.lowest_shell = f.lowest_shell
        cfg.highest_shell = -1 if f.highest_shell == math.inf else f.highest_shell
    if isinstance(f, StochasticOU):
        cfg.velocity_forcing = L.STOCHASTIC_OU
    elif isinstance(f, (ConstantPower, Linear)):
        cfg.velocity_forcing = L.CONSTANT_POWER if isinstance(f, ConstantPower) else L.LINEAR
        cfg.injection_rate = f.injection_rate if isinstance(f, ConstantPower) else f.rate
    elif isinstance(f, VarianceController):
        cfg.velocity_forcing = L.VARIANCE_CONTROLLER
        cfg.forcing_target, cfg.forcing_gain, cfg.forcing_time = f.target, f.gain, f.turnover_time
    elif isinstance(f, VarianceRescaling):
        cfg.velocity_forcing = L.VARIANCE_RESCALING
        cfg.forcing_target = f.target
        cfg.forcing_over = L.OVER_TOTAL if f.over == "total" else L.OVER_BAND
    sf = conditions.scalar_forcing
    if not carries_scalar and not isinstance(sf, NoForcing):
        raise ValueError(f"a scalar forcing ({sf!r}) needs a state that carries the scalar")
    if carries_scalar and not isinstance(sf, NoForcing):
        cfg.scalar_lowest_shell = sf.lowest_shell
        cfg.scalar_highest_shell = -1 if sf.highest_shell == math.inf else sf.highest_shell
        if isinstance(sf, StochasticOU):
            cfg.scalar_forcing = L.STOCHASTIC_OU
        elif isinstance(sf, ConstantPower):
            cfg.scalar_forcing, cfg.scalar_rate = L.CONSTANT_POWER, sf.injection_rate
        elif isinstance(sf, Linear):
            cfg.scalar_forcing, cfg.scalar_rate = L.LINEAR, sf.rate
        elif isinstance(sf, VarianceController):
            cfg.scalar_forcing = L.VARIANCE_CONTROLLER
            cfg.scalar_target, cfg.scalar_gain, cfg.scalar_time = sf.target, sf.gain, sf.turnover_time
        else:
            cfg.scalar_forcing, cfg.scalar_target = L.VARIANCE_RESCALING, sf.target
            cfg.scalar_over = L.OVER_TOTAL if sf.over == "total" else L.OVER_BAND
    cfg.mean_gradient = float(mean_gradient)
    ps = particle_scheme if particle_scheme is not None else Heun()
    cfg.particle_scheme = ps.code
    cfg.particle_steps = ps.steps
    cap = tracers.count if tracers is not None else 0
    cfg.tracer_capacity = int(tracer_capacity if tracer_capacity is not None else cap)
    cfg.unwrapped = int(tracers is not None and tracers.unwrapped)
    keep = keep_physical or tracers is not None
    cfg.store_physical = (1 | (2 if carries_scalar else 0) if keep else 0) | (4 if keep_vorticity else 0)
    return cfg


# the kernel families the autotuner schedules, in the order of hdns_schedule's entries
SCHEDULE_FAMILIES = ("y_passes", "z_forward", "z_inverse", "x_pass", "y_radix_plan")
SCHEDULE_VARIANTS = 4   # the autotune report's row per family (HDNS_SCHED_VARIANTS)



REVOLVE_OPS = ("advance", "snapshot", "restore", "first_turn", "adjoint")


def revolve_schedule(steps: int, snapshots: int) -> list:
    """The binomial checkpointing schedule hdns_adjoint_gradient_revolve runs: a list of
    (action, step, slot) with action in REVOLVE_OPS (host arithmetic, no device)."""
    count = c_int64()
    L.check(L.lib().hdns_revolve_schedule(int(steps), int(snapshots), None, 0, ctypes.byref(count)))
    buf = np.zeros(3 * count.value, dtype=np.int32)
    L.check(L.lib().hdns_revolve_schedule(int(steps), int(snapshots), buf.ctypes.data_as(ctypes.POINTER(c_int32)),
                                          count.value, ctypes.byref(count)))
    return [(REVOLVE_OPS[op], int(step), int(slot)) for op, step, slot in buf.reshape(-1, 3)]


class Simulation:
    """A run: the device context (generator, registers, cache), the tracers and the clock.

    ``Simulation(domain; state = ResolvedFields(u_hat, phi_hat))`` (src/SimulationRun.jl:750-931):
    the state is adopted and deposited, and the tracer trail gets its first sample at t = 0.
    ``mean_gradient`` is G of an imposed mean scalar gradient along z (an extension; the reference
    defers it, docs/ARCHITECTURE.md:769-771).
    """

    def __init__(self, domain: FixedDomain, *, conditions: FlowConditions, state: ResolvedFields,
                 field_integration_scheme=IFLowStorageRK4, tracers: Tracers | None = None,
                 particle_integration_scheme=None, mean_gradient: float = 0.0, cfl: float | None = None,
                 max_dt: float = math.inf, keep_vorticity: bool = False, keep_physical: bool = True,
                 tracer_capacity: int | None = None, shortened_first_steps: bool = False,
                 traversal_mode: str = "automatic", use_graph: bool = True):
        carries_scalar = state.scalar_spectral is not None
        u, p = _state_arrays(state, domain)
        self.mean_gradient = float(mean_gradient)
        if mean_gradient and not carries_scalar:
            raise ValueError("a mean scalar gradient needs a scalar in the state")
        self.domain = domain
        self.conditions = conditions
        self.field_integration_scheme = field_integration_scheme
        self.particle_integration_scheme = particle_integration_scheme or (Heun() if tracers is not None else None)
        if self.particle_integration_scheme is not None and not isinstance(self.particle_integration_scheme, (ABM, AB)):
            raise ValueError("the particle integration scheme must be Heun(), ABM(k) or AB(k), got "
                             f"{self.particle_integration_scheme!r}")
        self.shortened_first_steps = bool(shortened_first_steps)
        # src/SimulationRun.jl:242-244: the gathers at the tracers visit them in a tile order rebuilt
        # every iteration (automatic), rebuilt by the caller (manual), or in tracer order (none)
        traversal_mode = str(traversal_mode).lstrip(":")
        if traversal_mode not in ("automatic", "manual", "none"):
            raise ValueError(f"`traversal_mode` must be `:automatic`, `:manual` or `:none`; got {traversal_mode!r}")
        self.traversal_mode = traversal_mode
        # advance() replays the step as one captured HIP graph (bitwise the eager step; one capture
        # serves every dt), or launches it kernel by kernel
        self.use_graph = bool(use_graph)
        self.tracers = tracers
        self.cfl = recommended_cfl(field_integration_scheme, domain) if cfl is None else float(cfl)
        self.max_dt = float(max_dt)
        self.time = 0.0
        self.step = 0
        self.carries_scalar = carries_scalar
        cfg = _config(domain, conditions, field_integration_scheme, carries_scalar, mean_gradient, tracers,
                      tracer_capacity, keep_physical, keep_vorticity, self.particle_integration_scheme)
        h = c_void_p()
        L.check(L.lib().hdns_ctx_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._ctx = h
        self._cfg = cfg
        # StochasticOU forcing states, stationary at the start (src/Forcing.jl:1388-1392)
        self.ou = {}
        if isinstance(conditions.velocity_forcing, StochasticOU):
            self.ou[0] = OUProcess(conditions.velocity_forcing, 1)
        if carries_scalar and isinstance(conditions.scalar_forcing, StochasticOU):
            self.ou[1] = OUProcess(conditions.scalar_forcing, 0)
        self.fields_epoch = 0
        self.particles_epoch = 0
        self.trail_epoch = None
        self.trail_time = None
        try:
            self._call("hdns_state_upload", L.dptr(u), L.dptr(p) if p is not None else None)
            if tracers is not None:
                self._call("hdns_tracers_upload", tracers.count, L.dptr(tracers.positions))
                if self.traversal_mode != "none":   # src/SimulationRun.jl:921-922
                    self._call("hdns_rebuild_traversal")
                self._call("hdns_deposit_tracers")
                self.trail_epoch = (self.fields_epoch, self.particles_epoch)
                self.trail_time = 0.0
        except Exception:
            self.close()
            raise

    # plumbing
    def _call(self, name: str, *args) -> None:
        if self._ctx is None:
            raise RuntimeError("this Simulation has been closed")
        L.check(getattr(L.lib(), name)(self._ctx, *args), self._ctx)

    def close(self) -> None:
        if getattr(self, "_ctx", None) is not None:
            L.lib().hdns_ctx_destroy(self._ctx)
            self._ctx = None

    def __del__(self):  # pragma: no cover - interpreter teardown order
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # observables of the last deposit
    def max_speed(self) -> float:
        """max|u| of the cached deposit (src/Cache.jl:243)."""
        out = c_double()
        self._call("hdns_max_speed", ctypes.byref(out))
        return out.value

    def reductions(self) -> dict:
        """k, eps, <phi^2>, eps_phi and the forced-band energy (FP64, compensated on the device)."""
        out = np.zeros(5)
        self._call("hdns_reductions", L.dptr(out))
        keys = ("kinetic_energy", "dissipation", "scalar_variance", "scalar_dissipation", "band_energy")
        return dict(zip(keys, map(float, out)))

    def distance(self, other: "Simulation") -> dict:
        """RMS difference of this run's state from `other`'s (same grid and device; the element
        types may differ), over the grid points, with `other`'s RMS alongside and their ratios:
        the FP32-bulk error report against an FP64 run of the same flow."""
        if self._ctx is None or other._ctx is None:
            raise ValueError("distance compares two open standalone runs (not ensemble boxes)")
        out = np.zeros(4)
        L.check(L.lib().hdns_state_distance(self._ctx, other._ctx, L.dptr(out)), self._ctx)
        r = {"velocity_rms_difference": out[0], "velocity_rms": out[1],
             "velocity_relative": out[0] / out[1] if out[1] > 0 else math.nan}
        if self.carries_scalar:
            r.update({"scalar_rms_difference": out[2], "scalar_rms": out[3],
                      "scalar_relative": out[2] / out[3] if out[3] > 0 else math.nan})
        return r

    def derivative_skewness(self) -> float:
        """C * 3 <det S> / <S:S>^(3/2) of the current state, C = (4/7) sqrt(15/2): about 0 for a
        random-phase field, about -0.5 for developed turbulence (src/Observables.jl:406,465-477)."""
        out = np.zeros(1)
        self._call("hdns_derivative_skewness", L.dptr(out))
        return float(out[0])

    def kinetic_energy_and_dissipation_below_scale(self) -> dict:
        """(scales, kinetic_energy, dissipation) below every side l = n L / N of a top-hat box
        (src/Observables.jl:195-264): k(l) = k - (1/2) <|<u>_l|^2>, eps(l) = eps - nu <|<w>_l|^2>, with
        the box mean's transfer W_l = prod_a sinc(k_a l / 2).  O(N^4) work: an occasional diagnostic."""
        n = self.domain.grid_points_per_axis
        sc, k, e = np.zeros(n), np.zeros(n), np.zeros(n)
        self._call("hdns_below_scale", L.dptr(sc), L.dptr(k), L.dptr(e))
        return {"scales": sc, "kinetic_energy": k, "dissipation": e}

    def integral_length(self) -> float:
        """L11 of the current state's energy spectrum (src/Observables.jl:556-557)."""
        return integral_length(self.energy_spectrum())

    def energy_spectrum(self, which: int = L.FIELD_VELOCITY):
        """(shell centres, E(kappa)) of the state (src/Observables.jl:502-514)."""
        n = self.domain.grid_points_per_axis
        nb = int(math.floor(math.sqrt(3.0) * n / 2)) + 2
        out = np.zeros(nb)
        self._call("hdns_spectrum", which, L.dptr(out), nb)
        return np.arange(nb) * (2 * math.pi / self.domain.side_length), out

    def state(self) -> ResolvedFields:
        """The device state, copied back in the reference layout."""
        n = self.domain.grid_points_per_axis
        u = np.zeros((3, n, n, n // 2 + 1), dtype=np.complex128)
        p = np.zeros((n, n, n // 2 + 1), dtype=np.complex128) if self.carries_scalar else None
        self._call("hdns_state_download", L.dptr(u), L.dptr(p) if p is not None else None)
        return ResolvedFields(u, p)

    def physical(self, which: int = L.FIELD_VELOCITY) -> np.ndarray:
        """A deposited physical field: velocity [3, N, N, N], scalar [N, N, N] or vorticity [3, N, N, N]."""
        n = self.domain.grid_points_per_axis
        shape = (n, n, n) if which == L.FIELD_SCALAR else (3, n, n, n)
        out = np.zeros(shape)
        self._call("hdns_field_download", which, L.dptr(out))
        return out

    def tracer_positions(self) -> np.ndarray:
        out = np.zeros((3, self.tracers.count))
        self._call("hdns_tracers_download", L.dptr(out))
        return out

    def unwrapped_positions(self) -> np.ndarray:
        """unwrapped_positions (src/Particles.jl:317-323): the winding-preserving positions [3, M];
        raises on a tracer set built without ``unwrapped=True``."""
        if self.tracers is None or not self.tracers.unwrapped:
            raise ValueError("this tracer set carries no unwrapped positions: build it with `unwrapped=True` "
                             "(a `Tracers` keyword) to track winding alongside the wrapped positions")
        out = np.zeros((3, self.tracers.count))
        self._call("hdns_tracers_unwrapped_download", L.dptr(out))
        return out

    def set_unwrapped_positions(self, positions: np.ndarray) -> None:
        """Replace the unwrapped record (a checkpoint's stored one), [3, M]."""
        if self.tracers is None or not self.tracers.unwrapped:
            raise ValueError("this tracer set carries no unwrapped positions: build it with `unwrapped=True`")
        x = np.ascontiguousarray(positions, dtype=np.float64)
        if x.shape != (3, self.tracers.count):
            raise ValueError(f"unwrapped positions must have shape {(3, self.tracers.count)}, got {x.shape}")
        self._call("hdns_tracers_unwrapped_upload", self.tracers.count, L.dptr(x))

    def tracer_samples(self):
        """(velocity [3, M], scalar [M] or None) of the newest trail sample."""
        m = self.tracers.count
        v = np.zeros((3, m))
        s = np.zeros(m) if self.carries_scalar else None
        self._call("hdns_tracer_samples", L.dptr(v), L.dptr(s) if s is not None else None)
        return v, s

    def tracer_cells(self) -> np.ndarray:
        keys = np.zeros(self.tracers.count, dtype=np.int64)
        self._call("hdns_tracer_cells", keys.ctypes.data_as(ctypes.POINTER(c_int64)))
        return keys

    def rebuild_traversal(self) -> "Simulation":
        """rebuild_traversal!(sim) (src/SimulationRun.jl:1257-1266): rebuild the tracer traversal
        order from the current positions; a run without one refuses."""
        if self.tracers is None or self.traversal_mode == "none":
            why = "without tracers" if self.tracers is None else "with `traversal_mode = :none`"
            raise ValueError(f"this run carries no particle traversal order to rebuild: it was built {why}. Build it "
                             "with `traversal_mode = :manual` to rebuild by hand, or leave the default, which "
                             "rebuilds every iteration.")
        self._call("hdns_rebuild_traversal")
        return self

    def traversal_order(self) -> np.ndarray:
        order = np.zeros(self.tracers.count, dtype=np.int32)
        self._call("hdns_traversal_order", order.ctypes.data_as(ctypes.POINTER(c_int32)))
        return order

    def frame_tables(self):
        n = self.domain.grid_points_per_axis
        mx = np.zeros(n // 2 + 1, dtype=np.int32)
        my = np.zeros(n, dtype=np.int32)
        mz = np.zeros(n, dtype=np.int32)
        mask = np.zeros((n, n, n // 2 + 1), dtype=np.uint8)
        P = ctypes.POINTER
        self._call("hdns_frame_tables", mx.ctypes.data_as(P(c_int32)), my.ctypes.data_as(P(c_int32)),
                   mz.ctypes.data_as(P(c_int32)), mask.ctypes.data_as(P(c_uint8)))
        return mx, my, mz, mask

    def memory_bytes(self) -> int:
        b = c_int64()
        self._call("hdns_memory_bytes", ctypes.byref(b))
        return b.value

    def synchronize(self) -> None:
        self._call("hdns_synchronize")

    def stream(self) -> int:
        """The hipStream_t the context launches on (for event timing on that stream)."""
        s = c_void_p()
        self._call("hdns_stream", ctypes.byref(s))
        return s.value or 0

    def autotune(self, reps: int = 3) -> dict:
        """Time every launch schedule of every stage kernel family and keep the fastest (the state is
        left bitwise unchanged); returns {family: {"ms": [ms of each variant], "chosen": index}}."""
        ms = np.zeros(SCHEDULE_VARIANTS * len(SCHEDULE_FAMILIES))
        chosen = np.zeros(len(SCHEDULE_FAMILIES), dtype=np.int32)
        self._call("hdns_autotune", int(reps), L.dptr(ms), chosen.ctypes.data_as(ctypes.POINTER(c_int32)))
        out = {}
        for i, name in enumerate(SCHEDULE_FAMILIES):
            row = ms[SCHEDULE_VARIANTS * i: SCHEDULE_VARIANTS * (i + 1)]
            out[name] = {"ms": [round(float(t), 5) for t in row if t >= 0], "chosen": int(chosen[i])}
        return out

    def schedule(self, set_to=None):
        """Read (and optionally set) the launch schedule: one variant index per kernel family
        (y passes 0-1, z forward 0-2, z inverse 0-3, x pass 0-1, y radix plan 0-1)."""
        get = np.zeros(len(SCHEDULE_FAMILIES), dtype=np.int32)
        P = ctypes.POINTER(c_int32)
        s = None if set_to is None else np.ascontiguousarray(set_to, dtype=np.int32)
        if s is not None and s.shape != (len(SCHEDULE_FAMILIES),):
            raise ValueError(f"a schedule has {len(SCHEDULE_FAMILIES)} entries, one per kernel family")
        self._call("hdns_schedule", s.ctypes.data_as(P) if s is not None else None, get.ctypes.data_as(P))
        return tuple(int(v) for v in get)

    def tracer_schedule(self, set_to: int | None = None) -> int:
        """Read (and optionally set) the tracer gathers' launch: 1 capped (128-thread blocks, four
        per CU), 0 uncapped, -1 by the precision and the tracer count (FP64 capped from 2^16
        tracers on, FP32 uncapped); returns the launch in force."""
        get = c_int32()
        self._call("hdns_tracer_schedule", -2 if set_to is None else int(set_to), ctypes.byref(get))
        return get.value

    def profile(self, on: bool) -> None:
        """Time every kernel launch with HIP events on the context's stream."""
        self._call("hdns_profile_enable", int(on))

    def profile_read(self, reset: bool = True):
        """(ms, launches) per kernel kind: y forward, z forward + update, reality, z inverse,
        finalize, y inverse, x pass, other."""
        ms = np.zeros(8)
        cnt = np.zeros(8, dtype=np.int64)
        self._call("hdns_profile_read", L.dptr(ms), cnt.ctypes.data_as(ctypes.POINTER(c_int64)), int(reset))
        return ms, cnt

    # the hot path
    def set_fields(self, state: ResolvedFields) -> None:
        """Adopt a new field state and deposit it, keeping the context (the adoption door of
        src/SimulationRun.jl:786-787 without rebuilding the run).  Tracers keep their positions;
        the clock and the trail are the caller's to reset."""
        if (state.scalar_spectral is not None) != self.carries_scalar:
            raise ValueError("the new state must carry a scalar exactly when the run does")
        u, p = _state_arrays(state, self.domain)
        self._call("hdns_state_upload", L.dptr(u), L.dptr(p) if p is not None else None)
        self.fields_epoch += 1

    def step_fields(self, dt: float, graph: bool = False) -> None:
        """``step!`` then the field ``deposit!`` (src/Propagator.jl:304-419,808-842), asynchronously;
        ``graph`` replays the step as one captured HIP graph.  A StochasticOU forcing's process first
        takes its transition over dt and its band force goes to the device (prepare_forcing!)."""
        for field, proc in self.ou.items():
            proc.advance(dt)
            self._call("hdns_band_source_upload", int(field), L.dptr(proc.band_source(self.domain.grid_points_per_axis)))
        self._call("hdns_step_graph" if graph else "hdns_step", float(dt))
        self.fields_epoch += 1

    # the adjoint (the rrule of step!)
    def step_vjp(self, dt: float, lam_u: np.ndarray, lam_phi: np.ndarray | None = None):
        """The transpose of one step's derivative at the current state (which is left unchanged):
        the adjoints (lam_u, lam_phi) of the step's output, in the state layout, become those of
        its input.  Adjoints are gradients under the real inner product of the half spectrum with
        the Parseval weights (1 on kx = 0, 2 inside)."""
        n = self.domain.grid_points_per_axis
        lu = np.ascontiguousarray(lam_u, dtype=np.complex128)
        if lu.shape != (3, n, n, n // 2 + 1):
            raise ValueError("lam_u must have the velocity state's shape")
        lp = None
        if self.carries_scalar:
            lp = np.ascontiguousarray(lam_phi if lam_phi is not None else np.zeros((n, n, n // 2 + 1)),
                                      dtype=np.complex128)
        out_u = np.zeros_like(lu)
        out_p = np.zeros_like(lp) if lp is not None else None
        self._call("hdns_step_vjp", float(dt), L.dptr(lu), L.dptr(lp) if lp is not None else None, L.dptr(out_u),
                   L.dptr(out_p) if out_p is not None else None)
        return out_u, out_p

    def adjoint_gradient(self, dt: float, steps: int, checkpoint_every: int = 20, snapshots: int | None = None):
        """J = k + <phi^2> after `steps` steps from the current state, and its gradient with respect
        to the current u_hat (and phi_hat), by the hand-written adjoint with checkpoints every
        `checkpoint_every` steps, or, given `snapshots`, under binomial (revolve) checkpointing with
        that many state slots (the initial state's included); both give the same bits.  The run's state is
        left unchanged.  Returns (J, grad_u, grad_phi)."""
        n = self.domain.grid_points_per_axis
        J = c_double()
        gu = np.zeros((3, n, n, n // 2 + 1), dtype=np.complex128)
        gp = np.zeros((n, n, n // 2 + 1), dtype=np.complex128) if self.carries_scalar else None
        if snapshots is None:
            self._call("hdns_adjoint_gradient", float(dt), int(steps), int(checkpoint_every), ctypes.byref(J),
                       L.dptr(gu), L.dptr(gp) if gp is not None else None)
        else:
            self._call("hdns_adjoint_gradient_revolve", float(dt), int(steps), int(snapshots), ctypes.byref(J),
                       L.dptr(gu), L.dptr(gp) if gp is not None else None)
        return J.value, gu, gp

    def cfl_timestep(self) -> float:
        """The advective/resolution/max_dt bound and, under shortened_first_steps, the start-up
        bound of a multistep tracer scheme (src/SimulationRun.jl:1051-1129)."""
        c = self.conditions
        diffusivity = max(c.viscosity, c.scalar_diffusivity) if self.carries_scalar else c.viscosity
        dt = cfl_bound(self.max_speed(), self.domain.grid_spacing, self.cfl, diffusivity,
                       self.domain.side_length, self.max_dt, step=self.step)
        if self.shortened_first_steps and self.tracers is not None:
            eps = self.reductions()["dissipation"]
            dt = min(startup_bound(self.particle_integration_scheme, self.trail_times(), dt, c.viscosity, eps), dt)
        return dt

    def redeposit(self) -> None:
        """The field deposit! of the current state again (after a schedule change, say), and the
        tracer sample at the unmoved positions with it."""
        self._call("hdns_deposit")
        self.fields_epoch += 1
        if self.tracers is not None:
            self._call("hdns_deposit_tracers")
            self.trail_epoch = (self.fields_epoch, self.particles_epoch)

    def restore_trail(self, times, samples, t_next: float) -> None:
        """Put a stored trail back (read_simulation, ext/GBGHomogeneousTurbulenceHDF5Ext.jl:535-544):
        the earlier samples [n, 3, M] newest first with their times, then the newest measured again
        at the positions at t_next."""
        if self.tracers is None:
            raise ValueError("this run carries no tracers")
        samples = np.ascontiguousarray(samples, dtype=np.float64)
        times = np.ascontiguousarray(times, dtype=np.float64)
        if times.ndim != 1:
            raise ValueError("the trail's times are one number per stored sample")
        want = (times.shape[0], 3, self.tracers.count)
        # ext/GBGHomogeneousTurbulenceHDF5Ext.jl:268-283: a sample that does not fit the trail is refused
        if samples.shape != want:
            raise ValueError(f"the stored samples do not fit the trail they are read into: expected shape {want} "
                             f"(samples, components, tracers), got {samples.shape}")
        self._call("hdns_tracer_trail_restore", int(times.shape[0]), L.dptr(times), L.dptr(samples), float(t_next))
        self._call("hdns_deposit_tracers")
        self.trail_epoch = (self.fields_epoch, self.particles_epoch)
        self.trail_time = float(t_next)

    def trail_times(self) -> list:
        """The times of the velocity trail's samples at the tracers, newest first (src/Particles.jl:404-526)."""
        filled, cap = c_int32(), c_int32()
        times = np.zeros(8)
        self._call("hdns_tracer_trail", ctypes.byref(filled), ctypes.byref(cap), L.dptr(times))
        return [float(t) for t in times[: filled.value]]

    def trail_sample(self, offset: int = 0) -> np.ndarray:
        """The trail's velocity sample `offset` steps back from the newest (offset <= 0), [3, M]."""
        v = np.zeros((3, self.tracers.count))
        self._call("hdns_tracer_trail_sample", int(offset), L.dptr(v))
        return v

    def time_derivative(self, order: int = 1) -> np.ndarray:
        """time_derivative!(out, trail; order) (src/Particles.jl:597-631): d^m v/dt^m at the newest
        sample of the polynomial through every sample held; order 1 is the Lagrangian acceleration."""
        out = np.zeros((3, self.tracers.count))
        self._call("hdns_tracer_time_derivative", int(order), L.dptr(out))
        return out

    def advect(self, dt: float) -> None:
        """advect!(scheme, ...) (src/ParticlePropagator.jl:349-372) with its epoch check."""
        if self.tracers is None:
            raise ValueError("this run carries no tracers")
        if self.trail_epoch != (self.fields_epoch - 1, self.particles_epoch):
            raise RuntimeError("the fields must be exactly one step ahead of the tracer positions, and the "
                               "trail's newest sample taken at those positions")
        self._call("hdns_advect", float(dt))
        self.particles_epoch += 1

    def deposit_tracers(self, t: float) -> None:
        """The particle deposit! (src/ParticlePropagator.jl:378-386)."""
        if self.tracers is None:
            raise ValueError("this run carries no tracers")
        if not math.isfinite(t) or (self.trail_time is not None and t <= self.trail_time):
            raise ValueError(f"the sample time must be finite and later than the previous sample's, got {t}")
        self._call("hdns_deposit_tracers")
        self.trail_epoch = (self.fields_epoch, self.particles_epoch)
        self.trail_time = t

    def advect_and_deposit(self, dt: float, t: float) -> None:
        """advect(dt) then deposit_tracers(t) as one launch (hdns_advect_deposit): the iteration's
        push and closing sample, bit for bit the two calls, with both calls' checks made before
        anything moves."""
        if self.tracers is None:
            raise ValueError("this run carries no tracers")
        if self.trail_epoch != (self.fields_epoch - 1, self.particles_epoch):
            raise RuntimeError("the fields must be exactly one step ahead of the tracer positions, and the "
                               "trail's newest sample taken at those positions")
        if not math.isfinite(t) or (self.trail_time is not None and t <= self.trail_time):
            raise ValueError(f"the sample time must be finite and later than the previous sample's, got {t}")
        self._call("hdns_advect_deposit", float(dt))
        self.particles_epoch += 1
        self.trail_epoch = (self.fields_epoch, self.particles_epoch)
        self.trail_time = t

    def _check_trail_between_iterations(self) -> None:
        """The checked read at an iteration's start (src/Cache.jl:342-349)."""
        fe, pe = self.trail_epoch if self.trail_epoch is not None else (None, None)
        if pe != self.particles_epoch:
            raise RuntimeError("the tracers were pushed and their new velocity sample was never taken: deposit "
                               "the tracers before the next iteration")
        if fe != self.fields_epoch:
            ahead = self.fields_epoch - fe
            raise RuntimeError(f"the fields have taken {ahead} step(s) that the tracers have not, so the two belong "
                               "to different times. A run that has let them part cannot be repaired; build a new "
                               "one from its state and tracers")

    def advance(self, dt: float | None = None, check: bool = True) -> "Simulation":
        """One iteration (src/SimulationRun.jl:1155-1229): dt, step + deposit, traversal rebuild,
        tracer push and sample, then the blow-up guards and the clock."""
        ps = self.particle_integration_scheme
        if dt is not None and self.shortened_first_steps and ps is not None and ps.steps > 1:
            # src/SimulationRun.jl:1158-1161
            raise ValueError("this run was built with `shortened_first_steps = true`, which chooses the timestep "
                             "itself while the particle scheme starts, and a fixed `dt` was given: drop `dt`, or "
                             "build the run without `shortened_first_steps`")
        dt_ = self.cfl_timestep() if dt is None else float(dt)
        if not (math.isfinite(dt_) and dt_ > 0):
            raise ValueError(f"the timestep must be positive and finite, got {dt_}")
        # src/SimulationRun.jl:1177-1182 (trail_between_iterations, src/Cache.jl:342-349): a run whose
        # fields were stepped alone, or whose last push was never closed by a deposit, is refused
        # before anything moves
        if self.tracers is not None:
            self._check_trail_between_iterations()
        t_new = self.time + dt_
        if not t_new > self.time:
            raise ValueError(f"the timestep {dt_} is too small to advance the clock, which reads {self.time}")
        self.step_fields(dt_, graph=self.use_graph)
        if self.tracers is not None:
            if self.traversal_mode == "automatic":   # from the start-of-step positions (src/SimulationRun.jl:1197-1200)
                self._call("hdns_rebuild_traversal")
            self.advect_and_deposit(dt_, t_new)   # advect! then deposit! (src/SimulationRun.jl:1203-1207)
        if check:
            red = None
            if not math.isfinite(self.max_speed()):
                raise L.BlowUpError(f"non-finite velocity at step {self.step + 1}: the run appears to have blown up")
            if self.carries_scalar:
                red = self.reductions()
                if not math.isfinite(red["scalar_variance"]):
                    raise L.BlowUpError(f"non-finite scalar at step {self.step + 1}: the run appears to have blown up")
        self.time = t_new
        self.step += 1
        return self

    def run(self, callback=None, *, steps: int | None = None, dt: float | None = None) -> "Simulation":
        """run!(callback, sim; dt) (src/SimulationRun.jl:1395-1405): hand the state to `callback`,
        advance while it returns True, stop when it returns False.  `steps` (an extension) caps the
        number of iterations, and with no callback it is the run's length."""
        if callback is None and steps is None:
            raise ValueError("give a stopping criterion, or a number of steps")
        done = 0
        while steps is None or done < int(steps):
            if callback is not None:
                keep = callback(self)
                if not isinstance(keep, bool):
                    raise ValueError("the stopping criterion must return a Bool -- true to continue the run, "
                                     f"false to stop it; it returned a {type(keep).__name__}")
                if not keep:
                    break
            self.advance(dt)
            done += 1
        return self

    def __repr__(self) -> str:
        n = self.domain.grid_points_per_axis
        return f"Simulation({n}x{n}x{n} grid, time {self.time}, step {self.step})"


# --- ensembles --------------------------------------------------------------------------------

@dataclass
class EnsembleMember:
    """One box of an ensemble: its frame, conditions, initial state and optional tracers."""
    domain: FixedDomain
    conditions: FlowConditions
    state: ResolvedFields
    tracers: Tracers | None = None


class EnsembleBox(Simulation):
    """One box of an Ensemble, with the per-box reads of a Simulation (state, reductions,
    max_speed, spectra, physical fields, tracers).  The ensemble advances it; a box has no clock
    of its own."""

    def __init__(self, ensemble: "Ensemble", index: int, member: EnsembleMember, scheme):
        # no device context of its own: every call goes to the batch that holds this box
        self._ens = ensemble
        self.index = index
        self.domain = member.domain
        self.conditions = member.conditions
        self.field_integration_scheme = scheme
        self.tracers = member.tracers
        self.carries_scalar = member.state.scalar_spectral is not None
        self.time = 0.0
        self.step = 0
        self._ctx = None

    def _member(self) -> c_void_p:
        h = c_void_p()
        L.check(L.lib().hdns_ensemble_member(self._ens._h, self.index, ctypes.byref(h)), ensemble=self._ens._h)
        return h

    def _call(self, name: str, *args) -> None:
        if self._ens._h is None:
            raise RuntimeError("this Ensemble has been closed")
        h = self._member()
        L.check(getattr(L.lib(), name)(h, *args), h)

    def close(self) -> None:  # the ensemble owns the device memory
        pass

    def _advanced_by_the_ensemble(self, *args, **kwargs):
        raise RuntimeError("an ensemble box is advanced by its Ensemble (Ensemble.step / advance_tracers)")

    step_fields = advance = run = advect = deposit_tracers = autotune = schedule = _advanced_by_the_ensemble
    step_vjp = adjoint_gradient = _advanced_by_the_ensemble

    def __repr__(self) -> str:
        n = self.domain.grid_points_per_axis
        return f"EnsembleBox({self.index}: {n}x{n}x{n}, nu {self.conditions.viscosity:g})"


class Ensemble:
    """Boxes of uneven sizes and viscosities advanced together (the library's hdns_ensemble).

    Boxes of one size form a batch that every kernel advances in one launch; all batches share a
    stream, and ``step(dt, graph=True)`` advances every box with one captured HIP graph.  The
    boxes may differ in grid size, viscosity and initial state; scheme, precision, forcing, the
    scalar and the tracer count are common.
    """

    def __init__(self, members: list[EnsembleMember], *, field_integration_scheme=IFLowStorageRK3,
                 mean_gradient: float = 0.0, keep_physical: bool = False, particle_integration_scheme=None):
        if not members:
            raise ValueError("an ensemble needs at least one box")
        self.field_integration_scheme = field_integration_scheme
        self.particle_integration_scheme = particle_integration_scheme or Heun()
        if not isinstance(self.particle_integration_scheme, (ABM, AB)):
            raise ValueError("the particle integration scheme must be Heun(), ABM(k) or AB(k), got "
                             f"{self.particle_integration_scheme!r}")
        cfgs = (L.HdnsConfig * len(members))()
        for b, mb in enumerate(members):
            if isinstance(mb.conditions.velocity_forcing, StochasticOU) or isinstance(mb.conditions.scalar_forcing,
                                                                                     StochasticOU):
                raise NotImplementedError("an ensemble box cannot carry a StochasticOU forcing (its process is "
                                          "kept per standalone run)")
            carries = mb.state.scalar_spectral is not None
            cfgs[b] = _config(mb.domain, mb.conditions, field_integration_scheme, carries, mean_gradient, mb.tracers,
                              None, keep_physical, False, self.particle_integration_scheme)
        h = c_void_p()
        L.check(L.lib().hdns_ensemble_create(cfgs, len(members), ctypes.byref(h)), ensemble=None)
        self._h = h
        self.time = 0.0
        self.step_count = 0
        self.boxes = [EnsembleBox(self, b, mb, field_integration_scheme) for b, mb in enumerate(members)]
        try:
            for box, mb in zip(self.boxes, members):
                try:
                    u, p = _state_arrays(mb.state, mb.domain)
                except ValueError as e:
                    raise ValueError(f"box {box.index}: {e}") from None
                box._call("hdns_state_upload", L.dptr(u), L.dptr(p) if p is not None else None)
                if mb.tracers is not None:
                    box._call("hdns_tracers_upload", mb.tracers.count, L.dptr(mb.tracers.positions))
            if members[0].tracers is not None:   # the trail's first sample, at t = 0
                for box in self.boxes:
                    box._call("hdns_rebuild_traversal")
                    box._call("hdns_deposit_tracers")
        except Exception:
            self.close()
            raise

    def _ecall(self, name: str, *args) -> None:
        if self._h is None:
            raise RuntimeError("this Ensemble has been closed")
        L.check(getattr(L.lib(), name)(self._h, *args), ensemble=self._h)

    def __len__(self) -> int:
        return len(self.boxes)

    def box(self, b: int) -> EnsembleBox:
        return self.boxes[b]

    def step(self, dt: float, graph: bool = True) -> "Ensemble":
        """One IF-RK step (+ the field deposit) of every box; ``graph`` replays one captured graph."""
        if not (math.isfinite(dt) and dt > 0):
            raise ValueError(f"the timestep must be positive and finite, got {dt}")
        self._ecall("hdns_ensemble_step_graph" if graph else "hdns_ensemble_step", float(dt))
        self.time += dt
        self.step_count += 1
        return self

    def advance_tracers(self, dt: float) -> None:
        """Traversal rebuild, push (the ensemble's particle scheme over each batch's trail) and sample
        of every box's tracers (after a field step)."""
        self._ecall("hdns_ensemble_advance_tracers", float(dt))

    def synchronize(self) -> None:
        self._ecall("hdns_ensemble_synchronize")

    def _batches(self) -> dict:
        """{grid size: the first box of that size}: each size is one batch (one device context)."""
        heads = {}
        for box in self.boxes:
            heads.setdefault(box.domain.grid_points_per_axis, box)
        return heads

    def autotune(self, reps: int = 3) -> dict:
        """Time every launch schedule of every stage kernel family on each batch of equal-size boxes
        (every box of the batch in one launch, as the step runs them) and keep the fastest; the
        states are left bitwise unchanged and the next graph step captures the new schedules.
        Returns {grid size: {family: {"ms": [ms of each variant], "chosen": index}}}."""
        if self._h is None:
            raise RuntimeError("this Ensemble has been closed")
        return {n: Simulation.autotune(head, reps) for n, head in sorted(self._batches().items())}

    def schedules(self, set_to: dict | None = None) -> dict:
        """Read (and optionally set, {grid size: one variant per family}) each batch's launch
        schedule, as Simulation.schedule does for one run."""
        if self._h is None:
            raise RuntimeError("this Ensemble has been closed")
        heads = self._batches()
        for n in set_to or {}:
            if n not in heads:
                raise ValueError(f"the ensemble has no box of size {n}: its sizes are {sorted(heads)}")
        return {n: Simulation.schedule(head, None if set_to is None else set_to.get(n))
                for n, head in sorted(heads.items())}

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            L.lib().hdns_ensemble_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):  # pragma: no cover - interpreter teardown order
        try:
            self.close()
        except Exception:
            pass

    def __repr__(self) -> str:
        sizes = sorted({b.domain.grid_points_per_axis for b in self.boxes})
        return f"Ensemble({len(self.boxes)} boxes of sizes {sizes}, time {self.time})"



# --- spin-up to stationarity ---------------------------------------------------------------

class StationarityMonitor:
    """The flatness test of the warm-up detector (src/SimulationRun.jl:1416-1485): signals stored
    at least window/40 turnovers apart over a trailing window; a series is flat when the means of
    the window's two halves differ by at most tolerance * |mean|."""

    def __init__(self, *, min_turnovers, window_turnovers, tolerance, gate_scalar, gate_skewness, skewness_floor):
        self.min_turnovers = float(min_turnovers)
        self.window_turnovers = float(window_turnovers)
        self.tolerance = float(tolerance)
        self.gate_scalar = bool(gate_scalar)
        self.gate_skewness = bool(gate_skewness)
        self.skewness_floor = float(skewness_floor)
        self.sample_gap = self.window_turnovers / 40
        self.turn, self.eps, self.phi2, self.skew = [], [], [], []
        self.converged = False

    @staticmethod
    def is_flat(turn, y, now, W, tolerance) -> bool:
        mid = now - W / 2
        s1 = s2 = 0.0
        n1 = n2 = 0
        for t, v in zip(turn, y):
            if t < now - W:
                continue
            if t < mid:
                s1 += v
                n1 += 1
            else:
                s2 += v
                n2 += 1
        if n1 == 0 or n2 == 0:
            return False
        mean_all = (s1 + s2) / (n1 + n2)
        return abs(s2 / n2 - s1 / n1) <= tolerance * abs(mean_all)

    @staticmethod
    def window_mean(turn, y, now, W) -> float:
        vals = [v for t, v in zip(turn, y) if t >= now - W]
        return sum(vals) / len(vals) if vals else 0.0

    def due(self, elapsed_turn) -> bool:
        return not self.turn or elapsed_turn - self.turn[-1] >= self.sample_gap

    def push_and_test(self, elapsed_turn, eps, phi2, skew) -> bool:
        if self.due(elapsed_turn):
            for series, v in ((self.turn, elapsed_turn), (self.eps, eps), (self.phi2, phi2), (self.skew, skew)):
                series.append(v)
            while len(self.turn) > 2 and self.turn[0] < elapsed_turn - self.window_turnovers - self.sample_gap:
                for series in (self.turn, self.eps, self.phi2, self.skew):
                    series.pop(0)
        if elapsed_turn < self.min_turnovers:
            return False
        if not self.turn[0] <= elapsed_turn - self.window_turnovers:
            return False
        W, tol = self.window_turnovers, self.tolerance
        flat_v = self.is_flat(self.turn, self.eps, elapsed_turn, W, tol)
        flat_s = not self.gate_scalar or self.is_flat(self.turn, self.phi2, elapsed_turn, W, tol)
        flat_k = not self.gate_skewness or (self.is_flat(self.turn, self.skew, elapsed_turn, W, tol) and
                                            abs(self.window_mean(self.turn, self.skew, elapsed_turn, W)) >= self.skewness_floor)
        self.converged = flat_v and flat_s and flat_k
        return self.converged


class StationarityCriterion:
    """run() callback that spins a run up to a statistically stationary state
    (src/SimulationRun.jl:1560-1628): the eddy-turnover clock advances by dt * eps / k per
    iteration; the run stops once the dissipation (and, for a forced scalar, <phi^2>; with
    gate_skewness, the skewness, at least skewness_floor in size) is flat over the trailing window,
    or at max_turnovers.  A scalar driven by the imposed mean gradient counts as forced."""

    def __init__(self, *, gate_scalar: bool = True, gate_skewness: bool = False, min_turnovers: float = 5,
                 window_turnovers: float = 4, tolerance: float = 0.05, skewness_floor: float = 0.1,
                 max_turnovers: float = 60):
        self.monitor = StationarityMonitor(min_turnovers=min_turnovers, window_turnovers=window_turnovers,
                                           tolerance=tolerance, gate_scalar=gate_scalar, gate_skewness=gate_skewness,
                                           skewness_floor=skewness_floor)
        self.max_turnovers = float(max_turnovers)
        self.gate_scalar = bool(gate_scalar)
        self.elapsed_turn = 0.0
        self.last_time = math.nan
        self.started = False

    def __call__(self, sim: Simulation) -> bool:
        if not self.started:
            forced = sim.carries_scalar and (getattr(sim, "mean_gradient", 0.0) != 0.0 or
                                             not isinstance(sim.conditions.scalar_forcing, NoForcing))
            self.monitor.gate_scalar = self.gate_scalar and forced
            self.last_time = sim.time
            self.started = True
        r = sim.reductions()
        k, eps = r["kinetic_energy"], r["dissipation"]
        dt = sim.time - self.last_time
        self.last_time = sim.time
        if eps > 0:
            self.elapsed_turn += dt * eps / k
        else:
            warnings.warn(f"StationarityCriterion: the dissipation rate is not positive (eps = {eps}), so the "
                          "eddy-turnover clock cannot advance and stationarity can never be detected -- stopping. "
                          "A run with no dissipation has no cascade to become stationary.")
            return False
        m = self.monitor
        if m.due(self.elapsed_turn):
            phi2 = (r["scalar_variance"] if sim.carries_scalar else math.nan) if m.gate_scalar else 0.0
            skew = sim.derivative_skewness() if m.gate_skewness else 0.0
            if m.push_and_test(self.elapsed_turn, eps, phi2, skew):
                return False
        if self.elapsed_turn >= self.max_turnovers:
            warnings.warn(f"StationarityCriterion: hit the max_turnovers = {self.max_turnovers} cap without "
                          "detecting stationarity")
            return False
        return True


def warmup(sim: Simulation, *, dt: float | None = None, **kwargs) -> Simulation:
    """warmup!(sim; dt, kwargs...) = run!(StationarityCriterion(; kwargs...), sim; dt)
    (src/SimulationRun.jl:1630-1660)."""
    return sim.run(StationarityCriterion(**kwargs), dt=dt)
