"""FP64 restatement of the reference DNS step (test infrastructure only; see oracle/__init__.py).

Array conventions follow the field library's layout as the reference's tests reveal it
(test/regressions/r5-free-modules-01-mean-mode.jl:21,44; r7-reality-projection-forced-run.jl:20-29):
Julia column-major ``Complex{T}[N/2+1, N, N, c]`` is numpy C-order ``[c, kz, ky, kx]`` with
``kx`` the folded (rfft) axis, and a physical vector ``T[N, N, N, c]`` is ``[c, z, y, x]``.
The forward transform is unnormalised and the inverse carries 1/N^3
(docs/src/validation.md:73-74).  Frequencies use ``wn(i, n) = i <= n//2 ? i : i - n``
(Nyquist at +N/2, test/regressions/r6-inclusive-ball-exact.jl:29).

Canonical choices where the reference delegates to the field library and its source is not
available (each one is made identically by the HIP library):
  * 2/3 mask: keep a mode iff 9*(mx^2+my^2+mz^2) <= N^2, the closed ball of radius
    (2/3)(N/2)(2pi/L) in exact integers (src/SimulationDomain.jl:342-376); PhaseShift's ball
    9|m|^2 <= 2 N^2; HouLi's filter exp(-s (|m|/(N/2))^p) evaluated in FP64.
  * shell band of shells lo..hi: 4|m|^2 >= (2lo-1)^2 and 4|m|^2 < (2hi+1)^2 (src/Forcing.jl:435-439);
    no lattice mode lies on an edge, so the open/closed choice is immaterial.
  * spectrum shell of a mode: the integer n with (2n-1)^2 <= 4|m|^2 < (2n+1)^2.
  * derivative multiplier i*kappa uses the wn() value on every axis (masked modes only
    ever meet it, so the Nyquist convention never reaches a dealiased state).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from fractions import Fraction

import numpy as np

try:  # pocketfft with threads; numpy's pocketfft is the fallback
    import scipy.fft as _sfft
except ImportError:  # pragma: no cover
    _sfft = None

_WORKERS = 1


def set_workers(n: int) -> None:
    """Threads the oracle's transforms may use (the cpu_baseline leg sets all host cores)."""
    global _WORKERS
    _WORKERS = max(1, int(n))


def rfft3(a: np.ndarray) -> np.ndarray:
    """Forward real 3D transform over the last three axes, unnormalised (PF ``to_spectral!``)."""
    if _sfft is not None:
        return _sfft.rfftn(a, axes=(-3, -2, -1), workers=_WORKERS)
    return np.fft.rfftn(a, axes=(-3, -2, -1))


def irfft3(a: np.ndarray, n: int) -> np.ndarray:
    """Inverse real 3D transform with the 1/N^3 factor (PF ``to_physical!``).

    The complex passes over z and y run first and the c2r pass over x last, so the imaginary
    part of the kx = 0 (and kx = N/2) coefficients after the first two passes is ignored: an
    anti-Hermitian part on those planes is invisible to the physical field, as
    src/Propagator.jl:403-410 describes.
    """
    if _sfft is not None:
        return _sfft.irfftn(a, s=(n, n, n), axes=(-3, -2, -1), workers=_WORKERS)
    return np.fft.irfftn(a, s=(n, n, n), axes=(-3, -2, -1))


def wn(n: int) -> np.ndarray:
    """Signed integer frequency of each FFT bin (Nyquist at +n/2)."""
    i = np.arange(n)
    return np.where(i <= n // 2, i, i - n)


class Domain:
    """The reference ``FixedDomain`` tables (src/SimulationDomain.jl:295-328).

    ``kappa`` per axis is (2pi/L)*m; ``k2`` is the sum of the per-axis squares (the reference's
    squared_wavenumber_{1,2,3}, test/scalar.jl:55-56).  ``mask`` is the canonical closed 2/3
    ball (a18); ``weights`` are the Parseval weights of the half spectrum (2 on interior kx
    planes, 1 on kx = 0 and kx = N/2).
    """

    def __init__(self, n: int, side_length: float = 2 * math.pi, dealiasing: str = "two_thirds",
                 filter_order: int = 36, filter_strength: float | None = None, element_type=np.float64):
        """``dealiasing`` is "two_thirds", "hou_li" (with ``filter_order`` and ``filter_strength``,
        None for -log(eps(element_type))) or "phase_shift" (src/SimulationDomain.jl:44-195)."""
        if dealiasing not in ("two_thirds", "hou_li", "phase_shift"):
            raise ValueError(f"unknown dealiasing {dealiasing!r}")
        self.dealiasing = dealiasing
        self.n = int(n)
        self.side_length = float(side_length)
        self.k0 = 2 * math.pi / self.side_length
        self.mx = np.arange(self.n // 2 + 1)
        self.my = wn(self.n)
        self.mz = wn(self.n)
        self.kx = self.k0 * self.mx
        self.ky = self.k0 * self.my
        self.kz = self.k0 * self.mz
        mx = self.mx[None, None, :]
        my = self.my[None, :, None]
        mz = self.mz[:, None, None]
        self.m2 = (mx * mx + my * my) + mz * mz
        kx = self.kx[None, None, :]
        ky = self.ky[None, :, None]
        kz = self.kz[:, None, None]
        self.kvec = (np.broadcast_to(kx, self.m2.shape),
                     np.broadcast_to(ky, self.m2.shape),
                     np.broadcast_to(kz, self.m2.shape))
        self.k2 = (kx * kx + ky * ky) + kz * kz
        # the retained ball: the closed 2/3 ball, PhaseShift's ball of radius sqrt(2) N/3
        # (src/SimulationDomain.jl:344-346), or every mode under HouLi
        if dealiasing == "hou_li":
            self.mask = np.ones(self.m2.shape, dtype=bool)
            s = -math.log(float(np.finfo(element_type).eps)) if filter_strength is None else float(filter_strength)
            self.filter_strength = s
            # src/SimulationDomain.jl:365-376: exp(-strength (|k|/k_Ny)^order), |k|/k_Ny = |m|/(N/2)
            self.dealias = np.exp(-s * (np.sqrt(self.m2.astype(np.float64)) / (self.n / 2)) ** int(filter_order))
        else:
            self.mask = (9 * self.m2 <= (2 if dealiasing == "phase_shift" else 1) * self.n * self.n)
            self.dealias = self.mask.astype(np.float64)
        w = np.full(self.n // 2 + 1, 2.0)
        w[0] = 1.0
        if self.n % 2 == 0:
            w[-1] = 1.0
        self.weights = np.broadcast_to(w[None, None, :], self.m2.shape)
        self.norm = float(self.n) ** 6

    @property
    def max_retained_wavenumber(self) -> float:
        """f_ret (N/2)(2pi/L), f_ret = 2/3, 1 (HouLi) or 2 sqrt 2 / 3 (PhaseShift) (src/SimulationDomain.jl:181-195)."""
        f = {"two_thirds": 2.0 / 3.0, "hou_li": 1.0, "phase_shift": 2.0 * math.sqrt(2.0) / 3.0}[self.dealiasing]
        return f * (self.n / 2) * self.k0

    def half_cell_shift(self, sign: int) -> np.ndarray:
        """e^{i sign kappa . delta} with delta half a cell along every axis, kappa . delta =
        pi (mx + my + mz) / N: sign +1 samples a field on the displaced grid, -1 shifts back
        (PhaseShift's translation kernels, src/Generator.jl:166-175).  The dealiased average does not
        depend on the sign convention: the aliases it cancels carry (-1)^(wraps) either way."""
        msum = self.mx[None, None, :] + self.my[None, :, None] + self.mz[:, None, None]
        return np.exp(1j * sign * np.pi * msum / self.n)

    def shell_band(self, lowest: int, highest) -> np.ndarray:
        """Indicator of shells lowest..highest, [h-1/2, h+1/2) in units of 2pi/L (src/Forcing.jl:428-445)."""
        four_m2 = 4 * self.m2
        lo = (2 * lowest - 1) ** 2
        if highest == math.inf:
            return four_m2 >= lo
        hi = (2 * int(highest) + 1) ** 2
        return (four_m2 >= lo) & (four_m2 < hi)


# --- field-library operators (restated from their contracts) -------------------------------

def curl(d: Domain, uh: np.ndarray) -> np.ndarray:
    """i kappa x u_hat (PF ``curl!`` inside ``vorticity_physical!``, src/Observables.jl:329-333)."""
    kx, ky, kz = d.kvec
    ux, uy, uz = uh
    return np.stack([1j * (ky * uz - kz * uy),
                     1j * (kz * ux - kx * uz),
                     1j * (kx * uy - ky * ux)])


def leray(d: Domain, f: np.ndarray) -> np.ndarray:
    """P(k) f = f - k (k.f)/|k|^2, identity at k = 0 (PF ``helmholtz_projection!``)."""
    kx, ky, kz = d.kvec
    k2 = d.k2.copy()
    k2[0, 0, 0] = 1.0
    s = (kx * f[0] + ky * f[1] + kz * f[2]) / k2
    s[0, 0, 0] = 0.0
    return np.stack([f[0] - kx * s, f[1] - ky * s, f[2] - kz * s])


def weighted_inner(d: Domain, f: np.ndarray, w=None) -> float:
    """<f, w f> with the normalised inner product, by Parseval over the half spectrum.

    PF ``weighted_inner`` (src/Forcing.jl:1109, src/Observables.jl:84,170-173).  The oracle sums
    in FP64 with math.fsum per component (exactly rounded), the compensated baseline the
    device reductions are held to.
    """
    a = np.abs(f) ** 2 * d.weights
    if w is not None:
        a = a * w
    if a.ndim == 4:
        return math.fsum(math.fsum(np.ravel(a[c])) for c in range(a.shape[0])) / d.norm
    return math.fsum(np.ravel(a)) / d.norm


def reality_projection(d: Domain, f: np.ndarray) -> np.ndarray:
    """Make the self-conjugate planes Hermitian (PF ``reality_projection!``, src/Propagator.jl:411-412).

    On kx = 0 (and kx = N/2 for even N) each pair becomes (f(k) + conj f(-k))/2.
    """
    g = f.copy()
    n = d.n
    neg = (-np.arange(n)) % n
    planes = [0] + ([n // 2] if n % 2 == 0 else [])
    for p in planes:
        s = g[..., p]                     # [.., kz, ky]
        partner = s[..., neg, :][..., :, neg]
        g[..., p] = 0.5 * (s + np.conj(partner))
    return g


def max_speed(u: np.ndarray) -> float:
    """sqrt(max over points of |u|^2) (src/Observables.jl:304-305)."""
    return float(np.sqrt(np.max(u[0] * u[0] + u[1] * u[1] + u[2] * u[2])))


# --- physics ------------------------------------------------------------------------------

@dataclass
class Conditions:
    """``FlowConditions`` (src/Generator.jl:81-147) plus the forcing and the mean-gradient extension.

    forcing: None (NoForcing) or (injection_rate, lowest_shell, highest_shell) for ConstantPower
    (src/Forcing.jl:714-719,1069-1117).  mean_gradient: G of the imposed uniform mean scalar
    gradient along z (extension: src/Generator.jl:652-653 hands a scalar forcing only phi_hat,
    and docs/ARCHITECTURE.md:769-771 defers the mean gradient; the source is -G u_z).
    band_eps: eps(T) of the run's element type, used by the empty-band clamp.
    """
    viscosity: float
    schmidt: float = 1.0
    forcing: tuple | None = None
    mean_gradient: float = 0.0
    linear: tuple | None = None       # Linear forcing (rate A, lowest_shell, highest_shell)
    # VarianceController (target k_inf, gain G, turnover_time t_l, lowest_shell, highest_shell)
    controller: tuple | None = None
    # VarianceRescaling (target, over "band" | "total", lowest_shell, highest_shell)
    rescaling: tuple | None = None
    # the scalar forcing (src/Generator.jl:652-653, the same strategies with the variance factor 1
    # and the diffusivity nu/Sc, src/Forcing.jl:456): one of
    #   ("constant_power", injection_rate, lo, hi), ("linear", rate, lo, hi),
    #   ("controller", target, gain, turnover_time, lo, hi), ("rescaling", target, over, lo, hi)
    scalar_forcing: tuple | None = None
    # an additive force held over the step (StochasticOU's band force, src/Forcing.jl:1404-1405), in
    # the unnormalised spectral convention: [3, N, N, N/2+1] for the velocity, [N, N, N/2+1] for phi
    velocity_source: np.ndarray | None = None
    scalar_source: np.ndarray | None = None
    band_eps: float = float(np.finfo(np.float64).eps)

    @property
    def scalar_diffusivity(self) -> float:
        return self.viscosity / self.schmidt


def linear_factor(d: Domain, cond: Conditions, sub: float) -> np.ndarray:
    """The velocity's integrating factor over a substep: exp(-nu |k|^2 sub), with a Linear forcing's
    rate A folded in on its band, exp((A - nu |k|^2) sub) (src/Generator.jl:565-581,
    src/Forcing.jl:600-642)."""
    expo = -cond.viscosity * sub * d.k2
    if cond.linear is not None:
        rate, lo, hi = cond.linear
        expo = expo + rate * sub * d.shell_band(lo, hi)
    return np.exp(expo)


def scalar_factor(d: Domain, cond: Conditions, sub: float) -> np.ndarray:
    """The scalar's integrating factor over a substep: exp(-(nu/Sc) |k|^2 sub), with a Linear scalar
    forcing's rate folded in on its band (src/Generator.jl:583-595)."""
    expo = -cond.scalar_diffusivity * sub * d.k2
    sf = cond.scalar_forcing
    if sf is not None and sf[0] == "linear":
        expo = expo + sf[1] * sub * d.shell_band(sf[2], sf[3])
    return np.exp(expo)


def scalar_amplitude(d: Domain, cond: Conditions, ph: np.ndarray) -> float:
    """A of a ConstantPower or VarianceController scalar forcing at the stage scalar, with the
    variance factor 1: eps / (2 <phi^2>_band), or (2 kappa <phi, |k|^2 phi> - G (<phi^2> - target) / t_l)
    / (2 <phi^2>_band); 0 for an empty band (src/Forcing.jl:1041-1117)."""
    sf = cond.scalar_forcing
    lo, hi = sf[-2], sf[-1]
    band_obs = weighted_inner(d, ph, d.shell_band(lo, hi))
    whole = weighted_inner(d, ph)
    if sf[0] == "constant_power":
        num = sf[1]
    else:
        _, target, gain, tl, _, _ = sf
        num = 2 * cond.scalar_diffusivity * weighted_inner(d, ph, d.k2) - gain * (whole - target) / tl
    if band_is_empty(band_obs, whole, cond.band_eps):
        return 0.0
    return num / (2 * band_obs)


def scalar_rescaling(d: Domain, cond: Conditions, ph: np.ndarray) -> np.ndarray:
    """VarianceRescaling of the scalar after the step (src/Forcing.jl:575-596, variance factor 1)."""
    _, target, over, lo, hi = cond.scalar_forcing
    band = d.shell_band(lo, hi)
    band_obs = weighted_inner(d, ph, band)
    whole = weighted_inner(d, ph)
    if band_is_empty(band_obs, whole, cond.band_eps):
        return ph
    target_band = max(0.0, target - whole + band_obs) if over == "total" else target
    return ph + (math.sqrt(target_band / band_obs) - 1) * band * ph


def finalize_fields(d: Domain, cond: Conditions, uh: np.ndarray, ph: np.ndarray | None):
    """finalize! (the post-step projections of the forcings), then the state Leray projection and
    the reality projection of both fields (src/Propagator.jl:394-412)."""
    if cond.rescaling is not None:
        uh = variance_rescaling(d, cond, uh)
    if ph is not None and cond.scalar_forcing is not None and cond.scalar_forcing[0] == "rescaling":
        ph = scalar_rescaling(d, cond, ph)
    uh = reality_projection(d, leray(d, uh))
    if ph is not None:
        ph = reality_projection(d, ph)
    return uh, ph


def band_is_empty(band_value: float, whole_value: float, eps: float) -> bool:
    """src/Forcing.jl:110-111."""
    return not (math.isfinite(band_value) and band_value > 10 * eps * whole_value)


def constant_power_amplitude(d: Domain, cond: Conditions, uh: np.ndarray) -> float:
    """A = eps/(2 k_band) at the stage field, 0 for an empty band (src/Forcing.jl:1091-1117)."""
    rate, lo, hi = cond.forcing
    band = d.shell_band(lo, hi)
    band_obs = 0.5 * weighted_inner(d, uh, band)
    whole_obs = 0.5 * weighted_inner(d, uh)
    if band_is_empty(band_obs, whole_obs, cond.band_eps):
        return 0.0
    return rate / (2 * band_obs)


def controller_amplitude(d: Domain, cond: Conditions, uh: np.ndarray) -> float:
    """A = (eps - G (k - k_inf) / t_l) / (2 k_band) at the stage field, with eps = nu <u, |k|^2 u>,
    k the whole field's and k_band the band's kinetic energy; 0 for an empty band
    (src/Forcing.jl:1041-1065, the clamp :1093-1103)."""
    target, gain, tl, lo, hi = cond.controller
    band = d.shell_band(lo, hi)
    whole = 0.5 * weighted_inner(d, uh)
    band_obs = 0.5 * weighted_inner(d, uh, band)
    eps = cond.viscosity * weighted_inner(d, uh, d.k2)
    restoring = gain * (whole - target) / tl
    if band_is_empty(band_obs, whole, cond.band_eps):
        return 0.0
    return (eps - restoring) / (2 * band_obs)


def variance_rescaling(d: Domain, cond: Conditions, uh: np.ndarray) -> np.ndarray:
    """The post-step rescale of the band (src/Forcing.jl:575-596): s = sqrt(target_band / k_band)
    with target_band = target (over "band") or max(0, target - k + k_band) (over "total"); band
    modes become u + (s - 1) u; an empty band is left alone."""
    target, over, lo, hi = cond.rescaling
    band = d.shell_band(lo, hi)
    band_obs = 0.5 * weighted_inner(d, uh, band)
    whole = 0.5 * weighted_inner(d, uh)
    if band_is_empty(band_obs, whole, cond.band_eps):
        return uh
    target_band = max(0.0, target - whole + band_obs) if over == "total" else target
    scale = math.sqrt(target_band / band_obs)
    return uh + (scale - 1) * band * uh


def velocity_rhs(d: Domain, cond: Conditions, uh: np.ndarray, u: np.ndarray, w: np.ndarray) -> np.ndarray:
    """N_u = P mask FFT(u x w), mean mode zeroed, plus the ConstantPower source.

    src/Generator.jl:484-495 (velocity_rhs!), :498-501 (mask strategies), :528-534
    (_velocity_advective_term!: cross!(w, u, w) then to_spectral! then projection with the
    dealias kernel folded in), mean_free_projection! at :486, add_forcing_source! after it.
    """
    c = np.cross(u, w, axis=0)
    f = rfft3(c)
    if d.dealiasing == "phase_shift":   # src/Generator.jl:502-519: average with the displaced evaluation
        us_hat = uh * d.half_cell_shift(+1)
        us, ws = irfft3(us_hat, d.n), irfft3(curl(d, us_hat), d.n)
        f = 0.5 * (f + d.half_cell_shift(-1) * rfft3(np.cross(us, ws, axis=0)))
    nl = leray(d, f * d.dealias)
    nl[:, 0, 0, 0] = 0.0
    if cond.forcing is not None:
        amp = constant_power_amplitude(d, cond, uh)
        band = d.shell_band(cond.forcing[1], cond.forcing[2])
        nl = nl + amp * band * uh
    if cond.controller is not None:
        amp = controller_amplitude(d, cond, uh)
        nl = nl + amp * d.shell_band(cond.controller[3], cond.controller[4]) * uh
    if cond.velocity_source is not None:
        nl = nl + cond.velocity_source
    return nl


def scalar_rhs(d: Domain, cond: Conditions, uh: np.ndarray, u: np.ndarray, phi: np.ndarray,
               ph: np.ndarray | None = None) -> np.ndarray:
    """N_phi = -mask i kappa . FFT(u phi) - G u_z_hat, plus a scalar forcing's source A 1_band phi_hat.

    src/Generator.jl:648-661 (scalar_rhs!), :663-666 (mask strategies, alpha = -1), :708-730
    (_scalar_flux_divergence!, _scalar_flux!).  The -G u_z term is the mean-gradient extension
    (the reference has none), evaluated at the stage velocity like every other source.  ``ph`` is
    the stage scalar in spectral space, needed by a ConstantPower or VarianceController scalar forcing.
    """
    flux = u * phi[None]
    f = rfft3(flux)
    if d.dealiasing == "phase_shift":   # src/Generator.jl:675-704: average with the displaced flux
        sh = d.half_cell_shift(+1)
        us, phs = irfft3(uh * sh, d.n), irfft3(ph * sh, d.n)
        f = 0.5 * (f + d.half_cell_shift(-1) * rfft3(us * phs[None]))
    kx, ky, kz = d.kvec
    div = 1j * (kx * f[0] + ky * f[1] + kz * f[2])
    nl = -(d.dealias * div)
    if cond.mean_gradient != 0.0:
        nl = nl - cond.mean_gradient * uh[2]
    sf = cond.scalar_forcing
    if sf is not None and sf[0] in ("constant_power", "controller"):
        nl = nl + scalar_amplitude(d, cond, ph) * d.shell_band(sf[-2], sf[-1]) * ph
    if cond.scalar_source is not None:
        nl = nl + cond.scalar_source
    return nl


# --- the scheme ---------------------------------------------------------------------------

def _round(x: Fraction, dtype) -> float:
    """The element-type value T(p//q): the rational rounded once, to nearest (ties to even)."""
    v = x.numerator / x.denominator          # Python int division is correctly rounded to FP64
    if dtype is not np.float32:
        return v
    c = np.float32(v)
    best = None
    for cand in (np.nextafter(c, np.float32(-np.inf)), c, np.nextafter(c, np.float32(np.inf))):
        err = abs(Fraction(float(cand)) - x)
        key = (err, int(np.array(cand).view(np.uint32)) & 1)
        if best is None or key < best[0]:
            best = (key, cand)
    return float(best[1])


def scheme_coefficients(scheme: str, dtype=np.float64):
    """(a, b, d) of the low-storage schemes, rounded to the element type like the reference.

    IFLowStorageRK3: src/Propagator.jl:264-274 (Williamson).  IFLowStorageRK4:
    src/Propagator.jl:304-320 (Carpenter-Kennedy rationals; d = differences of the rounded
    abscissae, formed in T).
    """
    F = Fraction
    if scheme == "rk3":
        a = (F(0), F(-5, 9), F(-153, 128))
        b = (F(1, 3), F(15, 16), F(8, 15))
        dd = (F(1, 3), F(5, 12), F(1, 4))
        return tuple(_round(x, dtype) for x in a), tuple(_round(x, dtype) for x in b), tuple(_round(x, dtype) for x in dd)
    if scheme == "rk4":
        a = (F(0), F(-567301805773, 1357537059087), F(-2404267990393, 2016746695238),
             F(-3550918686646, 2091501179385), F(-1275806237668, 842570457699))
        b = (F(1432997174477, 9575080441755), F(5161836677717, 13612068292357),
             F(1720146321549, 2090206949498), F(3134564353537, 4481467310338),
             F(2277821191437, 14882151754819))
        c2 = dtype(_round(F(1432997174477, 9575080441755), dtype))
        c3 = dtype(_round(F(2526269341429, 6820363962896), dtype))
        c4 = dtype(_round(F(2006345519317, 3224310063776), dtype))
        c5 = dtype(_round(F(2802321613138, 2924317926251), dtype))
        one = dtype(1)
        dd = (float(c2), float(dtype(c3 - c2)), float(dtype(c4 - c3)), float(dtype(c5 - c4)), float(dtype(one - c5)))
        return tuple(_round(x, dtype) for x in a), tuple(_round(x, dtype) for x in b), dd
    raise ValueError(f"unknown scheme {scheme!r}")


@dataclass
class Deposit:
    """The field deposit (src/Propagator.jl:808-842): physical u, w, phi and max|u|."""
    u: np.ndarray
    w: np.ndarray
    phi: np.ndarray | None
    max_speed: float


def deposit(d: Domain, uh: np.ndarray, ph: np.ndarray | None) -> Deposit:
    """Recover u, w = curl u and phi in physical space, then max_speed (src/Propagator.jl:808-842)."""
    u = irfft3(uh, d.n)
    w = irfft3(curl(d, uh), d.n)
    phi = irfft3(ph, d.n) if ph is not None else None
    return Deposit(u, w, phi, max_speed(u))


def step(d: Domain, cond: Conditions, scheme: str, uh: np.ndarray, ph: np.ndarray | None,
         dt: float, dep: Deposit | None = None, coeff_dtype=np.float64):
    """One ``step!`` of the 2N low-storage loop, then the closing field ``deposit!``.

    src/Propagator.jl:326-419 (_low_storage_2n_step!): per stage i
        h  <- a_i h + dt N(q)          (a_1 = 0)
        q  <- E(d_i dt) (q + b_i h)
        h  <- E(d_i dt) h              (skipped at the last stage)
    for u_hat and phi_hat in lockstep, both right-hand sides evaluated at the stage state
    before either field moves; E = exp(-D |k|^2 d_i dt) with D = nu or nu/Sc
    (src/Generator.jl:565-581, src/Propagator.jl:140-150).  After the stages: Leray
    projection of the state (:401) and the reality projection of both fields (:411-412).
    Returns (u_hat, phi_hat, deposit).
    """
    if scheme == "ssprk3":
        return ssp_step(d, cond, uh, ph, dt, dep, coeff_dtype)
    if scheme == "heun":
        return heun_step(d, cond, uh, ph, dt, dep)
    a, b, dd = scheme_coefficients(scheme, coeff_dtype)
    S = len(a)
    uh = uh.copy()
    ph = ph.copy() if ph is not None else None
    if dep is None:
        dep = deposit(d, uh, ph)
    hu = None
    hp = None
    for i in range(S):
        if i == 0:
            u, w, phi = dep.u, dep.w, dep.phi
        else:
            u = irfft3(uh, d.n)
            w = irfft3(curl(d, uh), d.n)
            phi = irfft3(ph, d.n) if ph is not None else None
        nu_hat = velocity_rhs(d, cond, uh, u, w)
        np_hat = scalar_rhs(d, cond, uh, u, phi, ph) if ph is not None else None
        hu = dt * nu_hat if i == 0 else a[i] * hu + dt * nu_hat
        sub = dd[i] * dt
        eu = linear_factor(d, cond, sub)
        uh = eu * (uh + b[i] * hu)
        if i < S - 1:
            hu = eu * hu
        if ph is not None:
            hp = dt * np_hat if i == 0 else a[i] * hp + dt * np_hat
            ep = scalar_factor(d, cond, sub)
            ph = ep * (ph + b[i] * hp)
            if i < S - 1:
                hp = ep * hp
    uh, ph = finalize_fields(d, cond, uh, ph)
    return uh, ph, deposit(d, uh, ph)


def ssp_step(d: Domain, cond: Conditions, uh: np.ndarray, ph: np.ndarray | None, dt: float,
             dep: Deposit | None = None, coeff_dtype=np.float64):
    """One IFSSPRK3 ``step!`` (Shu-Osher), then the closing deposit (src/Propagator.jl:599-691).

    With E(t) the integrating factor over t and N(q) the right-hand side at the stage state q:
        q1    = E(dt) (u + dt N(u))
        q2    = 3/4 E(dt/2) u + 1/4 E(-dt/2) (q1 + dt N(q1))
        u_new = 1/3 E(dt) u + 2/3 E(dt/2) (q2 + dt N(q2))
    for u_hat and phi_hat in lockstep (the scalar's factor uses nu/Sc), then the state Leray
    projection (:677) and the reality projection of both fields (:687-688).  The weights are the
    element type's roundings of 3/4, 1/4, 1/3, 2/3.
    """
    w34, w14, w13, w23 = (_round(Fraction(p, q), coeff_dtype) for p, q in ((3, 4), (1, 4), (1, 3), (2, 3)))
    if dep is None:
        dep = deposit(d, uh, ph)

    def factors(t):
        return linear_factor(d, cond, t), scalar_factor(d, cond, t)

    def rhs(q, p, first):
        if first:
            u, w, phi = dep.u, dep.w, dep.phi
        else:
            u, w = irfft3(q, d.n), irfft3(curl(d, q), d.n)
            phi = irfft3(p, d.n) if p is not None else None
        nu_hat = velocity_rhs(d, cond, q, u, w)
        np_hat = scalar_rhs(d, cond, q, u, phi, p) if p is not None else None
        return nu_hat, np_hat

    eu_full, ep_full = factors(dt)
    eu_half, ep_half = factors(dt / 2)
    eu_back, ep_back = factors(-dt / 2)
    n0, m0 = rhs(uh, ph, True)
    q = eu_full * (uh + dt * n0)
    p = ep_full * (ph + dt * m0) if ph is not None else None
    n1, m1 = rhs(q, p, False)
    q = w14 * (eu_back * (q + dt * n1)) + eu_half * (w34 * uh)
    if ph is not None:
        p = w14 * (ep_back * (p + dt * m1)) + ep_half * (w34 * ph)
    n2, m2 = rhs(q, p, False)
    q = eu_full * (w13 * uh) + w23 * (eu_half * (q + dt * n2))
    if ph is not None:
        p = ep_full * (w13 * ph) + w23 * (ep_half * (p + dt * m2))
    q, p = finalize_fields(d, cond, q, p)   # finalize! before the state projections (:674-688)
    return q, p, deposit(d, q, p)


def heun_step(d: Domain, cond: Conditions, uh: np.ndarray, ph: np.ndarray | None, dt: float,
              dep: Deposit | None = None):
    """One IFHeun ``step!``, then the closing deposit (src/Propagator.jl:456-544):
        u* = E(dt) (u + dt N(u)),   u_new = E(dt) u + dt/2 (E(dt) N(u) + N(u*))
    for u_hat and phi_hat in lockstep, then finalize! and the state projections (:519-538)."""
    if dep is None:
        dep = deposit(d, uh, ph)
    eu, ep = linear_factor(d, cond, dt), scalar_factor(d, cond, dt)
    n1 = velocity_rhs(d, cond, uh, dep.u, dep.w)
    m1 = scalar_rhs(d, cond, uh, dep.u, dep.phi, ph) if ph is not None else None
    us = eu * (uh + dt * n1)
    ps = ep * (ph + dt * m1) if ph is not None else None
    u, w = irfft3(us, d.n), irfft3(curl(d, us), d.n)
    n2 = velocity_rhs(d, cond, us, u, w)
    m2 = scalar_rhs(d, cond, us, u, irfft3(ps, d.n), ps) if ph is not None else None
    q = eu * uh + dt / 2 * (eu * n1) + dt / 2 * n2
    p = ep * ph + dt / 2 * (ep * m1) + dt / 2 * m2 if ph is not None else None
    q, p = finalize_fields(d, cond, q, p)
    return q, p, deposit(d, q, p)


# --- observables --------------------------------------------------------------------------

def reductions(d: Domain, cond: Conditions, uh: np.ndarray, ph: np.ndarray | None) -> dict:
    """k, eps, <phi^2>, eps_phi (src/Observables.jl:84, :155-158, :345-400)."""
    out = {
        "kinetic_energy": 0.5 * weighted_inner(d, uh),
        "dissipation": cond.viscosity * weighted_inner(d, uh, d.k2),
    }
    if ph is not None:
        out["scalar_variance"] = weighted_inner(d, ph)
        out["scalar_dissipation"] = 2 * cond.scalar_diffusivity * weighted_inner(d, ph, d.k2)
    return out


def shell_index(d: Domain) -> np.ndarray:
    """Canonical shell of each mode: n - 1/2 <= |m| < n + 1/2, in exact integers (n >= 0)."""
    four = 4 * d.m2
    n = np.floor(np.sqrt(four.astype(np.float64)) / 2 + 0.5).astype(np.int64)
    # exact integer correction of the float estimate
    n = np.where((n > 0) & ((2 * n - 1) ** 2 > four), n - 1, n)
    n = np.where((2 * n + 1) ** 2 <= four, n + 1, n)
    return n


def energy_spectrum(d: Domain, f: np.ndarray, vector: bool = True):
    """E(kappa_n) = (shell sum of the weighted |f_hat|^2 / N^6) / (2 dk) (src/Observables.jl:502-514).

    For a scalar field (vector=False) the factor 1/2 is dropped (the variance spectrum).
    Returns (shell centres kappa_n, density).
    """
    sh = shell_index(d)
    nmax = int(sh.max())
    a = np.abs(f) ** 2 * d.weights
    if a.ndim == 4:
        a = a.sum(axis=0)
    spec = np.bincount(sh.ravel(), weights=a.ravel(), minlength=nmax + 1) / d.norm
    dk = d.k0
    factor = 2 * dk if vector else dk
    return np.arange(nmax + 1) * dk, spec / factor


def divergence_residual(d: Domain, uh: np.ndarray) -> float:
    """max |k . u_hat| / max |k||u_hat| -- zero for a solenoidal state."""
    kx, ky, kz = d.kvec
    div = np.abs(kx * uh[0] + ky * uh[1] + kz * uh[2])
    scale = np.sqrt(d.k2) * np.sqrt(np.sum(np.abs(uh) ** 2, axis=0))
    return float(div.max() / max(scale.max(), 1e-300))
