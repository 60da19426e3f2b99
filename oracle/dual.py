"""Forward-mode (dual-number) derivative of the oracle step (test infrastructure only).

A dual number x + eps x' with eps^2 = 0 carries a value and a tangent; evaluating the forward step
on duals gives the step's output and its derivative along the seeded tangent, with no truncation
error and no transposition.  That makes it an independent pin of the adjoint: for any direction
v, the reverse-mode gradient g must satisfy <g, v>_w = dJ(u0 + eps v)/d eps to round-off.

`Dual` holds numpy arrays (or Python floats) for the value and the tangent.  Linear maps (the
transforms, curl, Leray, the masks, the reality projection, the integrating factor) act on both
parts; products follow the product rule; the ConstantPower amplitude A = eps/(2 k_band) follows
the quotient rule, with the empty-band clamp decided on the value (a constant 0 has no tangent).
`step` restates dns.step (src/Propagator.jl:326-419) operation for operation on duals, so its value
part reproduces dns.step and its tangent part is that step's Jacobian-vector product.
"""
from __future__ import annotations

import numpy as np

from . import dns


class Dual:
    """value + eps * tangent, eps^2 = 0."""

    __slots__ = ("v", "t")
    __array_ufunc__ = None   # ndarray (op) Dual defers to the Dual's reflected operator

    def __init__(self, v, t):
        self.v = v
        self.t = t

    def __add__(self, o):
        if isinstance(o, Dual):
            return Dual(self.v + o.v, self.t + o.t)
        return Dual(self.v + o, self.t)

    __radd__ = __add__

    def __sub__(self, o):
        if isinstance(o, Dual):
            return Dual(self.v - o.v, self.t - o.t)
        return Dual(self.v - o, self.t)

    def __rsub__(self, o):
        return Dual(o - self.v, -self.t)

    def __neg__(self):
        return Dual(-self.v, -self.t)

    def __mul__(self, o):
        if isinstance(o, Dual):
            return Dual(self.v * o.v, self.v * o.t + self.t * o.v)
        return Dual(self.v * o, self.t * o)

    __rmul__ = __mul__

    def __getitem__(self, k):
        return Dual(self.v[k], self.t[k])


def linear(f, x: Dual) -> Dual:
    """A linear map applied to both parts."""
    return Dual(f(x.v), f(x.t))


def cross(a: Dual, b: Dual) -> Dual:
    """a x b over the component axis, by the product rule."""
    return Dual(np.cross(a.v, b.v, axis=0), np.cross(a.t, b.v, axis=0) + np.cross(a.v, b.t, axis=0))


def weighted_inner(d: dns.Domain, f: Dual, w=None) -> Dual:
    """dns.weighted_inner (<f, w f>) and its derivative 2 <f, w f'>."""
    a = d.weights * np.real(np.conj(f.v) * f.t)
    if w is not None:
        a = a * w
    return Dual(dns.weighted_inner(d, f.v, w), 2.0 * float(np.sum(a)) / d.norm)


def constant_power_amplitude(d: dns.Domain, cond: dns.Conditions, uh: Dual) -> Dual:
    """A = eps/(2 k_band) (src/Forcing.jl:1091-1117) with dA = -eps dk_band / (2 k_band^2)."""
    rate, lo, hi = cond.forcing
    band = d.shell_band(lo, hi)
    band_obs = 0.5 * weighted_inner(d, uh, band)
    whole_obs = 0.5 * dns.weighted_inner(d, uh.v)
    if dns.band_is_empty(band_obs.v, whole_obs, cond.band_eps):
        return Dual(0.0, 0.0)
    return Dual(rate / (2 * band_obs.v), -rate * band_obs.t / (2 * band_obs.v ** 2))


def velocity_rhs(d: dns.Domain, cond: dns.Conditions, uh: Dual, u: Dual, w: Dual) -> Dual:
    """dns.velocity_rhs on duals (src/Generator.jl:484-495,528-534)."""
    f = linear(lambda a: dns.leray(d, dns.rfft3(a) * d.mask), cross(u, w))
    f.v[:, 0, 0, 0] = 0.0
    f.t[:, 0, 0, 0] = 0.0
    if cond.forcing is not None:
        amp = constant_power_amplitude(d, cond, uh)
        band = d.shell_band(cond.forcing[1], cond.forcing[2])
        f = f + amp * (band * uh)
    return f


def scalar_rhs(d: dns.Domain, cond: dns.Conditions, uh: Dual, u: Dual, phi: Dual) -> Dual:
    """dns.scalar_rhs on duals (src/Generator.jl:648-661,708-730 plus the mean-gradient source)."""
    kx, ky, kz = d.kvec

    def div(flux):
        f = dns.rfft3(flux)
        return -(d.mask * (1j * (kx * f[0] + ky * f[1] + kz * f[2])))

    nl = linear(div, u * Dual(phi.v[None], phi.t[None]))
    if cond.mean_gradient != 0.0:
        nl = nl - cond.mean_gradient * uh[2]
    return nl


def step(d: dns.Domain, cond: dns.Conditions, scheme: str, uh: Dual, ph: Dual | None, dt: float):
    """One dns.step on duals: (u_hat, phi_hat) after the step, each with its tangent."""
    a, b, dd = dns.scheme_coefficients(scheme)
    S = len(a)
    hu = hp = None
    for i in range(S):
        u = linear(lambda f: dns.irfft3(f, d.n), uh)
        w = linear(lambda f: dns.irfft3(dns.curl(d, f), d.n), uh)
        phi = linear(lambda f: dns.irfft3(f, d.n), ph) if ph is not None else None
        nu_hat = velocity_rhs(d, cond, uh, u, w)
        np_hat = scalar_rhs(d, cond, uh, u, phi) if ph is not None else None
        hu = dt * nu_hat if i == 0 else a[i] * hu + dt * nu_hat
        sub = dd[i] * dt
        eu = np.exp(-cond.viscosity * sub * d.k2)
        uh = eu * (uh + b[i] * hu)
        if i < S - 1:
            hu = eu * hu
        if ph is not None:
            hp = dt * np_hat if i == 0 else a[i] * hp + dt * np_hat
            ep = np.exp(-cond.scalar_diffusivity * sub * d.k2)
            ph = ep * (ph + b[i] * hp)
            if i < S - 1:
                hp = ep * hp
    uh = linear(lambda f: dns.reality_projection(d, dns.leray(d, f)), uh)
    if ph is not None:
        ph = linear(lambda f: dns.reality_projection(d, f), ph)
    return uh, ph


def objective(d: dns.Domain, uh: Dual, ph: Dual | None) -> Dual:
    """J = k + <phi^2> (the adjoint's objective) with its derivative."""
    j = 0.5 * weighted_inner(d, uh)
    if ph is not None:
        j = j + weighted_inner(d, ph)
    return j


def directional_derivative(d: dns.Domain, cond: dns.Conditions, scheme: str, uh0: np.ndarray,
                           ph0: np.ndarray | None, dt: float, steps: int, du: np.ndarray,
                           dp: np.ndarray | None = None):
    """(J, dJ/d eps) of J(T) after `steps` steps from (u_hat0 + eps du, phi_hat0 + eps dp)."""
    uh = Dual(uh0.copy(), du.copy())
    ph = None
    if ph0 is not None:
        ph = Dual(ph0.copy(), dp.copy() if dp is not None else np.zeros_like(ph0))
    for _ in range(steps):
        uh, ph = step(d, cond, scheme, uh, ph, dt)
    j = objective(d, uh, ph)
    return j.v, j.t
