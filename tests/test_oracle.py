"""Pins of the FP64 oracle against closed forms, invariants and brute force.

The reference stores no golden arrays (every expected value in its test/ is computed in-test);
these are its own test oracles re-expressed: observables on single-shell fields
(test/observables_analytic.jl:1-87), exact diffusion (test/scalar.jl:42-60), Cauchy ratios
(test/integrators.jl:49-78,165-188), the 2N zero-padded dealiasing oracle
(test/dealiasing.jl:69-114, test/regressions/r6-inclusive-ball-exact.jl), the exact mean mode
(r5-free-modules-01-mean-mode.jl), the reality projection (r7-reality-projection-forced-run.jl),
<u.f> = eps for ConstantPower (test/forcing.jl:92-104), plus brute-force DFTs and the
Carpenter-Kennedy stability polynomial (src/Propagator.jl:287-292).
"""
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import dns

TWO_PI = 2 * math.pi


def _grid(n):
    x = np.arange(n) * TWO_PI / n
    z, y, xx = np.meshgrid(x, x, x, indexing="ij")
    return xx, y, z


# --- transforms and frame ------------------------------------------------------------------

def test_transform_convention_matches_brute_force_dft():
    n = 4
    rng = np.random.default_rng(1)
    a = rng.standard_normal((n, n, n))
    got = dns.rfft3(a)
    idx = np.arange(n)
    want = np.zeros((n, n, n // 2 + 1), dtype=complex)
    for kz in range(n):
        for ky in range(n):
            for kx in range(n // 2 + 1):
                ph = np.exp(-2j * np.pi * (kx * idx[None, None, :] + ky * idx[None, :, None] + kz * idx[:, None, None]) / n)
                want[kz, ky, kx] = np.sum(a * ph)
    assert np.max(np.abs(got - want)) < 1e-12
    assert np.max(np.abs(dns.irfft3(got, n) - a)) < 1e-14          # the 1/N^3 sits in the inverse


@pytest.mark.parametrize("n", [12, 16, 48, 64])
def test_mask_is_the_closed_two_thirds_ball_in_exact_arithmetic(n):
    d = dns.Domain(n)
    cut = Fraction(2, 3) * Fraction(n, 2)                        # (2/3)(N/2) in units of 2 pi/L
    for (iz, iy, ix), keep in np.ndenumerate(d.mask):
        m2 = int(d.mx[ix]) ** 2 + int(d.my[iy]) ** 2 + int(d.mz[iz]) ** 2
        assert keep == (Fraction(m2) <= cut * cut)
    if n % 3 == 0:                                               # the boundary shell is kept
        assert d.mask[0, 0, n // 3]
    assert d.my[n // 2] == n // 2 and d.my[n // 2 + 1] == -(n // 2 - 1)


# --- observables (test/observables_analytic.jl) ----------------------------------------------

A, B, C, D = 0.7, -0.4, 1.1, 0.9
NU, KAP = 0.03, 0.05


def _analytic(n=16):
    x, y, z = _grid(n)
    u = np.stack([A * np.sin(y), B * np.sin(z), C * np.sin(x)])
    phi = D * np.sin(x)
    return dns.Domain(n), dns.rfft3(u), dns.rfft3(phi), u


def test_observables_closed_forms():
    d, uh, ph, u = _analytic()
    sq = A * A + B * B + C * C
    r = dns.reductions(d, dns.Conditions(viscosity=NU, schmidt=NU / KAP), uh, ph)
    assert r["kinetic_energy"] == pytest.approx(sq / 4, rel=1e-13)
    assert r["dissipation"] == pytest.approx(NU * sq / 2, rel=1e-13)
    assert r["scalar_variance"] == pytest.approx(D * D / 2, rel=1e-13)
    assert r["scalar_dissipation"] == pytest.approx(KAP * D * D, rel=1e-13)
    dep = dns.deposit(d, uh, ph)
    assert dep.max_speed == pytest.approx(math.sqrt(sq), rel=1e-13)
    kap, e = dns.energy_spectrum(d, uh)
    assert e[1] * d.k0 == pytest.approx(sq / 4, rel=1e-13)       # all energy in shell 1
    assert np.sum(e) * d.k0 == pytest.approx(sq / 4, rel=1e-13)  # Parseval over the shells
    # the vorticity of the deposit is curl u = -(B cos z, C cos x, A cos y)
    x, y, z = _grid(16)
    w = -np.stack([B * np.cos(z), C * np.cos(x), A * np.cos(y)])
    assert np.max(np.abs(dep.w - w)) < 1e-13


# --- the nonlinear term ----------------------------------------------------------------------

def _band_limited(n, seed, shell_frac=1.0):
    """A random solenoidal, Hermitian field filling the 2/3 ball (boundary shell included)."""
    d = dns.Domain(n)
    rng = np.random.default_rng(seed)
    f = (rng.standard_normal((3,) + d.m2.shape) + 1j * rng.standard_normal((3,) + d.m2.shape))
    f = f * d.mask * (9 * d.m2 <= (shell_frac * n) ** 2)
    f[:, 0, 0, 0] = 0
    f = dns.leray(d, f)
    return d, dns.reality_projection(d, f)


@pytest.mark.parametrize("n", [12, 16])
def test_dealiased_rhs_equals_the_zero_padded_2n_oracle(n):
    d, uh = _band_limited(n, 3)
    dep = dns.deposit(d, uh, None)
    cond = dns.Conditions(viscosity=1.0)
    got = dns.velocity_rhs(d, cond, uh, dep.u, dep.w)
    # embed in a 2N grid (no aliasing there), form the raw product, project, restrict
    n2 = 2 * n
    d2 = dns.Domain(n2)
    big = np.zeros((3, n2, n2, n + 1), dtype=complex)
    idx = np.array([i if i <= n // 2 else i + n for i in range(n)])
    big[:, idx[:, None], idx[None, :], : n // 2 + 1] = 8 * uh
    dep2 = dns.deposit(d2, big, None)
    raw = dns.leray(d2, dns.rfft3(np.cross(dep2.u, dep2.w, axis=0)))
    want = raw[:, idx[:, None], idx[None, :], : n // 2 + 1] / 8
    want = want * d.mask
    want[:, 0, 0, 0] = 0
    scale = np.max(np.abs(want))
    assert np.max(np.abs(got - want)) / scale < 1e-12
    # the comparison can see aliasing: a field reaching past the ball aliases into it
    rng = np.random.default_rng(4)
    wide = rng.standard_normal((3, n, n, n)) * 0.1
    wh = dns.leray(d, dns.rfft3(wide))
    wh[:, 0, 0, 0] = 0
    dw = dns.deposit(d, wh, None)
    gotw = dns.velocity_rhs(d, cond, wh, dw.u, dw.w)
    bigw = np.zeros((3, n2, n2, n + 1), dtype=complex)
    bigw[:, idx[:, None], idx[None, :], : n // 2 + 1] = 8 * wh
    bigw[:, :, :, n // 2] = 0
    bigw[:, n // 2 + n, :, :] = 0
    bigw[:, :, n // 2 + n, :] = 0
    dw2 = dns.deposit(d2, bigw, None)
    truew = dns.leray(d2, dns.rfft3(np.cross(dw2.u, dw2.w, axis=0)))[:, idx[:, None], idx[None, :], : n // 2 + 1] / 8
    truew = truew * d.mask
    truew[:, 0, 0, 0] = 0
    assert np.max(np.abs(gotw - truew)) / np.max(np.abs(truew)) > 1e-3


def test_rhs_mean_mode_exactly_zero_and_solenoidal():
    d, uh = _band_limited(16, 5)
    dep = dns.deposit(d, uh, None)
    nl = dns.velocity_rhs(d, dns.Conditions(viscosity=1.0), uh, dep.u, dep.w)
    assert np.all(nl[:, 0, 0, 0] == 0)
    assert dns.divergence_residual(d, nl) < 1e-14
    # the scalar flux divergence has no mean by construction
    phi = np.random.default_rng(2).standard_normal((16, 16, 16))
    ns = dns.scalar_rhs(d, dns.Conditions(viscosity=1.0), uh, dep.u, phi)
    assert ns[0, 0, 0] == 0


def test_constant_power_injects_exactly_eps():
    d, uh = _band_limited(16, 6)
    eps = 0.1
    cond = dns.Conditions(viscosity=0.01, forcing=(eps, 1, 2))
    amp = dns.constant_power_amplitude(d, cond, uh)
    f = amp * d.shell_band(1, 2) * uh
    power = math.fsum(np.ravel((np.conj(uh) * f).real * d.weights)) / d.norm     # <u, f>
    assert power == pytest.approx(eps, rel=1e-12)
    # an empty band is clamped to zero (src/Forcing.jl:110-111)
    empty = uh * (~d.shell_band(1, 2))
    assert dns.constant_power_amplitude(d, cond, empty) == 0.0


def test_shell_band_edges():
    d = dns.Domain(16)
    band = d.shell_band(1, 2)
    m = np.sqrt(d.m2)
    assert np.all(band == ((m >= 0.5) & (m < 2.5)))
    assert np.all(d.shell_band(1, math.inf) == (d.m2 > 0))


# --- the scheme --------------------------------------------------------------------------------

@pytest.mark.parametrize("scheme,order", [("rk3", 3), ("rk4", 4)])
def test_stability_polynomial(scheme, order):
    """The 2N loop applied to q' = z q (no integrating factor) has R(z) = sum_k<=p z^k/k! (+ z^5/200)."""
    a, b, _ = dns.scheme_coefficients(scheme)
    for z in (0.3 + 0.1j, -1.2 + 0.5j, 1.5j):
        q, h = 1.0 + 0j, 0j
        for i in range(len(a)):
            h = (0 if i == 0 else a[i] * h) + z * q
            q = q + b[i] * h
        want = sum(z ** k / math.factorial(k) for k in range(order + 1))
        if scheme == "rk4":
            want += z ** 5 / 200
        assert abs(q - want) < 1e-14


def test_rk4_abscissae_match_the_b_partial_sums():
    """d are differences of the nodes c, with c2 = b1 (Carpenter-Kennedy)."""
    a, b, d = dns.scheme_coefficients("rk4")
    c = np.cumsum((0.0,) + d)
    assert c[1] == b[0]
    assert c[-1] == pytest.approx(1.0, abs=1e-16)
    assert c[2] == pytest.approx(0.3704, abs=1e-4) and c[4] == pytest.approx(0.9583, abs=1e-4)


def test_exact_diffusion_at_zero_velocity():
    n, sc, nu = 16, 0.7, 0.02
    d = dns.Domain(n)
    _, f = _band_limited(n, 9)
    ph = f[0]
    uh = np.zeros_like(f)
    cond = dns.Conditions(viscosity=nu, schmidt=sc, mean_gradient=1.0)
    dt, steps = 0.02, 25
    u, p, dep = uh, ph, None
    for _ in range(steps):
        u, p, dep = dns.step(d, cond, "rk3", u, p, dt, dep)
    want = ph * np.exp(-(nu / sc) * d.k2 * dt * steps)
    assert np.linalg.norm(p - want) / np.linalg.norm(want) < 1e-10


def _integrate(scheme, n, steps, T, uh, ph, cond):
    d = dns.Domain(n)
    u, p, dep = uh, ph, None
    for _ in range(steps):
        u, p, dep = dns.step(d, cond, scheme, u, p, T / steps, dep)
    return u, p


@pytest.mark.parametrize("scheme,lo,hi", [("rk3", 5.0, 12.0), ("rk4", 10.0, 24.0), ("ssprk3", 5.0, 12.0),
                                           ("heun", 3.0, 6.0)])
def test_temporal_self_convergence(scheme, lo, hi):
    n = 16
    d, uh = _band_limited(n, 1, shell_frac=0.5)
    uh = uh * math.sqrt(0.5 / dns.reductions(d, dns.Conditions(viscosity=1), uh, None)["kinetic_energy"])
    _, ph = _band_limited(n, 2, shell_frac=0.5)
    ph = ph[0]
    cond = dns.Conditions(viscosity=0.05, mean_gradient=0.5, forcing=(0.1, 1, 2))
    T = 0.4
    runs = [_integrate(scheme, n, k, T, uh, ph, cond) for k in (8, 16, 32)]
    for f in (0, 1):
        ratio = np.linalg.norm(runs[0][f] - runs[1][f]) / np.linalg.norm(runs[1][f] - runs[2][f])
        assert lo < ratio < hi, (f, ratio)


def test_mean_mode_and_reality_hold_over_steps():
    n = 16
    d, uh = _band_limited(n, 4, shell_frac=0.6)
    ph = np.zeros((n, n, n // 2 + 1), dtype=complex)
    cond = dns.Conditions(viscosity=0.02, forcing=(0.1, 1, 2), mean_gradient=1.0)
    u, p, dep = uh, ph, None
    for _ in range(5):
        u, p, dep = dns.step(d, cond, "rk4", u, p, 0.01, dep)
    assert np.all(u[:, 0, 0, 0] == 0)
    assert np.array_equal(dns.reality_projection(d, u), u)
    assert dns.divergence_residual(d, u) < 1e-15


def test_reality_projection_removes_a_planted_antihermitian_pair():
    d, uh = _band_limited(16, 8)
    planted = uh.copy()
    planted[0, 2, 1, 0] += 1j * 0.3          # mode (0, +1, +2)
    planted[0, 14, 15, 0] += 1j * 0.3        # its partner (0, -1, -2): anti-Hermitian pair
    fixed = dns.reality_projection(d, planted)
    assert np.max(np.abs(fixed - uh)) < 1e-15
    assert np.max(np.abs(dns.irfft3(planted, 16) - dns.irfft3(uh, 16))) < 1e-15   # invisible to c2r


def test_mean_gradient_source_is_linear_in_g():
    n = 16
    d, uh = _band_limited(n, 11, shell_frac=0.6)
    ph = np.zeros((n, n, n // 2 + 1), dtype=complex)
    outs = []
    for G in (0.0, 1.0, 2.5):
        cond = dns.Conditions(viscosity=0.02, mean_gradient=G)
        u, p, dep = uh, ph, None
        for _ in range(3):
            u, p, dep = dns.step(d, cond, "rk4", u, p, 0.01, dep)
        outs.append(p)
    assert np.all(outs[0] == 0)                                         # phi stays zero at G = 0
    assert np.linalg.norm(outs[2] - 2.5 * outs[1]) / np.linalg.norm(outs[2]) < 1e-13


def test_scalar_variance_budget():
    """d<phi^2>/dt = -2 G <u_z phi> - eps_phi, checked at O(dt^4) across one RK4 step."""
    n = 16
    d, uh = _band_limited(n, 12, shell_frac=0.6)
    _, pf = _band_limited(n, 13, shell_frac=0.6)
    ph = pf[0]
    G = 0.7
    cond = dns.Conditions(viscosity=0.03, mean_gradient=G)

    def rate(u, p):
        dep = dns.deposit(d, u, p)
        prod = -2 * G * float(np.mean(dep.u[2] * dep.phi))
        return prod - dns.reductions(d, cond, u, p)["scalar_dissipation"]

    dt = 1e-3
    u1, p1, _ = dns.step(d, cond, "rk4", uh, ph, dt)
    v0 = dns.reductions(d, cond, uh, ph)["scalar_variance"]
    v1 = dns.reductions(d, cond, u1, p1)["scalar_variance"]
    mid = 0.5 * (rate(uh, ph) + rate(u1, p1))
    assert (v1 - v0) / dt == pytest.approx(mid, rel=1e-5)


def test_linear_forcing_grows_band_modes_at_a_minus_nu_k2_exactly():
    """Linear forcing folded into the integrating factor (src/Forcing.jl:600-642): the shear waves
    u_x = a sin y (shell 1, forced) + b sin 3z (shell 3, not forced) carry no nonlinear term, so each
    evolves exactly, at exp((A - nu) t) and exp(-9 nu t)."""
    n, nu, A, dt, steps = 16, 0.05, 0.3, 1e-2, 25
    d = dns.Domain(n)
    x = np.arange(n) * 2 * math.pi / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    u = np.zeros((3, n, n, n))
    u[0] = 0.8 * np.sin(Y) + 0.5 * np.sin(3 * Z)
    uh = dns.rfft3(u) * d.mask
    cond = dns.Conditions(viscosity=nu, linear=(A, 1, 2))
    band = d.shell_band(1, 2)
    expected = uh * np.where(band, np.exp((A - nu * d.k2) * steps * dt), np.exp(-nu * d.k2 * steps * dt))
    for scheme in ("rk3", "rk4", "ssprk3", "heun"):
        v = uh
        for _ in range(steps):
            v, _, _ = dns.step(d, cond, scheme, v, None, dt)
        assert np.max(np.abs(v - expected)) <= 1e-13 * np.max(np.abs(expected))


# --- IFSSPRK3 ------------------------------------------------------------------------------------

def _lawson_ssprk3(d, cond, uh, ph, dt):
    """IFSSPRK3 written independently of the oracle's Shu-Osher arrangement: the classical SSPRK3
    Butcher tableau (c = 0, 1, 1/2; b = 1/6, 1/6, 2/3) applied to the transformed variable
    v = E(-tau) q, whose equation is v' = E(-tau) N(E(tau) v); then q = E(dt) v, the state Leray
    projection and the reality projection."""
    kap = cond.scalar_diffusivity

    def g(tau, v, w):
        q = dns.linear_factor(d, cond, tau) * v
        p = np.exp(-kap * tau * d.k2) * w
        dep = dns.deposit(d, q, p)
        nu = dns.velocity_rhs(d, cond, q, dep.u, dep.w)
        npp = dns.scalar_rhs(d, cond, q, dep.u, dep.phi)
        return dns.linear_factor(d, cond, -tau) * nu, np.exp(kap * tau * d.k2) * npp

    k1 = g(0.0, uh, ph)
    k2 = g(dt, uh + dt * k1[0], ph + dt * k1[1])
    k3 = g(dt / 2, uh + dt / 4 * (k1[0] + k2[0]), ph + dt / 4 * (k1[1] + k2[1]))
    v = uh + dt / 6 * (k1[0] + k2[0] + 4 * k3[0])
    w = ph + dt / 6 * (k1[1] + k2[1] + 4 * k3[1])
    q = dns.linear_factor(d, cond, dt) * v
    p = np.exp(-kap * dt * d.k2) * w
    return dns.reality_projection(d, dns.leray(d, q)), dns.reality_projection(d, p)


@pytest.mark.parametrize("forcing", ["none", "constant_power", "linear"])
def test_ssprk3_equals_the_lawson_butcher_form(forcing):
    """The Shu-Osher stages of src/Propagator.jl:599-691 against the Butcher form of the same
    scheme on the transformed variable: a swapped weight, a wrong factor sign or a wrong stage
    time fails this at 1e-13."""
    n = 16
    d, uh = _band_limited(n, 4, shell_frac=0.6)
    _, ph = _band_limited(n, 5, shell_frac=0.6)
    ph = ph[0]
    kw = {"constant_power": {"forcing": (0.2, 1, 2)}, "linear": {"linear": (0.4, 1, 2)}}.get(forcing, {})
    cond = dns.Conditions(viscosity=0.03, schmidt=0.8, mean_gradient=0.7, **kw)
    dt = 0.05
    got_u, got_p, _ = dns.step(d, cond, "ssprk3", uh, ph, dt)
    want_u, want_p = _lawson_ssprk3(d, cond, uh, ph, dt)
    assert np.max(np.abs(got_u - want_u)) <= 1e-13 * np.max(np.abs(want_u))
    assert np.max(np.abs(got_p - want_p)) <= 1e-13 * np.max(np.abs(want_p))


def test_ssprk3_and_rk3_are_distinct_tableaux_of_one_order():
    """test/integrators.jl:71-76: the two third-order schemes agree as dt -> 0 while their
    single-dt states differ."""
    n = 16
    d, uh = _band_limited(n, 1, shell_frac=0.5)
    uh = uh * math.sqrt(0.5 / dns.reductions(d, dns.Conditions(viscosity=1), uh, None)["kinetic_energy"])
    cond = dns.Conditions(viscosity=0.05)
    T = 0.4
    u1, _ = _integrate("rk3", n, 8, T, uh, None, cond)
    u4, _ = _integrate("rk3", n, 32, T, uh, None, cond)
    w1, _ = _integrate("ssprk3", n, 8, T, uh, None, cond)
    w4, _ = _integrate("ssprk3", n, 32, T, uh, None, cond)
    assert np.linalg.norm(u1 - w1) > 0
    assert np.linalg.norm(u4 - w4) < np.linalg.norm(u1 - w1)


def test_ssprk3_weights_round_to_the_element_type():
    """T(3//4), T(1//4), T(1//3), T(2//3) (src/Propagator.jl:643-667): exact in FP64 for the
    quarters, correctly rounded for the thirds."""
    assert dns._round(Fraction(1, 3), np.float32) == float(np.float32(1 / 3))
    assert dns._round(Fraction(2, 3), np.float32) == float(np.float32(2 / 3))
    assert dns._round(Fraction(3, 4), np.float32) == 0.75


# --- VarianceController and VarianceRescaling ------------------------------------------------------

def _physical_mean_square(d, fh):
    """<|f|^2> over the grid points of the physical field whose spectrum is fh (no Parseval)."""
    f = dns.irfft3(fh, d.n)
    return float(np.mean(np.sum(f * f, axis=0)))


def test_controller_amplitude_is_bassenne_eq20_in_physical_space():
    """test/forcing.jl:322-347: A = (eps - G (k - k_inf)/t_l) / (2 k_band), the restoring term on the
    WHOLE field's k and the normalisation on the band's, each measured here in physical space
    (eps = nu <|grad u|^2> from the physical gradient)."""
    n, nu = 16, 0.03
    d, uh = _band_limited(n, 11, shell_frac=0.7)
    uh = uh * math.sqrt(0.4 / dns.reductions(d, dns.Conditions(viscosity=1), uh, None)["kinetic_energy"])
    target, gain, tl = 1.5, 1.0, 0.8
    cond = dns.Conditions(viscosity=nu, controller=(target, gain, tl, 1, 2))
    band = d.shell_band(1, 2)
    k = 0.5 * _physical_mean_square(d, uh)
    kb = 0.5 * _physical_mean_square(d, uh * band)
    kv = d.kvec
    grad = sum(_physical_mean_square(d, 1j * kv[i] * uh) for i in range(3))
    eps = nu * grad
    want = (eps - gain * (k - target) / tl) / (2 * kb)
    got = dns.controller_amplitude(d, cond, uh)
    assert got == pytest.approx(want, rel=1e-12)
    wrong = (eps - gain * (kb - target) / tl) / (2 * kb)
    assert abs(got - wrong) > 1e-3 * abs(wrong)


def _controlled_run(scheme, gain, dt, steps, k0=0.5, target=1.5, tl=0.8):
    n = 16
    d, uh = _band_limited(n, 3, shell_frac=0.5)
    uh = uh * math.sqrt(k0 / dns.reductions(d, dns.Conditions(viscosity=1), uh, None)["kinetic_energy"])
    cond = dns.Conditions(viscosity=0.05, controller=(target, gain, tl, 1, 2))
    u, dep = uh, None
    ks = []
    for _ in range(steps):
        u, _, dep = dns.step(d, cond, scheme, u, None, dt, dep)
        ks.append(dns.reductions(d, cond, u, None)["kinetic_energy"])
    return np.array(ks)


def test_controller_relaxes_k_exponentially():
    """The dealiased nonlinear term carries no energy at N = 16 (no product mode aliases into the
    ball), so dk/dt = -eps + 2 A k_band = -G (k - k_inf)/t_l: k(t) = k_inf + (k0 - k_inf) exp(-G t/t_l)."""
    dt, steps, gain, tl = 5e-3, 100, 2.0, 0.8
    ks = _controlled_run("rk4", gain, dt, steps, tl=tl)
    t = dt * np.arange(1, steps + 1)
    want = 1.5 + (0.5 - 1.5) * np.exp(-gain * t / tl)
    assert np.max(np.abs(ks - want)) < 1e-9


def test_controller_marginal_law_conserves_k():
    """test/forcing.jl:389-405: gain = 0 reinjects eps exactly, so k stays at its initial 0.5, far
    from the target 1.5."""
    ks = _controlled_run("rk3", 0.0, 5e-3, 100)
    assert np.max(np.abs(ks - 0.5)) < 1e-8   # the scheme's uncorrected O(dt^3) bias: 4.3e-9 here


def test_controller_keeps_third_order():
    """src/Forcing.jl:521-526: the controller enters as a per-stage source, so IFLowStorageRK3 keeps
    its order (ratio about 8 between successive halvings)."""
    n = 16
    d, uh = _band_limited(n, 1, shell_frac=0.5)
    uh = uh * math.sqrt(0.5 / dns.reductions(d, dns.Conditions(viscosity=1), uh, None)["kinetic_energy"])
    cond = dns.Conditions(viscosity=0.05, controller=(0.8, 1.0, 0.5, 1, 2))
    runs = [_integrate("rk3", n, k, 0.4, uh, None, cond)[0] for k in (8, 16, 32)]
    ratio = np.linalg.norm(runs[0] - runs[1]) / np.linalg.norm(runs[1] - runs[2])
    assert 5.0 < ratio < 12.0, ratio


@pytest.mark.parametrize("n", [16, 24])
@pytest.mark.parametrize("over", ["band", "total"])
def test_rescaling_every_mode_scales_the_field_to_the_target(n, over):
    """test/forcing.jl:440-456: with highest_shell = Inf the band is every mode but the mean and both
    modes scale the whole field by sqrt(target / k)."""
    d, uh = _band_limited(n, n, shell_frac=0.8)
    cond = dns.Conditions(viscosity=0.01, rescaling=(1.3, over, 1, math.inf))
    k = dns.reductions(d, cond, uh, None)["kinetic_energy"]
    got = dns.variance_rescaling(d, cond, uh)
    assert dns.reductions(d, cond, got, None)["kinetic_energy"] == pytest.approx(1.3, rel=1e-12)
    want = uh * math.sqrt(1.3 / k)
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))


def test_rescaling_band_and_total_hit_their_targets():
    n = 16
    d, uh = _band_limited(n, 8, shell_frac=0.8)
    band = d.shell_band(1, 2)
    k = dns.reductions(d, dns.Conditions(viscosity=1), uh, None)["kinetic_energy"]
    kb = 0.5 * dns.weighted_inner(d, uh, band)
    # over band: the band's energy becomes the target, the rest is untouched
    got = dns.variance_rescaling(d, dns.Conditions(viscosity=1, rescaling=(0.2, "band", 1, 2)), uh)
    assert 0.5 * dns.weighted_inner(d, got, band) == pytest.approx(0.2, rel=1e-12)
    assert np.array_equal(got * ~band, uh * ~band)
    # over total: the whole field's energy becomes the target
    tgt = k + 0.1
    got = dns.variance_rescaling(d, dns.Conditions(viscosity=1, rescaling=(tgt, "total", 1, 2)), uh)
    assert dns.reductions(d, dns.Conditions(viscosity=1), got, None)["kinetic_energy"] == pytest.approx(tgt, rel=1e-12)
    # over total with the outside alone above the target: the band is zeroed
    got = dns.variance_rescaling(d, dns.Conditions(viscosity=1, rescaling=(0.5 * (k - kb), "total", 1, 2)), uh)
    assert np.max(np.abs(got * band)) == 0.0
    # an empty band is left alone
    v = uh * ~band
    got = dns.variance_rescaling(d, dns.Conditions(viscosity=1, rescaling=(0.3, "band", 1, 2)), v)
    assert np.array_equal(got, v)


@pytest.mark.parametrize("scheme", ["rk3", "ssprk3"])
def test_rescaling_after_each_step_holds_the_target(scheme):
    """finalize! runs after the stages and before the state projections, which keep the band's
    energy: the held value is exact after every step."""
    n = 16
    d, uh = _band_limited(n, 2, shell_frac=0.6)
    cond = dns.Conditions(viscosity=0.02, rescaling=(0.7, "total", 1, 2))
    u, dep = uh, None
    for _ in range(5):
        u, _, dep = dns.step(d, cond, scheme, u, None, 1e-2, dep)
        assert dns.reductions(d, cond, u, None)["kinetic_energy"] == pytest.approx(0.7, rel=1e-13)


def _lawson_heun(d, cond, uh, ph, dt):
    """IFHeun as the Butcher tableau (c = 0, 1; b = 1/2, 1/2) on v = E(-tau) q, independent of the
    reference's predictor/corrector arrangement."""
    kap = cond.scalar_diffusivity

    def g(tau, v, w):
        q = dns.linear_factor(d, cond, tau) * v
        p = np.exp(-kap * tau * d.k2) * w
        dep = dns.deposit(d, q, p)
        return (dns.linear_factor(d, cond, -tau) * dns.velocity_rhs(d, cond, q, dep.u, dep.w),
                np.exp(kap * tau * d.k2) * dns.scalar_rhs(d, cond, q, dep.u, dep.phi))

    k1 = g(0.0, uh, ph)
    k2 = g(dt, uh + dt * k1[0], ph + dt * k1[1])
    q = dns.linear_factor(d, cond, dt) * (uh + dt / 2 * (k1[0] + k2[0]))
    p = np.exp(-kap * dt * d.k2) * (ph + dt / 2 * (k1[1] + k2[1]))
    return dns.reality_projection(d, dns.leray(d, q)), dns.reality_projection(d, p)


@pytest.mark.parametrize("forcing", ["none", "constant_power", "linear"])
def test_heun_equals_the_lawson_butcher_form(forcing):
    n = 16
    d, uh = _band_limited(n, 4, shell_frac=0.6)
    _, ph = _band_limited(n, 5, shell_frac=0.6)
    ph = ph[0]
    kw = {"constant_power": {"forcing": (0.2, 1, 2)}, "linear": {"linear": (0.4, 1, 2)}}.get(forcing, {})
    cond = dns.Conditions(viscosity=0.03, schmidt=0.8, mean_gradient=0.7, **kw)
    got_u, got_p, _ = dns.step(d, cond, "heun", uh, ph, 0.05)
    want_u, want_p = _lawson_heun(d, cond, uh, ph, 0.05)
    assert np.max(np.abs(got_u - want_u)) <= 1e-13 * np.max(np.abs(want_u))
    assert np.max(np.abs(got_p - want_p)) <= 1e-13 * np.max(np.abs(want_p))


def test_heun_is_the_shu_osher_ssprk2():
    """The device runs IFHeun as SSPRK(2,2): u' = 1/2 E u + 1/2 (u* + dt N(u*)), u* = E (u + dt N(u));
    the rearrangement equals the reference's E u + dt/2 (E N1 + N2) to roundoff."""
    n = 16
    d, uh = _band_limited(n, 6, shell_frac=0.6)
    cond = dns.Conditions(viscosity=0.03, forcing=(0.2, 1, 2))
    dt = 0.05
    e = dns.linear_factor(d, cond, dt)
    dep = dns.deposit(d, uh, None)
    us = e * (uh + dt * dns.velocity_rhs(d, cond, uh, dep.u, dep.w))
    ds = dns.deposit(d, us, None)
    so = 0.5 * (e * uh) + 0.5 * (us + dt * dns.velocity_rhs(d, cond, us, ds.u, ds.w))
    want = dns.reality_projection(d, dns.leray(d, so))
    got, _, _ = dns.step(d, cond, "heun", uh, None, dt)
    assert np.max(np.abs(got - want)) <= 1e-14 * np.max(np.abs(want))


# --- scalar forcings ---------------------------------------------------------------------------

def _scalar_wave(n, amp=0.6):
    """phi = amp (sin y + 0.5 cos 3z): shells 1 and 3, with u = 0 nothing advects it."""
    x = np.arange(n) * TWO_PI / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    return dns.rfft3(amp * (np.sin(Y) + 0.5 * np.cos(3 * Z)))


def test_scalar_constant_power_injects_exactly_its_rate():
    """<phi, A 1_band phi> = eps_phi / 2 with the variance factor 1 (src/Forcing.jl:456, :1104-1117):
    d<phi^2>/dt gets exactly eps_phi from the source."""
    n = 16
    d = dns.Domain(n)
    _, f = _band_limited(n, 12, shell_frac=0.8)
    ph = f[0]
    cond = dns.Conditions(viscosity=0.02, scalar_forcing=("constant_power", 0.3, 1, 2))
    src = dns.scalar_amplitude(d, cond, ph) * d.shell_band(1, 2) * ph
    rate = 2 * sum(np.real(np.sum(d.weights * np.conj(ph) * src, axis=None)) for _ in [0]) / d.norm
    assert rate == pytest.approx(0.3, rel=1e-12)


@pytest.mark.parametrize("scheme", ["rk3", "rk4", "ssprk3", "heun"])
def test_scalar_linear_forcing_grows_its_band_exactly(scheme):
    """With u = 0 the scalar is linear: shell 1 grows as exp((A - kappa) t), shell 3 decays as
    exp(-9 kappa t) (the rate folded into the scalar integrating factor)."""
    n, nu, sc, A, dt, steps = 16, 0.05, 0.5, 0.4, 1e-2, 20
    d = dns.Domain(n)
    ph = _scalar_wave(n)
    kap = nu / sc
    cond = dns.Conditions(viscosity=nu, schmidt=sc, scalar_forcing=("linear", A, 1, 2))
    want = ph * np.where(d.shell_band(1, 2), np.exp((A - kap * d.k2) * steps * dt), np.exp(-kap * d.k2 * steps * dt))
    u, p, dep = np.zeros((3,) + ph.shape, dtype=complex), ph, None
    for _ in range(steps):
        u, p, dep = dns.step(d, cond, scheme, u, p, dt, dep)
    assert np.max(np.abs(p - want)) <= 1e-13 * np.max(np.abs(want))


def test_scalar_controller_relaxes_the_variance_exponentially():
    """u = 0: d<phi^2>/dt = -2 kappa <phi, |k|^2 phi> + 2 A <phi^2>_band = -G (<phi^2> - target)/t_l."""
    n, nu, sc, dt, steps, G, tl, target = 16, 0.05, 0.5, 5e-3, 60, 2.0, 0.8, 0.6
    d = dns.Domain(n)
    ph = _scalar_wave(n)
    cond = dns.Conditions(viscosity=nu, schmidt=sc, scalar_forcing=("controller", target, G, tl, 1, 2))
    v0 = dns.weighted_inner(d, ph)
    u, p, dep = np.zeros((3,) + ph.shape, dtype=complex), ph, None
    for s in range(1, steps + 1):
        u, p, dep = dns.step(d, cond, "rk4", u, p, dt, dep)
        want = target + (v0 - target) * math.exp(-G * s * dt / tl)
        assert dns.weighted_inner(d, p) == pytest.approx(want, rel=2e-8)   # IF-RK4's truncation error


@pytest.mark.parametrize("over", ["band", "total"])
def test_scalar_rescaling_holds_its_target_after_every_step(over):
    n = 16
    d, uh = _band_limited(n, 13, shell_frac=0.6)
    _, f = _band_limited(n, 14, shell_frac=0.6)
    ph = f[0] * math.sqrt(0.3 / dns.weighted_inner(d, f[0]))
    cond = dns.Conditions(viscosity=0.02, mean_gradient=0.5, scalar_forcing=("rescaling", 0.45, over, 1, 2))
    u, p, dep = uh, ph, None
    for _ in range(4):
        u, p, dep = dns.step(d, cond, "rk3", u, p, 1e-2, dep)
        held = dns.weighted_inner(d, p, d.shell_band(1, 2)) if over == "band" else dns.weighted_inner(d, p)
        assert held == pytest.approx(0.45, rel=1e-13)


# --- HouLi and PhaseShift (src/SimulationDomain.jl:76-195, src/Generator.jl:153-182,502-519,668-704) ---

def _field_in(d, seed, keep):
    """A random solenoidal, Hermitian velocity and a random scalar, nonzero only where `keep`."""
    rng = np.random.default_rng(seed)
    shape = (3,) + d.m2.shape
    f = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * keep
    f[:, 0, 0, 0] = 0
    f = dns.reality_projection(d, dns.leray(d, f))
    p = (rng.standard_normal(d.m2.shape) + 1j * rng.standard_normal(d.m2.shape)) * keep
    p[0, 0, 0] = 0
    return f, dns.reality_projection(d, p[None])[0]


def _exact_products(d, uh, ph):
    """The unaliased spectra of u x w and of -i k . (u phi) on the N grid, from a 2N grid on which
    no product of the N grid's modes aliases (every N-grid mode, Nyquist planes included, is placed
    at its own signed wavenumber)."""
    n, n2 = d.n, 2 * d.n
    d2 = dns.Domain(n2)
    idx = np.array([i if i < n // 2 else i + n for i in range(n)])   # signed m -> bin of the 2N grid
    big = np.zeros((3, n2, n2, n + 1), dtype=complex)
    big[:, idx[:, None], idx[None, :], : n // 2 + 1] = 8 * uh
    pbig = np.zeros((n2, n2, n + 1), dtype=complex)
    pbig[idx[:, None], idx[None, :], : n // 2 + 1] = 8 * ph
    dep2 = dns.deposit(d2, big, pbig)
    lamb = dns.rfft3(np.cross(dep2.u, dep2.w, axis=0))[:, idx[:, None], idx[None, :], : n // 2 + 1] / 8
    flux = dns.rfft3(dep2.u * dep2.phi[None])[:, idx[:, None], idx[None, :], : n // 2 + 1] / 8
    kx, ky, kz = d.kvec
    return lamb, -1j * (kx * flux[0] + ky * flux[1] + kz * flux[2])


@pytest.mark.parametrize("n", [12, 16])
def test_phase_shift_dealiases_the_quadratic_exactly(n):
    """PhaseShift keeps the ball |m| <= sqrt(2) N/3 and, for fields inside it, its nonlinear terms
    equal the unaliased products truncated to the ball (src/SimulationDomain.jl:140-176): the
    half-cell average cancels the odd aliases and the ball excludes the two-axis ones.  A
    single-grid evaluation with the same ball (the displaced term dropped) is visibly aliased."""
    d = dns.Domain(n, dealiasing="phase_shift")
    assert np.array_equal(d.mask, 9 * d.m2 <= 2 * n * n)
    uh, ph = _field_in(d, 11, d.mask)
    dep = dns.deposit(d, uh, ph)
    cond = dns.Conditions(viscosity=1.0)
    lamb, div = _exact_products(d, uh, ph)
    want_u = dns.leray(d, lamb * d.mask)
    want_u[:, 0, 0, 0] = 0
    got_u = dns.velocity_rhs(d, cond, uh, dep.u, dep.w)
    assert np.max(np.abs(got_u - want_u)) / np.max(np.abs(want_u)) < 1e-13
    got_p = dns.scalar_rhs(d, cond, uh, dep.u, dep.phi, ph)
    assert np.max(np.abs(got_p - div * d.mask)) / np.max(np.abs(div)) < 1e-13
    plain = dns.leray(d, dns.rfft3(np.cross(dep.u, dep.w, axis=0)) * d.mask)
    plain[:, 0, 0, 0] = 0
    assert np.max(np.abs(plain - want_u)) / np.max(np.abs(want_u)) > 1e-2


@pytest.mark.parametrize("n", [16, 24])
def test_hou_li_filters_the_unaliased_products(n):
    """HouLi keeps every mode and multiplies the nonlinear terms by exp(-s (|k|/k_Ny)^p)
    (src/SimulationDomain.jl:365-376).  For fields with |m_a| < N/4 nothing aliases, so the terms
    are the filtered exact products; the filter is 1 at the mean, e^{-s} at the axis Nyquist mode
    and eps(T) there by default."""
    order, s = 36, 2.5
    d = dns.Domain(n, dealiasing="hou_li", filter_order=order, filter_strength=s)
    assert d.mask.all()
    assert d.dealias[0, 0, 0] == 1.0 and d.dealias[0, 0, n // 2] == pytest.approx(math.exp(-s), rel=1e-15)
    q = n // 4 - 1
    r = math.sqrt(3) * q / (n / 2)              # the mode (q, q, q): |k|/k_Ny = sqrt(3) q / (N/2)
    assert d.dealias[q, q, q] == pytest.approx(math.exp(-s * r ** order), rel=1e-14)
    assert dns.Domain(n, dealiasing="hou_li").dealias[0, 0, n // 2] == pytest.approx(np.finfo(np.float64).eps, rel=1e-12)
    assert dns.Domain(n, dealiasing="hou_li", element_type=np.float32).dealias[0, 0, n // 2] == \
        pytest.approx(np.finfo(np.float32).eps, rel=1e-6)
    quarter = (np.abs(d.mx)[None, None, :] < n // 4) & (np.abs(d.my)[None, :, None] < n // 4) & \
        (np.abs(d.mz)[:, None, None] < n // 4)
    uh, ph = _field_in(d, 12, quarter)
    dep = dns.deposit(d, uh, ph)
    cond = dns.Conditions(viscosity=1.0)
    lamb, div = _exact_products(d, uh, ph)
    filt = np.exp(-s * (np.sqrt(d.m2.astype(float)) / (n / 2)) ** order)
    want_u = dns.leray(d, lamb) * filt
    want_u[:, 0, 0, 0] = 0
    got_u = dns.velocity_rhs(d, cond, uh, dep.u, dep.w)
    assert np.max(np.abs(got_u - want_u)) / np.max(np.abs(want_u)) < 1e-13
    got_p = dns.scalar_rhs(d, cond, uh, dep.u, dep.phi, ph)
    assert np.max(np.abs(got_p - div * filt)) / np.max(np.abs(div)) < 1e-13
    # the filter is felt: at |k| > 1.2 k_Ny the products the terms carry are damped away
    edge = (d.m2 > (0.6 * n) ** 2) & (np.abs(lamb).sum(axis=0) > 1e-8 * np.abs(lamb).max())
    assert edge.any() and np.all(filt[edge] < 1e-3) and np.all(np.abs(got_u[:, edge]) < 1e-3 * np.abs(lamb[:, edge]))


def test_the_wide_strategies_keep_the_step_hermitian_solenoidal_and_at_its_order():
    """A step under each strategy keeps the state real (Hermitian kx = 0 and N/2 planes) and
    solenoidal, and IF-RK4 under HouLi (default strength, Nyquist damped to eps) and PhaseShift
    keeps its fourth order (Cauchy ratio 16)."""
    n = 16
    for kind in ("hou_li", "phase_shift"):
        d = dns.Domain(n, dealiasing=kind)
        uh, ph = _field_in(d, 21, d.mask & (9 * d.m2 <= n * n))
        uh, ph = uh * 40.0, ph * 40.0
        cond = dns.Conditions(viscosity=0.02, mean_gradient=1.0)
        u1, p1, _ = dns.step(d, cond, "rk4", uh, ph, 2e-3)
        # solenoidal away from the Nyquist indices, where +N/2 and -N/2 are one mode of a real field
        # and the reality projection pairs them (the filter holds them at e^{-s} there)
        inner = (np.abs(d.mx)[None, None, :] < n // 2) & (np.abs(d.my)[None, :, None] < n // 2) & \
            (np.abs(d.mz)[:, None, None] < n // 2)
        kx, ky, kz = d.kvec
        div = np.abs(kx * u1[0] + ky * u1[1] + kz * u1[2])
        assert np.max(div[inner]) < 1e-13 * np.max(np.sqrt(d.k2) * np.abs(u1).sum(axis=0))
        for f in (u1, p1[None]):
            assert np.allclose(f, dns.reality_projection(d, f), rtol=0, atol=1e-15 * np.abs(f).max())

        def run(steps, t=0.08):
            a, b = uh, ph
            for _ in range(steps):
                a, b, _ = dns.step(d, cond, "rk4", a, b, t / steps)
            return a

        ref = run(64)
        e1, e2 = np.linalg.norm(run(4) - ref), np.linalg.norm(run(8) - ref)
        assert 13 < e1 / e2 < 19, (kind, e1 / e2)
