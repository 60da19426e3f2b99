"""The oracle's spin-up observables against what mathematics fixes (CPU), after the reference's
test/observables_analytic.jl:43-100: the closed-form field, the hand-written spectrum, oddness in
the velocity, the Dirichlet route of the denominator, and the determinant itself."""
import math

import numpy as np
import pytest

from oracle import dns
from oracle import observables as obs

A, B, C = 0.7, -1.3, 0.4


def _analytic(n):
    """u = (A sin y, B sin z, C sin x): one shell, and a strain with zero mean determinant."""
    d = dns.Domain(n)
    x = np.arange(n) * 2 * math.pi / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    u = np.stack([A * np.sin(Y), B * np.sin(Z), C * np.sin(X)])
    return d, dns.rfft3(u)


def _random(n, seed=5):
    d = dns.Domain(n)
    rng = np.random.default_rng(seed)
    k = np.sqrt(d.k2)
    env = np.where(d.mask, k ** 2 * np.exp(-(k / 3.0) ** 2), 0.0)
    uh = dns.rfft3(rng.standard_normal((3, n, n, n))) * env
    uh[:, 0, 0, 0] = 0
    return d, dns.reality_projection(d, dns.leray(d, uh))


def test_closed_form_field():
    d, uh = _analytic(16)
    kappa, E = dns.energy_spectrum(d, uh)
    assert E[1] == pytest.approx((A * A + B * B + C * C) / 4, rel=1e-13)
    assert obs.integral_length(kappa, E) == pytest.approx(3 * math.pi / 4, rel=1e-12)
    assert abs(obs.derivative_skewness(d, uh)) < 1e-12


def test_integral_length_of_a_hand_written_spectrum():
    assert obs.integral_length([0.0, 1.0, 2.0], [5.0, 2.0, 1.0]) == pytest.approx(
        3 * math.pi / 4 * (2.0 / 1.0 + 1.0 / 2.0) / (2.0 + 1.0), rel=1e-14)
    with pytest.raises(ValueError, match="no energy"):
        obs.integral_length([0.0, 1.0], [5.0, 0.0])


def test_skewness_is_odd_in_the_velocity_and_nonzero_for_a_random_field():
    d, uh = _random(16)
    s = obs.derivative_skewness(d, uh)
    assert s != 0
    assert obs.derivative_skewness(d, -uh) == pytest.approx(-s, rel=1e-12)


def test_the_dirichlet_route_equals_the_direct_strain_norm():
    """<S:S> from u_hat equals the grid mean of S_ij S_ij of the physical strain (the cross term
    vanishes for a solenoidal field)."""
    d, uh = _random(16)
    s6 = obs.strain_physical(d, uh)
    direct = float(np.mean(s6[0] ** 2 + s6[1] ** 2 + s6[2] ** 2 + 2 * (s6[3] ** 2 + s6[4] ** 2 + s6[5] ** 2)))
    assert obs.strain_dirichlet(d, uh) == pytest.approx(direct, rel=1e-12)


def test_the_determinant_against_linalg_and_cayley_hamilton():
    """det from the closed form equals numpy.linalg.det pointwise, and, the strain being traceless,
    equals tr(S^3) / 3."""
    d, uh = _random(8, seed=9)
    s6 = obs.strain_physical(d, uh)
    a, b, c, dd, e, f = s6
    M = np.stack([np.stack([a, dd, e]), np.stack([dd, b, f]), np.stack([e, f, c])])   # [3, 3, z, y, x]
    Mt = np.moveaxis(M, (0, 1), (-2, -1))
    det = obs.det_symmetric(s6)
    np.testing.assert_allclose(det, np.linalg.det(Mt), rtol=1e-10, atol=1e-12 * np.abs(det).max())
    assert np.max(np.abs(a + b + c)) < 1e-12 * np.abs(s6).max()
    tr3 = np.einsum("...ij,...jk,...ki->...", Mt, Mt, Mt)
    np.testing.assert_allclose(det, tr3 / 3, rtol=1e-9, atol=1e-12 * np.abs(det).max())


def test_below_scale_of_a_shear_wave_is_the_sinc_closed_form():
    """u_x = sin(m y): W_l = sinc(m k0 l / 2) on its modes, so k(l) = k (1 - sinc^2) and
    eps(l) = eps (1 - sinc^2); at l = L the window kills every mode but the mean: k(L) = k."""
    n, m = 16, 3
    d = dns.Domain(n)
    x = np.arange(n) * 2 * math.pi / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    u = np.zeros((3, n, n, n))
    u[0] = np.sin(m * Y)
    uh = dns.rfft3(u) * d.mask
    nu = 0.01
    k = 0.5 * dns.weighted_inner(d, uh)
    eps = nu * dns.weighted_inner(d, uh, d.k2)
    sc, kl, el = obs.below_scale(d, uh, nu)
    s = np.sin(m * sc / 2) / (m * sc / 2)
    assert np.allclose(kl, k * (1 - s ** 2), rtol=0, atol=1e-14)
    assert np.allclose(el, eps * (1 - s ** 2), rtol=0, atol=1e-14)
    assert kl[-1] == pytest.approx(k, rel=1e-14) and el[-1] == pytest.approx(eps, rel=1e-14)


def test_below_scale_is_the_energy_about_the_box_mean_by_quadrature():
    """k(l) at l = 2 cells of an 8^3 field is k less half the mean square of the box mean
    <u>_l(x) = l^-3 int_{box at x} u (Pope 2000, Ex. 13.22). Here the box mean comes from direct
    quadrature of the band-limited field on a 16x finer grid (Simpson's rule per axis), not from
    the sinc transfer."""
    n = 8
    d = dns.Domain(n)
    rng = np.random.default_rng(3)
    f = (rng.standard_normal((3,) + d.m2.shape) + 1j * rng.standard_normal((3,) + d.m2.shape)) * d.mask
    f[:, 0, 0, 0] = 0
    uh = dns.reality_projection(d, dns.leray(d, f))
    _, kl, _ = obs.below_scale(d, uh, 1.0)
    fine = 16 * n                                     # the band-limited field on a finer grid
    big = np.zeros((3, fine, fine, fine // 2 + 1), dtype=complex)
    idx = np.array([i if i <= n // 2 else i + fine - n for i in range(n)])
    big[:, idx[:, None], idx[None, :], : n // 2 + 1] = uh * (fine / n) ** 3
    uf = dns.irfft3(big, fine)
    w = 2 * fine // n                                 # a side of 2 coarse cells, in fine steps

    def box_mean(a, axis):                            # Simpson's rule over [x, x + l] along one axis
        out = a + np.roll(a, -w, axis=axis)
        for j in range(1, w):
            out = out + (4 if j % 2 else 2) * np.roll(a, -j, axis=axis)
        return out / (3 * w)

    mean = box_mean(box_mean(box_mean(uf, 1), 2), 3)
    want = 0.5 * (np.mean(np.sum(uf ** 2, axis=0)) - np.mean(np.sum(mean ** 2, axis=0)))
    assert kl[1] == pytest.approx(want, rel=1e-5)          # Simpson's O(h^4) error
