"""The oracle adjoint against central finite differences of the oracle step (CPU).

These pin oracle/adjoint.py: the transposed step is held to the derivative of the forward
oracle along random directions, for both schemes, with ConstantPower forcing, the scalar and
the mean-gradient source all active, and the checkpointed gradient of J = k + <phi^2> over
several steps is held to finite differences along random solenoidal directions.
"""
import math

import numpy as np
import pytest

from oracle import adjoint as adj
from oracle import dns


def _setup(n, forcing=0.1, G=1.0, seed=3):
    d = dns.Domain(n)
    rng = np.random.default_rng(seed)
    k = np.sqrt(d.k2)
    env = np.where(d.mask, k ** 2 * np.exp(-(k / 3.0) ** 2), 0.0)
    u = dns.rfft3(rng.standard_normal((3, n, n, n))) * env
    u[:, 0, 0, 0] = 0
    u = dns.leray(d, u)
    u *= math.sqrt(0.5 / (0.5 * adj.inner(d, u, u) / d.norm))
    p = dns.rfft3(rng.standard_normal((n, n, n))) * env * 0.3
    p[0, 0, 0] = 0
    cond = dns.Conditions(viscosity=0.02, schmidt=0.7, forcing=(forcing, 1, 2) if forcing else None, mean_gradient=G)
    return d, cond, u, p


def fd4(f, eps):
    """Fourth-order central difference: truncation O(eps^4), round-off O(1e-16/eps)."""
    return (-f(2 * eps) + 8 * f(eps) - 8 * f(-eps) + f(-2 * eps)) / (12 * eps)


def _direction(d, seed, scalar=False):
    rng = np.random.default_rng(seed)
    if scalar:
        f = dns.rfft3(rng.standard_normal((d.n,) * 3)) * d.mask
        return f / math.sqrt(adj.inner(d, f, f))
    return adj.solenoidal_direction(d, seed)


@pytest.mark.parametrize("scheme", ["rk3", "rk4"])
@pytest.mark.parametrize("forcing", [None, 0.1])
def test_step_vjp_equals_the_directional_derivative(scheme, forcing):
    d, cond, u, p = _setup(16, forcing)
    dt = 2e-2
    lu, lp = _direction(d, 11), _direction(d, 12, scalar=True)       # the output's adjoint
    gu, gp = adj.step_vjp(d, cond, scheme, u, p, dt, lu, lp)
    for seed in (21, 22):
        du, dp = _direction(d, seed), _direction(d, seed + 100, scalar=True)

        def f(e):
            uo, po, _ = dns.step(d, cond, scheme, u + e * du, p + e * dp, dt)
            return adj.inner(d, lu, uo) + adj.inner(d, lp, po)

        fd = fd4(f, 0.1)   # |delta| = 1 against |u_hat| ~ N^3: still deep in the linear regime
        an = adj.inner(d, gu, du) + adj.inner(d, gp, dp)
        assert an == pytest.approx(fd, rel=1e-8, abs=1e-12)


def test_checkpointed_gradient_matches_central_differences():
    d, cond, u, p = _setup(16)
    dt, steps = 1e-2, 6
    J, gu, gp = adj.gradient(d, cond, "rk4", u, p, dt, steps, every=2)
    assert J == pytest.approx(adj.objective(d, *adj.run(d, cond, "rk4", u, p, dt, steps)), rel=1e-15)
    for seed in range(4):
        du = adj.solenoidal_direction(d, 40 + seed)
        fd = fd4(lambda e: adj.objective(d, *adj.run(d, cond, "rk4", u + e * du, p, dt, steps)), 0.1)
        assert adj.inner(d, gu, du) == pytest.approx(fd, rel=1e-8)


def test_checkpoint_spacing_does_not_change_the_gradient():
    d, cond, u, p = _setup(16, seed=5)
    ref = adj.gradient(d, cond, "rk3", u, p, 1e-2, 6, every=1)
    for every in (2, 3, 6):
        J, gu, gp = adj.gradient(d, cond, "rk3", u, p, 1e-2, 6, every=every)
        assert J == ref[0]
        assert np.array_equal(gu, ref[1]) and np.array_equal(gp, ref[2])


def test_transposes_of_the_transforms():
    """rfft3^T = N^3 irfft3 and curl^T = curl under <,>_w, for arbitrary (non-Hermitian) spectra."""
    d = dns.Domain(12)
    rng = np.random.default_rng(1)
    a = rng.standard_normal((d.n, d.n, d.n // 2 + 1)) + 1j * rng.standard_normal((d.n, d.n, d.n // 2 + 1))
    x = rng.standard_normal((d.n, d.n, d.n))
    lhs = float(np.sum(dns.irfft3(a, d.n) * x))
    rhs = adj.inner(d, a, dns.rfft3(x)) / d.n ** 3
    assert lhs == pytest.approx(rhs, rel=1e-12)
    b = dns.rfft3(rng.standard_normal((3, d.n, d.n, d.n)))
    c = dns.rfft3(rng.standard_normal((3, d.n, d.n, d.n)))
    assert adj.inner(d, b, dns.curl(d, c)) == pytest.approx(adj.inner(d, dns.curl(d, b), c), rel=1e-12)
