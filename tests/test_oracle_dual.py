"""The dual-number (forward-mode) derivative of the oracle step (CPU).

Three pins: the dual step's value part is the oracle step; its tangent part is the directional
derivative that fourth-order central differences of the oracle give; and the reverse-mode
gradient of the oracle adjoint, dotted with a direction, equals the forward-mode derivative
along it (the transposition identity <g, v> = J' v, which no finite difference limits).
"""
import numpy as np
import pytest

from oracle import adjoint as adj
from oracle import dns
from oracle import dual

from test_oracle_adjoint import _direction, _setup, fd4


@pytest.mark.parametrize("scheme", ["rk3", "rk4"])
def test_dual_value_is_the_oracle_step(scheme):
    d, cond, u, p = _setup(16)
    uo, po, _ = dns.step(d, cond, scheme, u, p, 2e-2)
    du, dp = _direction(d, 5), _direction(d, 6, scalar=True)
    ud, pd = dual.step(d, cond, scheme, dual.Dual(u, du), dual.Dual(p, dp), 2e-2)
    assert np.max(np.abs(ud.v - uo)) <= 1e-13 * np.max(np.abs(uo))
    assert np.max(np.abs(pd.v - po)) <= 1e-13 * np.max(np.abs(po))


@pytest.mark.parametrize("scheme", ["rk3", "rk4"])
@pytest.mark.parametrize("forcing", [None, 0.1])
def test_dual_derivative_equals_finite_differences(scheme, forcing):
    d, cond, u, p = _setup(16, forcing)
    dt, steps = 2e-2, 3
    du, dp = _direction(d, 31), _direction(d, 32, scalar=True)
    J, dJ = dual.directional_derivative(d, cond, scheme, u, p, dt, steps, du, dp)

    def f(e):
        uo, po = adj.run(d, cond, scheme, u + e * du, p + e * dp, dt, steps)
        return adj.objective(d, uo, po)

    assert J == pytest.approx(f(0.0), rel=1e-13)
    fd = fd4(f, 0.1)
    assert abs(dJ - fd) <= 1e-7 * abs(fd)


@pytest.mark.parametrize("scheme", ["rk3", "rk4"])
def test_adjoint_gradient_equals_the_dual_derivative(scheme):
    d, cond, u, p = _setup(16, 0.1)
    dt, steps = 2e-2, 5
    J, gu, gp = adj.gradient(d, cond, scheme, u, p, dt, steps, 2)
    for seed in (41, 42, 43):
        du, dp = _direction(d, seed), _direction(d, seed + 100, scalar=True)
        Jd, dJ = dual.directional_derivative(d, cond, scheme, u, p, dt, steps, du, dp)
        assert Jd == pytest.approx(J, rel=1e-13)
        an = adj.inner(d, gu, du) + adj.inner(d, gp, dp)
        assert abs(an - dJ) <= 1e-11 * abs(dJ)


def test_a_planted_adjoint_error_is_caught_by_the_dual_check(monkeypatch):
    """Dropping the forcing's amplitude derivative in the adjoint (a plausible slip) moves
    <g, v> far outside the dual check's bar."""
    d, cond, u, p = _setup(16, 0.1)
    dt, steps = 2e-2, 3
    du = _direction(d, 51)
    _, dJ = dual.directional_derivative(d, cond, "rk4", u, p, dt, steps, du)
    orig = adj.rhs_vjp

    def slipped(d_, cond_, uh, uu, w, phi, amp, eb, mu_u, mu_p):
        return orig(d_, cond_, uh, uu, w, phi, amp, float("inf"), mu_u, mu_p)   # drops the -A/e_b term

    monkeypatch.setattr(adj, "rhs_vjp", slipped)
    _, gu, _ = adj.gradient(d, cond, "rk4", u, p, dt, steps, 3)
    assert abs(adj.inner(d, gu, du) - dJ) > 1e-6 * abs(dJ)
