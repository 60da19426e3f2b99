"""Pins of the tracer oracle: the reference's particle tests re-expressed
(test/particles.jl:27-40,94-140; test/regressions/02-api-contracts-02.jl, 02-free-modules-02.jl)."""
import math

import numpy as np
import pytest

from oracle import tracers as trc

L = 2 * math.pi


def test_adams_weights_are_the_classical_ones():
    h = 0.1
    assert trc.adams_weights((0.0,), h)[0] == pytest.approx(h, rel=1e-15)                 # AB1
    b = trc.adams_weights((h, 0.0), h)                                                   # trapezoid
    assert b == pytest.approx((h / 2, h / 2), rel=1e-15)
    ab2 = trc.adams_weights((0.0, -h), h)
    assert ab2 == pytest.approx((1.5 * h, -0.5 * h), rel=1e-14)
    ab3 = trc.adams_weights((0.0, -h, -2 * h), h)
    assert ab3 == pytest.approx((23 * h / 12, -16 * h / 12, 5 * h / 12), rel=1e-14)


def test_catmull_rom_partition_of_unity_and_linear_reproduction():
    t = np.linspace(0, 1, 17, endpoint=False)
    w = np.array(trc.catmull_rom(t))
    assert np.max(np.abs(w.sum(axis=0) - 1)) < 1e-15
    offsets = np.array([-1, 0, 1, 2])[:, None]
    assert np.max(np.abs((w * offsets).sum(axis=0) - t)) < 1e-15                         # reproduces x
    assert np.all(np.array(trc.catmull_rom(0.0)) == [0, 1, 0, 0])                        # interpolates


def test_constant_field_samples_exactly_and_linear_field_is_reproduced():
    n = 16
    X = trc.uniform_positions(40, L, 2).T.copy()
    const = np.full((n, n, n), 2.5)
    assert np.max(np.abs(trc.interpolate(const, X, L) - 2.5)) < 1e-14
    # u_x linear in x away from the periodic jump: reproduced exactly
    x = np.arange(n) * L / n
    lin = np.broadcast_to((0.3 * x)[None, None, :], (n, n, n))
    Xin = X.copy()
    Xin[:, 0] = np.linspace(2 * L / n, L - 3 * L / n, len(X))
    assert np.max(np.abs(trc.interpolate(lin, Xin, L) - 0.3 * Xin[:, 0])) < 1e-13


def test_uniform_flow_advects_exactly():
    n = 16
    X0 = np.array([[0.1, 0.2, 0.3], [1.0, 2.0, 3.0], [2.0, 3.0, 4.0]])
    c, steps, dt = 0.5, 20, 0.005
    u = np.zeros((3, n, n, n))
    u[0] = c
    X = X0.copy()
    v = trc.interpolate(u, X, L)
    for _ in range(steps):
        X, v = trc.heun_advect(X, v, u, dt, L)
    assert np.allclose(X[:, 0], np.mod(X0[:, 0] + c * steps * dt, L), rtol=0, atol=1e-14)
    assert np.array_equal(X[:, 1:], X0[:, 1:])


def test_heun_is_second_order_on_a_linear_strain():
    """x' = s (x - xc), exact x(t) = xc + (x0 - xc) e^{s t}; Catmull-Rom reproduces the linear field."""
    n, s, xc, x0, T = 16, 1.0, math.pi, math.pi + 0.4, 0.5
    xg = np.arange(n) * L / n
    u = np.zeros((3, n, n, n))
    u[0] = s * (xg - xc)[None, None, :]
    errs = []
    for steps in (10, 20, 40):
        X = np.array([[x0, 1.0, 1.0]])
        v = trc.interpolate(u, X, L)
        for _ in range(steps):
            X, v = trc.heun_advect(X, v, u, T / steps, L)
        errs.append(abs(X[0, 0] - (xc + (x0 - xc) * math.exp(s * T))))
    assert 3.5 < errs[0] / errs[1] < 4.5 and 3.5 < errs[1] / errs[2] < 4.5


def test_wrap_contract():
    x = np.array([-0.5, 0.0, 1.0, L, L + 0.25, -L, np.nan, np.inf])
    r = trc.wrap(x, L)
    assert r[0] == pytest.approx(L - 0.5)
    assert r[1] == 0.0 and r[2] == 1.0 and r[3] == 0.0
    assert r[4] == pytest.approx(0.25)
    assert r[5] == 0.0
    assert np.isnan(r[6]) and np.isnan(r[7])            # never laundered into the box
    tiny = np.array([-1e-20])
    assert trc.wrap(tiny, L)[0] == 0.0                  # -tiny + L rounds to L, which wraps to 0


def test_cell_keys_in_float32_use_float32_arithmetic():
    n = 64
    X = np.array([[np.float32(L) * np.float32(0.999999), 0.5, 6.2]], dtype=np.float32)
    cell, _ = trc.cell_coordinates(X, n, L, np.float32)
    s = X.astype(np.float32) * np.float32(n / L)
    assert np.array_equal(cell, np.mod(np.floor(s).astype(np.int64), n))
    assert trc.cell_keys(X, n, L, np.float32)[0] == (cell[0, 2] * n + cell[0, 1]) * n + cell[0, 0]


# --- the Adams family, the trail and its time derivative ----------------------------------------

def test_adams_weights_integrate_polynomials_of_degree_below_the_node_count_exactly():
    """sum_j w_j p(t_j) = int_0^h p for deg p <= J - 1 on uneven nodes, with and without h among
    them (the AB and AM weights, src/ParticlePropagator.jl:256-280); J up to 8."""
    rng = np.random.default_rng(3)
    h = 0.07
    for J in range(1, 9):
        past = -np.cumsum(rng.uniform(0.5, 1.5, J - 1)) * h
        for nodes in ((0.0,) + tuple(past), (h, 0.0) + tuple(past[:-1]) if J > 1 else (h,)):
            w = trc.adams_weights(nodes, h)
            for deg in range(len(nodes)):
                coef = rng.standard_normal(deg + 1)
                p = np.polynomial.Polynomial(coef)
                want = p.integ()(h) - p.integ()(0.0)
                got = sum(wj * p(t) for wj, t in zip(w, nodes))
                assert got == pytest.approx(want, rel=1e-11, abs=1e-14), (J, deg)


def test_fornberg_weights_differentiate_polynomials_exactly_and_give_the_classical_stencils():
    w = trc.derivative_weights([0.0, -1.0, -2.0], 2)
    assert np.allclose(w[:, 1], [1.5, -2.0, 0.5], atol=1e-15)            # second-order backward
    assert np.allclose(w[:, 2], [1.0, -2.0, 1.0], atol=1e-15)
    rng = np.random.default_rng(4)
    for n in range(2, 9):
        nodes = np.concatenate([[0.0], -np.cumsum(rng.uniform(0.5, 1.5, n - 1)) * 0.1])
        w = trc.derivative_weights(nodes, min(3, n - 1))
        p = np.polynomial.Polynomial(rng.standard_normal(n))
        for m in range(min(3, n - 1) + 1):
            want = p.deriv(m)(0.0) if m else p(0.0)
            assert float(w[:, m] @ p(nodes)) == pytest.approx(want, rel=1e-8, abs=1e-9), (n, m)


def _strain_field(n, s, xc):
    xg = np.arange(n) * L / n
    u = np.zeros((3, n, n, n))
    u[0] = s * (xg - xc)[None, None, :]
    return u


def run_trail(kind, k, X, u, dts, samples=None, times=None):
    """The run's loop for a steady field u: advect over the trail, then sample at the new positions
    (src/ParticlePropagator.jl:349-372, :378-386, src/Particles.jl:454-490); the trail keeps
    k + 1 samples under ABM(k), k under AB(k)."""
    cap = k if kind == "ab" else k + 1
    samples = [trc.interpolate(u, X, L)] if samples is None else list(samples)
    times = [0.0] if times is None else list(times)
    for dt in dts:
        full = len(samples) >= k
        J = k if full else len(samples)
        X = trc.adams_advect(X, samples[:J], times[:J], u, dt, L, corrected=(not full) or kind == "abm")
        samples = [trc.interpolate(u, X, L)] + samples[: cap - 1]
        times = [times[0] + dt] + times[: cap - 1]
    return X, samples, times


@pytest.mark.parametrize("kind,k,order", [("ab", 2, 2), ("ab", 3, 3), ("ab", 4, 4), ("ab", 5, 5),
                                          ("abm", 1, 2), ("abm", 2, 3), ("abm", 3, 4), ("abm", 4, 5)])
def test_adams_schemes_reach_their_order_from_an_exact_history(kind, k, order):
    """x' = s (x - xc) with Catmull-Rom exact on the linear field: from a trail of exact samples at
    the exact past positions every step is the full scheme, and the error falls as dt^order."""
    n, s, xc, x0, T = 16, 1.0, math.pi, math.pi + 0.4, 0.5
    u = _strain_field(n, s, xc)
    errs = []
    for steps in (16, 32, 64):
        h = T / steps
        ts = [-j * h for j in range(k)]
        xs = [xc + (x0 - xc) * math.exp(s * t) for t in ts]
        samples = [np.array([[s * (x - xc), 0.0, 0.0]]) for x in xs]
        X = np.array([[x0, 1.0, 1.0]])
        X, _, _ = run_trail(kind, k, X, u, [h] * steps, samples, ts)
        errs.append(abs(X[0, 0] - (xc + (x0 - xc) * math.exp(s * T))))
    r = 2.0 ** order
    assert 0.75 * r < errs[0] / errs[1] < 1.3 * r and 0.75 * r < errs[1] / errs[2] < 1.3 * r, errs


def test_abm1_through_the_trail_is_the_heun_push():
    n = 16
    rng = np.random.default_rng(9)
    u = rng.standard_normal((3, n, n, n))
    X = trc.uniform_positions(50, L, 3)
    v = trc.interpolate(u, X, L)
    want, _ = trc.heun_advect(X, v, u, 0.01, L)
    got, _, _ = run_trail("abm", 1, X, u, [0.01])
    assert np.array_equal(got, want)


def test_time_derivative_of_a_trail_converges_to_the_acceleration():
    """Along x' = s (x - xc) the acceleration is s^2 (x - xc); the derivative of the polynomial
    through d uneven samples errs by O(dt^(d-1))."""
    s, xc, x0 = 1.3, 0.0, 0.4
    errs = []
    for h in (0.04, 0.02, 0.01):
        ts = [0.0, -h, -2.3 * h, -3.1 * h]
        samples = [np.array([s * (x0 * math.exp(s * t) - xc)]) for t in ts]
        a = trc.time_derivative(samples, ts, 1)
        errs.append(abs(a[0] - s * s * x0))
    assert 6.0 < errs[0] / errs[1] < 10.0 and 6.0 < errs[1] / errs[2] < 10.0
    assert trc.time_derivative(samples, ts, 0)[0] == samples[0][0]


@pytest.mark.parametrize("kind,k", [("ab", 3), ("abm", 2), ("abm", 1)])
def test_unwrapped_record_of_a_uniform_flow_is_the_straight_line(kind, k):
    """src/ParticlePropagator.jl:299,313: the unwrapped record takes every displacement term
    without the wrap, so in a uniform flow c it is X0 + c t across several windings (the Adams
    weights of each step sum to the step), and reduced to the box it is the wrapped position."""
    n = 16
    c = np.array([7.0, 0.0, -3.5])
    u = np.zeros((3, n, n, n))
    u[0], u[2] = c[0], c[2]
    X0 = trc.uniform_positions(20, L, 5)
    cap = k if kind == "ab" else k + 1
    dts = [0.05 * (1 + 0.3 * math.sin(i)) for i in range(40)]
    X, U = X0.copy(), X0.copy()
    samples, times = [trc.interpolate(u, X, L)], [0.0]
    for dt in dts:
        full = len(samples) >= k
        J = k if full else len(samples)
        X, U = trc.adams_advect(X, samples[:J], times[:J], u, dt, L, corrected=(not full) or kind == "abm",
                                unwrapped=U)
        samples = [trc.interpolate(u, X, L)] + samples[: cap - 1]
        times = [times[0] + dt] + times[: cap - 1]
    T = sum(dts)
    assert np.max(np.abs(U - (X0 + c * T))) < 1e-11
    assert np.array_equal(U[:, 1], X0[:, 1])
    winding = (U - X) / L
    assert np.max(np.abs(winding - np.round(winding))) < 1e-12
    assert set(np.round(winding[:, 0])) >= {2.0} and set(np.round(winding[:, 2])) <= {-1.0, -2.0}
    assert np.array_equal(trc.adams_advect(X0, samples[:1], times[:1], u, 0.1, L, corrected=True),
                          trc.adams_advect(X0, samples[:1], times[:1], u, 0.1, L, corrected=True, unwrapped=X0)[0])


# --- the start-up and the stability of the Adams family (test/particles.jl:30-85, 233-318) ---------
# x' = cos(w t) s (x - xc): Catmull-Rom is exact on the linear field, and the exact path is
# xc + (x0 - xc) exp(s sin(w t) / w).
S_, W_, XC_, X0_ = 1.0, 3.0, math.pi, math.pi + 0.4


def _strain_exact(t):
    return XC_ + (X0_ - XC_) * math.exp(S_ * math.sin(W_ * t) / W_)


def _strain_grid(t, n=16):
    u = np.zeros((3, n, n, n))
    u[0] = math.cos(W_ * t) * S_ * (np.arange(n) * L / n - XC_)[None, None, :]
    return u


def _strain_error(kind, k, steps, exact_start=False):
    """The largest error along a pass of the run's loop through `steps` (time-dependent field)."""
    cap = k if kind == "ab" else k + 1
    X = np.array([[X0_, 1.0, 1.0]])
    if exact_start:     # the exact earlier samples at -h, -2h, ... for the first step h
        h = steps[0]
        times = [-j * h for j in range(min(k, cap))]
        samples = [np.array([[math.cos(W_ * t) * S_ * (_strain_exact(t) - XC_), 0.0, 0.0]]) for t in times]
    else:
        times, samples = [0.0], [trc.interpolate(_strain_grid(0.0), X, L)]
    t, worst = 0.0, 0.0
    for dt in steps:
        full = len(samples) >= k
        J = k if full else len(samples)
        u_new = _strain_grid(t + dt)
        X = trc.adams_advect(X, samples[:J], times[:J], u_new, dt, L, corrected=(not full) or kind == "abm")
        t += dt
        samples = [trc.interpolate(u_new, X, L)] + samples[: cap - 1]
        times = [t] + times[: cap - 1]
        worst = max(worst, abs(X[0, 0] - _strain_exact(t)))
    return worst


def _uniform(n, T=2.0):
    return [T / n] * n


def _shortened(g, scheme, n, T=2.0):
    """The schedule shortened_first_steps gives with tau = 1/w (src/SimulationRun.jl:1099-1129),
    built with the host's own rule: nu / eps = tau^2."""
    dt, steps, t, trail = T / n, [], 0.0, [0.0]
    while t < T - 1e-12:
        h = min(dt, g.api.startup_bound(scheme, trail, dt, 1.0, W_ * W_), T - t)
        steps.append(h)
        t += h
        trail = [t] + trail[:1]
    return steps


def _order(errors):
    return math.log2(errors[-2] / errors[-1])


def test_the_ordinary_start_up_leaves_an_error_of_third_order():
    counts = (64, 128, 256)
    ref = [_strain_error("abm", 4, _uniform(n)) for n in counts]
    assert 2.7 < _order(ref) < 3.3
    for kind, k in (("ab", 6), ("ab", 8), ("abm", 7)):
        errs = [_strain_error(kind, k, _uniform(n)) for n in counts]
        assert 2.7 < _order(errs) < 3.3
        assert errs[-1] == pytest.approx(ref[-1], rel=0.05)
    errs = [_strain_error("ab", 2, _uniform(n)) for n in counts]
    assert 1.7 < _order(errs) < 2.3


def test_shortened_first_steps_restore_the_full_order(g):
    for kind, k in [("ab", k) for k in range(3, 9)] + [("abm", k) for k in range(2, 8)]:
        scheme = g.AB(k) if kind == "ab" else g.ABM(k)
        errs = [_strain_error(kind, k, _shortened(g, scheme, n)) for n in (32, 64, 128)]
        assert _order(errs) > scheme.order - 0.6, (kind, k, errs)
        assert errs[1] < 4 * _strain_error(kind, k, _uniform(64), exact_start=True), (kind, k)
    # at order 8 the doubling that serves order 5 loses the accuracy again
    dt, steps, t = 2.0 / 64, [], 0.0
    while t < 2.0 - 1e-12:
        h = min(dt, (1 / W_) * (dt * W_) ** (8 / 3)) if not steps else min(dt, 2.0 * steps[-1])
        h = min(h, 2.0 - t)
        steps.append(h)
        t += h
    assert _strain_error("ab", 8, steps) > 20 * _strain_error("ab", 8, _shortened(g, g.AB(8), 64))


def _real_axis_limit(k, corrected):
    """The largest h|lambda| at which the scheme at uniform step damps x' = lambda x: the recurrence
    y_{n+1} = sum_j c_j(z) y_{n+1-j} is stable when every root of its characteristic polynomial
    lies in the closed unit disc."""
    nodes = tuple(-float(i) for i in range(k))
    b = np.array(trc.adams_weights(nodes, 1.0))
    beta = np.array(trc.adams_weights((1.0,) + nodes, 1.0)) if corrected else None

    def stable(z):
        c = z * beta[1:] + z * z * beta[0] * b if corrected else z * b
        c = c.copy()
        c[0] += 1 + z * beta[0] if corrected else 1
        return np.max(np.abs(np.roots(np.concatenate([[1.0], -c])))) <= 1 + 1e-9

    lo, hi = 0.0, 3.0
    for x in np.arange(0.005, 3.0, 0.005):
        if not stable(-x):
            hi = x
            break
        lo = x
    for _ in range(30):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if stable(-mid) else (lo, mid)
    return lo


def test_the_stability_limits_on_the_real_axis_are_the_textbook_ones():
    for k, want in ((1, 2.0), (2, 1.0), (3, 6 / 11), (4, 3 / 10), (5, 90 / 551)):
        assert _real_axis_limit(k, False) == pytest.approx(want, abs=2e-3), k
    assert _real_axis_limit(1, True) == pytest.approx(2.0, abs=2e-3)       # Heun: 1 + z + z^2/2
    assert _real_axis_limit(6, False) == pytest.approx(0.088, abs=2e-3)
    assert _real_axis_limit(8, False) == pytest.approx(0.024, abs=2e-3)
    for k, want in ((3, 1.93), (5, 1.04), (7, 0.58)):
        assert _real_axis_limit(k, True) == pytest.approx(want, abs=0.02), k
