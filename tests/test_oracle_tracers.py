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
