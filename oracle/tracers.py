"""Tracer restatement: Catmull-Rom gather, Heun and Adams pushes, wrap, Adams and Fornberg
weights, cell keys.

Test infrastructure only (see oracle/__init__.py).  The reference delegates the gather and
the displacement to its field library (``interpolate!``, ``displace!``); their contracts are
restated here:

  * Catmull-Rom = LalescuSpline(4), the TurTLE n1m1 C^1 spline (src/Particles.jl:63-69,
    docs/src/validation.md:84): grid point j sits at x_j = j*L/N; with s = x*(N/L),
    i = floor(s), t = s - i, the 4 weights for points i-1, i, i+1, i+2 are
      w(-1) = t((2 - t)t - 1)/2, w(0) = (t^2(3t - 5) + 2)/2,
      w(+1) = t((4 - 3t)t + 1)/2, w(+2) = t^2(t - 1)/2,
    applied separably on x, y, z with periodic indices.
  * wrap (test/regressions/02-api-contracts-02.jl:1-6, 02-free-modules-02.jl:3-8):
      r = fmod(x, L); if r < 0: r += L; if r == L: r = 0  (NaN and Inf pass through).
  * cell key: the canonical cell index floor(x*(N/L)) mod N per axis, with (N/L) rounded to
    the element type once and the product formed in the element type.
"""
from __future__ import annotations

import math

import numpy as np

# src/ParticlePropagator.jl:261-262
_GAUSS_NODES = (-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526)
_GAUSS_WEIGHTS = (0.3478548451374538, 0.6521451548625461, 0.6521451548625461, 0.3478548451374538)


def adams_weights(nodes, h: float):
    """Integral over [0, h] of each Lagrange basis polynomial on ``nodes`` (src/ParticlePropagator.jl:264-280)."""
    out = []
    n = len(nodes)
    for j in range(n):
        weight = 0.0
        for xi, w in zip(_GAUSS_NODES, _GAUSS_WEIGHTS):
            t = h * (xi + 1) / 2
            basis = 1.0
            for m in range(n):
                if m == j:
                    continue
                basis *= (t - nodes[m]) / (nodes[j] - nodes[m])
            weight += w * h / 2 * basis
        out.append(weight)
    return tuple(out)


def catmull_rom(t):
    """Weights for offsets -1, 0, +1, +2 at fractional position t in [0, 1)."""
    return (t * ((2 - t) * t - 1) / 2,
            (t * t * (3 * t - 5) + 2) / 2,
            t * ((4 - 3 * t) * t + 1) / 2,
            t * t * (t - 1) / 2)


def cell_coordinates(X: np.ndarray, n: int, side_length: float, dtype=np.float64):
    """Integer cell (i, j, k) and fractional offsets of each position, in the element type."""
    Xd = np.asarray(X, dtype=dtype)
    inv_dx = dtype(n / side_length)
    s = Xd * inv_dx
    fl = np.floor(s)
    cell = np.mod(fl.astype(np.int64), n)
    frac = (s - fl).astype(np.float64)
    return cell, frac


def cell_keys(X: np.ndarray, n: int, side_length: float, dtype=np.float64) -> np.ndarray:
    """Linear cell index (k*N + j)*N + i of each tracer -- bit-exact contract with the device."""
    cell, _ = cell_coordinates(X, n, side_length, dtype)
    return ((cell[:, 2] * n + cell[:, 1]) * n + cell[:, 0]).astype(np.int64)


def interpolate(field: np.ndarray, X: np.ndarray, side_length: float, dtype=np.float64) -> np.ndarray:
    """Catmull-Rom gather of a physical field [c, z, y, x] (or [z, y, x]) at positions X[M, 3].

    PF ``interpolate!`` as called by src/Particles.jl:290-293 and src/ParticlePropagator.jl:309,382.
    """
    scalar = field.ndim == 3
    f = field[None] if scalar else field
    n = f.shape[-1]
    cell, frac = cell_coordinates(X, n, side_length, dtype)
    wx = catmull_rom(frac[:, 0])
    wy = catmull_rom(frac[:, 1])
    wz = catmull_rom(frac[:, 2])
    out = np.zeros((f.shape[0], X.shape[0]))
    for c in range(4):
        kk = (cell[:, 2] + c - 1) % n
        for b in range(4):
            jj = (cell[:, 1] + b - 1) % n
            row = np.zeros((f.shape[0], X.shape[0]))
            for a in range(4):
                ii = (cell[:, 0] + a - 1) % n
                row += wx[a] * f[:, kk, jj, ii]
            out += (wz[c] * wy[b]) * row
    return out[0] if scalar else out.T


def wrap(x: np.ndarray, side_length: float) -> np.ndarray:
    """Periodic reduction into [0, L) (the ``displace!`` wrap contract)."""
    L = np.asarray(side_length, dtype=x.dtype)
    r = np.fmod(x, L)
    r = np.where(r < 0, r + L, r)
    return np.where(r == L, np.zeros_like(r), r)


def heun_advect(X: np.ndarray, v_prev: np.ndarray, u_new: np.ndarray, h: float, side_length: float):
    """One Heun (= ABM(1)) push and the closing sample (src/ParticlePropagator.jl:288-315,349-386).

    X* = wrap(X + b1 v_prev);  f* = u_new(X*);  X' = wrap(X* + beta1 f* + (beta2 - b1) v_prev);
    v' = u_new(X').  Returns (X', v').
    """
    (b1,) = adams_weights((0.0,), h)
    beta1, beta2 = adams_weights((h, 0.0), h)
    Xs = wrap(X + b1 * v_prev, side_length)
    fs = interpolate(u_new, Xs, side_length)
    Xn = wrap(Xs + beta1 * fs + (beta2 - b1) * v_prev, side_length)
    vn = interpolate(u_new, Xn, side_length)
    return Xn, vn


def adams_advect(X: np.ndarray, samples, times, u_new: np.ndarray, h: float, side_length: float,
                 corrected: bool, unwrapped: np.ndarray | None = None):
    """One AB(J) step, or with ``corrected`` one ABM(J) step, of the positions X[M, 3]
    (src/ParticlePropagator.jl:287-316): ``samples`` are the trail's velocity samples [M, 3],
    newest first, taken at ``times``; the Adams weights come from the samples' actual times.

        X* = wrap(X + sum_j b_j f_j)
        X' = wrap(X* + beta_* f* + sum_j (beta_{j+1} - b_j) f_j),  f* = u_new(X*)   (corrected)

    Each sum is added left to right in the order written, as one fused displacement.  Given the
    unwrapped record U[M, 3] (src/ParticlePropagator.jl:299,313), returns (X', U') with U taking the
    same terms in the same order and never wrapped.
    """
    nodes = tuple(float(t) - float(times[0]) for t in times)
    b = adams_weights(nodes, h)
    acc, uacc = X, unwrapped
    for bj, f in zip(b, samples):
        acc = acc + bj * f
        if uacc is not None:
            uacc = uacc + bj * f
    Xs = wrap(acc, side_length)
    if not corrected:
        return Xs if unwrapped is None else (Xs, uacc)
    fs = interpolate(u_new, Xs, side_length)
    beta = adams_weights((h,) + nodes, h)
    acc = Xs + beta[0] * fs
    if uacc is not None:
        uacc = uacc + beta[0] * fs
    for j, f in enumerate(samples):
        acc = acc + (beta[j + 1] - b[j]) * f
        if uacc is not None:
            uacc = uacc + (beta[j + 1] - b[j]) * f
    return wrap(acc, side_length) if unwrapped is None else (wrap(acc, side_length), uacc)


def derivative_weights(nodes, order: int) -> np.ndarray:
    """Fornberg's finite-difference weights at 0 on arbitrary ``nodes`` (Math. Comp. 51 (1988)
    699-706): w[j, m] multiplies the value at node j in the derivative of order m, m = 0..order.
    The reference computes the same weights for time_derivative! (src/Particles.jl:533-565)."""
    x = [float(v) for v in nodes]
    n = len(x)
    c = np.zeros((n, order + 1))
    c[0, 0] = 1.0
    prod_prev = 1.0
    for i in range(1, n):
        top = min(i, order)
        prod = 1.0
        for j in range(i):
            gap = x[i] - x[j]
            prod *= gap
            if j == i - 1:   # the new node's row, from the previous node's (before that row moves)
                for k in range(top, 0, -1):
                    c[i, k] = prod_prev * (k * c[i - 1, k - 1] - x[i - 1] * c[i - 1, k]) / prod
                c[i, 0] = -prod_prev * x[i - 1] * c[i - 1, 0] / prod
            for k in range(top, 0, -1):
                c[j, k] = (x[i] * c[j, k] - k * c[j, k - 1]) / gap
            c[j, 0] = x[i] * c[j, 0] / gap
        prod_prev = prod
    return c


def time_derivative(samples, times, order: int = 1) -> np.ndarray:
    """d^m f/dt^m at the newest sample's time of the polynomial through every sample
    (src/Particles.jl:597-631); ``samples`` newest first."""
    nodes = [float(t) - float(times[0]) for t in times]
    w = derivative_weights(nodes, order)
    out = np.zeros_like(np.asarray(samples[0], dtype=np.float64))
    for j, f in enumerate(samples):
        out = out + w[j, order] * f
    return out


def uniform_positions(m: int, side_length: float, seed: int) -> np.ndarray:
    """Helper for tests: uniform positions in [0, L)^3 from numpy's PCG64 (not the product's seeder)."""
    rng = np.random.default_rng(seed)
    return rng.random((m, 3)) * side_length


def _sanity() -> None:  # pragma: no cover - manual check
    assert math.isclose(sum(catmull_rom(0.3)), 1.0)
