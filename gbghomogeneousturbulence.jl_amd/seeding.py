"""Seeded initial conditions of the benchmark configurations (host-side set-up, not the step).

The reference's own seeder draws host white noise, transforms it and shapes it with a Gaussian
envelope (src/State.jl:65-78,193-287); the configurations this library is measured on ask for a
different spectrum, E(k) ~ k^4 exp(-2 (k/k_p)^2), with random phases, which the reference refuses
by design (src/SimulationRun.jl:402-409) and adopts through
``Simulation(domain; state = ResolvedFields(u_hat, phi_hat))`` (src/SimulationRun.jl:769-788).

Random numbers come from a counter-based generator: every value is a hash of (seed, stream,
mode), so a mode draws the same phases on every grid size and every backend.
"""
from __future__ import annotations

import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def _mix64(z: np.ndarray) -> np.ndarray:
    z = z.astype(np.uint64, copy=True)
    with np.errstate(over="ignore"):
        z ^= z >> np.uint64(30)
        z *= np.uint64(0xBF58476D1CE4E5B9)
        z ^= z >> np.uint64(27)
        z *= np.uint64(0x94D049BB133111EB)
        z ^= z >> np.uint64(31)
    return z


def counter_uniform(seed: int, stream: int, counter: np.ndarray) -> np.ndarray:
    """Uniform doubles in [0, 1) from (seed, stream, counter), 53 random bits each."""
    key = _mix64(np.array([(seed & 0xFFFFFFFF) * 0x100 + (stream & 0xFF)], dtype=np.uint64))[0]
    with np.errstate(over="ignore"):
        z = counter.astype(np.uint64) * _GOLDEN + key
    return (_mix64(z) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def _mode_counter(mx, my, mz) -> np.ndarray:
    off = 1 << 20
    return ((mz.astype(np.int64) + off) << 42) | ((my.astype(np.int64) + off) << 21) | mx.astype(np.int64)


def _wn(n: int) -> np.ndarray:
    i = np.arange(n)
    return np.where(i <= n // 2, i, i - n)


def _ball_modes(n: int):
    """Signed indices (mx, my, mz) and storage indices (iz, iy, ix) of the nonzero modes inside the
    2/3 ball, in C order of the [N, N, N/2+1] layout, without materialising full index grids."""
    mys = _wn(n)
    m2 = (mys[:, None, None] ** 2 + mys[None, :, None] ** 2).astype(np.int32) + \
        (np.arange(n // 2 + 1, dtype=np.int32) ** 2)[None, None, :]
    iz, iy, ix = np.nonzero((9 * m2 <= n * n) & (m2 > 0))
    return ix, mys[iy], mys[iz], (iz, iy, ix)


def parseval_energy(u_hat: np.ndarray, n: int) -> float:
    """1/2 <u, u> of a spectral vector in the reference layout [3, N, N, N/2+1]."""
    w = np.full(n // 2 + 1, 2.0)
    w[0] = 1.0
    w[-1] = 1.0
    return 0.5 * float(np.sum(np.abs(u_hat) ** 2 * w)) / float(n) ** 6


def hermitian_kx0(f: np.ndarray, n: int) -> np.ndarray:
    """Make the kx = 0 and kx = N/2 planes conjugate-symmetric (the reality projection)."""
    neg = (-np.arange(n)) % n
    for p in (0, n // 2):
        s = f[..., p]
        f[..., p] = 0.5 * (s + np.conj(s[..., neg, :][..., :, neg]))
    return f


def random_phase_velocity(n: int, side_length: float = 2 * math.pi, energy: float = 0.5,
                          peak: float = 4.0, seed: int = 2026) -> np.ndarray:
    """Random-phase solenoidal u_hat with E(k) ~ k^4 exp(-2 (k/k_p)^2), zero mean, inside the 2/3 ball.

    Rogallo (1981) construction: u_hat = a(|k|) (alpha e1 + beta e2) with e1, e2 an orthonormal
    pair normal to k and (alpha, beta) = (e^{i t1} cos g, e^{i t2} sin g) from three uniforms;
    a^2 ~ E(k)/(4 pi k^2).  Normalised to 1/2 <|u|^2> = ``energy``; reference layout
    [3, N, N, N/2+1] (unnormalised forward-transform convention).
    """
    k0 = 2 * math.pi / side_length
    mx, my, mz, keep = _ball_modes(n)
    kx, ky, kz = (k0 * mx, k0 * my, k0 * mz)
    kk = np.sqrt(kx * kx + ky * ky + kz * kz)
    kp = peak * k0
    amp = np.sqrt(kk ** 2 * np.exp(-2.0 * (kk / kp) ** 2))          # sqrt(E/(4 pi k^2)) up to a constant
    ctr = _mode_counter(mx, my, mz)
    t1 = 2 * math.pi * counter_uniform(seed, 0, ctr)
    t2 = 2 * math.pi * counter_uniform(seed, 1, ctr)
    g = 2 * math.pi * counter_uniform(seed, 2, ctr)
    alpha = np.exp(1j * t1) * np.cos(g)
    beta = np.exp(1j * t2) * np.sin(g)
    kxy = np.sqrt(kx * kx + ky * ky)
    on_axis = kxy == 0
    safe = np.where(on_axis, 1.0, kxy)
    e1 = np.stack([np.where(on_axis, 1.0, ky / safe), np.where(on_axis, 0.0, -kx / safe), np.zeros_like(kx)])
    e2 = np.stack([np.where(on_axis, 0.0, kx * kz / (kk * safe)), np.where(on_axis, 1.0, ky * kz / (kk * safe)),
                   np.where(on_axis, 0.0, -kxy / kk)])
    u = np.zeros((3, n, n, n // 2 + 1), dtype=np.complex128)
    for c in range(3):
        u[c][keep] = amp * (alpha * e1[c] + beta * e2[c])
    hermitian_kx0(u, n)
    e = parseval_energy(u, n)
    if not e > 0 and energy > 0:
        raise ValueError(f"the {n}^3 grid's dealias ball holds no nonzero mode, so no field of kinetic energy "
                         f"{energy} can be seeded on it; use a finer grid")
    if e > 0:
        u *= math.sqrt(energy / e)
    return u


def random_phase_scalar(n: int, side_length: float = 2 * math.pi, variance: float = 0.5,
                        peak: float = 4.0, seed: int = 2027) -> np.ndarray:
    """Random-phase scalar with the same spectral shape, zero mean, inside the 2/3 ball, <phi^2> = variance."""
    k0 = 2 * math.pi / side_length
    mx, my, mz, keep = _ball_modes(n)
    kk = k0 * np.sqrt((mx * mx + my * my + mz * mz).astype(np.float64))
    amp = np.sqrt(kk ** 2 * np.exp(-2.0 * (kk / (peak * k0)) ** 2))
    t = 2 * math.pi * counter_uniform(seed, 3, _mode_counter(mx, my, mz))
    p = np.zeros((n, n, n // 2 + 1), dtype=np.complex128)
    p[keep] = amp * np.exp(1j * t)
    hermitian_kx0(p, n)
    v = 2.0 * parseval_energy(p[None], n)
    if not v > 0 and variance > 0:
        raise ValueError(f"the {n}^3 grid's dealias ball holds no nonzero mode, so no scalar of variance "
                         f"{variance} can be seeded on it; use a finer grid")
    if v > 0:
        p *= math.sqrt(variance / v)
    return p


def uniform_positions(m: int, side_length: float = 2 * math.pi, seed: int = 2028, dtype=np.float64) -> np.ndarray:
    """M tracer positions uniform in [0, L)^3, layout [3, M] (the (M, 3) column-major matrix)."""
    ctr = np.arange(m, dtype=np.int64)
    x = np.stack([counter_uniform(seed, 8 + d, ctr) for d in range(3)]) * side_length
    x = x.astype(dtype)
    return np.where(x >= dtype(side_length), dtype(0), x)
