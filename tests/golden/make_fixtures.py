"""Writes the golden fixtures of this directory from the FP64 oracle (run from the repo root:
``python tests/golden/make_fixtures.py``).

The reference holds no golden vectors and cannot run here (it is Julia), so these are the oracle's
own outputs on small seeded cases, committed so that the oracle cannot drift unnoticed and so that
the device is held to fixed numbers as well as to a live oracle run.  Inputs are stored with the
outputs; only the modes inside the 2/3 ball are stored (every other mode is zero).
"""
from __future__ import annotations

import math
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

from oracle import dns  # noqa: E402
from oracle import tracers as trc  # noqa: E402

L = 2 * math.pi

CASES = {
    # name: (n, scheme, viscosity, schmidt, forcing, mean_gradient, dt, seed)
    "dns16_rk4_forced": (16, "rk4", 0.02, 0.7, (0.1, 1, 2), 1.0, 1e-2, 101),
    "dns16_rk3_unforced": (16, "rk3", 0.02, 1.0, None, 1.0, 1e-2, 102),
}
STEPS = (1, 10)


def initial_state(d: dns.Domain, seed: int):
    """A random solenoidal velocity with energy 1/2 and a random scalar, inside the ball, zero mean."""
    rng = np.random.default_rng(seed)
    k = np.sqrt(d.k2)
    env = np.where(d.mask, k ** 2 * np.exp(-(k / 3.0) ** 2), 0.0)
    u = dns.rfft3(rng.standard_normal((3, d.n, d.n, d.n))) * env
    u[:, 0, 0, 0] = 0
    u = dns.reality_projection(d, dns.leray(d, u))
    u *= math.sqrt(1.0 / dns.weighted_inner(d, u))          # k = <u, u>/2 = 1/2
    p = dns.rfft3(rng.standard_normal((d.n,) * 3)) * env * 0.3
    p[0, 0, 0] = 0
    p = dns.reality_projection(d, p)
    return u, p


def ball(d: dns.Domain, f: np.ndarray) -> np.ndarray:
    return f[..., d.mask]


def conditions(case):
    n, scheme, nu, sc, forcing, G, dt, seed = case
    return dns.Conditions(viscosity=nu, schmidt=sc, forcing=forcing, mean_gradient=G)


def make_dns(name: str) -> None:
    case = CASES[name]
    n, scheme, nu, sc, forcing, G, dt, seed = case
    d = dns.Domain(n)
    cond = conditions(case)
    u, p = initial_state(d, seed)
    out = {"n": n, "scheme": scheme, "viscosity": nu, "schmidt": sc, "forcing": np.array(forcing or (0, 0, 0), float),
           "mean_gradient": G, "dt": dt, "u0": ball(d, u), "p0": ball(d, p)}
    uo, po, dep = u, p, None
    for s in range(1, max(STEPS) + 1):
        uo, po, dep = dns.step(d, cond, scheme, uo, po, dt, dep)
        if s in STEPS:
            out[f"u{s}"] = ball(d, uo)
            out[f"p{s}"] = ball(d, po)
            r = dns.reductions(d, cond, uo, po)
            out[f"reductions{s}"] = np.array([r["kinetic_energy"], r["dissipation"], r["scalar_variance"],
                                             r["scalar_dissipation"]])
            out[f"spectrum{s}"] = dns.energy_spectrum(d, uo)[1]
            out[f"max_speed{s}"] = dep.max_speed
    np.savez_compressed(HERE / f"{name}.npz", **out)


def make_tracers() -> None:
    """64 tracers in the 16^3 forced RK4 case: samples at t = 0 and positions and samples after
    3 Heun steps (the field steps alongside)."""
    case = CASES["dns16_rk4_forced"]
    n, scheme, nu, sc, forcing, G, dt, seed = case
    d = dns.Domain(n)
    cond = conditions(case)
    u, p = initial_state(d, seed)
    X = trc.uniform_positions(64, L, 104)                  # [M, 3]
    dep = dns.deposit(d, u, p)
    out = {"X0": X.copy(), "v0": trc.interpolate(dep.u, X, L), "s0": trc.interpolate(dep.phi, X, L)}
    v = out["v0"]
    for _ in range(3):
        u, p, dep = dns.step(d, cond, scheme, u, p, dt, dep)
        X, v = trc.heun_advect(X, v, dep.u, dt, L)
    out.update({"X3": X, "v3": v, "s3": trc.interpolate(dep.phi, X, L)})
    np.savez_compressed(HERE / "tracers16_rk4_forced.npz", **out)


if __name__ == "__main__":
    for name in CASES:
        make_dns(name)
    make_tracers()
    for f in sorted(HERE.glob("*.npz")):
        print(f.name, f.stat().st_size, "bytes")
