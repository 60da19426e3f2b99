"""The oracle against the committed golden fixtures (CPU).

tests/golden/*.npz hold the FP64 oracle's outputs on small seeded cases (written by
tests/golden/make_fixtures.py): states after 1 and 10 steps for IF-RK4 with ConstantPower forcing,
scalar and mean gradient, and IF-RK3 unforced; reductions, spectra and max|u|; tracer positions
and samples after 3 Heun steps.  The oracle must keep reproducing them.
"""
from pathlib import Path

import numpy as np
import pytest

from oracle import dns
from oracle import tracers as trc

GOLDEN = Path(__file__).resolve().parent / "golden"
L = 2 * np.pi
DNS_CASES = ["dns16_rk4_forced", "dns16_rk3_unforced"]


def load_case(name):
    """(domain, conditions, scheme, dt, fixture, unpack) with unpack(key) the full half spectrum."""
    z = np.load(GOLDEN / f"{name}.npz")
    n = int(z["n"])
    d = dns.Domain(n)
    f = tuple(z["forcing"])
    cond = dns.Conditions(viscosity=float(z["viscosity"]), schmidt=float(z["schmidt"]),
                          forcing=None if f[0] == 0 else (float(f[0]), int(f[1]), int(f[2])),
                          mean_gradient=float(z["mean_gradient"]))

    def unpack(key):
        v = z[key]
        full = np.zeros(v.shape[:-1] + d.mask.shape, dtype=np.complex128)
        full[..., d.mask] = v
        return full

    return d, cond, str(z["scheme"]), float(z["dt"]), z, unpack


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("name", DNS_CASES)
def test_oracle_reproduces_the_dns_fixture(name):
    d, cond, scheme, dt, z, unpack = load_case(name)
    u, p, dep = unpack("u0"), unpack("p0"), None
    for s in range(1, 11):
        u, p, dep = dns.step(d, cond, scheme, u, p, dt, dep)
        if s in (1, 10):
            assert rel(u, unpack(f"u{s}")) < 1e-13
            assert rel(p, unpack(f"p{s}")) < 1e-13
            r = dns.reductions(d, cond, u, p)
            got = [r["kinetic_energy"], r["dissipation"], r["scalar_variance"], r["scalar_dissipation"]]
            np.testing.assert_allclose(got, z[f"reductions{s}"], rtol=1e-13)
            np.testing.assert_allclose(dns.energy_spectrum(d, u)[1], z[f"spectrum{s}"], rtol=1e-12, atol=1e-300)
            assert dep.max_speed == pytest.approx(float(z[f"max_speed{s}"]), rel=1e-13)


def test_the_fixture_states_are_fields_the_step_accepts():
    """Inside the ball, solenoidal, Hermitian on kx = 0, and moving (the 10-step state is not the
    initial one), so the fixtures pin the step rather than a fixed point."""
    for name in DNS_CASES:
        d, cond, scheme, dt, z, unpack = load_case(name)
        u0, u10 = unpack("u0"), unpack("u10")
        assert dns.divergence_residual(d, u0) < 1e-12
        assert rel(dns.reality_projection(d, u0), u0) < 1e-15
        assert rel(u10, u0) > 1e-3


def test_oracle_reproduces_the_tracer_fixture():
    d, cond, scheme, dt, _, unpack = load_case("dns16_rk4_forced")
    t = np.load(GOLDEN / "tracers16_rk4_forced.npz")
    u, p = unpack("u0"), unpack("p0")
    dep = dns.deposit(d, u, p)
    X = t["X0"]
    v = trc.interpolate(dep.u, X, L)
    np.testing.assert_allclose(v, t["v0"], rtol=0, atol=1e-14)
    np.testing.assert_allclose(trc.interpolate(dep.phi, X, L), t["s0"], rtol=0, atol=1e-14)
    for _ in range(3):
        u, p, dep = dns.step(d, cond, scheme, u, p, dt, dep)
        X, v = trc.heun_advect(X, v, dep.u, dt, L)
    np.testing.assert_allclose(X, t["X3"], rtol=0, atol=1e-13)
    np.testing.assert_allclose(v, t["v3"], rtol=0, atol=1e-13)
    np.testing.assert_allclose(trc.interpolate(dep.phi, X, L), t["s3"], rtol=0, atol=1e-13)
