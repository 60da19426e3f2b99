"""The hand-written adjoint through the C ABI, on an MI355X.

Pins: the one-step VJP equals the oracle adjoint's to FP64 round-off; the checkpointed gradient
equals the oracle's and the forward-mode (dual-number) derivative of the oracle along random
directions; and at 128^3 over 200 forced IF-RK4 steps with checkpoints every 20 it matches
fourth-order central differences of device runs along 16 seeded solenoidal directions to 1e-6.
"""
import math
import time

import numpy as np
import pytest

from oracle import adjoint as oadj
from oracle import dns
from oracle import dual

from test_gpu_parity import L, make_sim, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scheme", ["rk3", "rk4"])
def test_step_vjp_matches_the_oracle_adjoint(g, scheme):
    n, dt = 32, 2e-3
    sch = g.IFLowStorageRK4 if scheme == "rk4" else g.IFLowStorageRK3
    sim, u, p, oc = make_sim(g, n, np.float64, scheme=sch, forcing=0.1, nu=0.02, seed=41)
    d = dns.Domain(n)
    lu, lp = oadj.solenoidal_direction(d, 1), dns.rfft3(np.random.default_rng(2).standard_normal((n, n, n))) * d.mask
    with sim:
        before = sim.state().velocity_spectral
        gu, gp = sim.step_vjp(dt, lu, lp)
        assert np.array_equal(sim.state().velocity_spectral, before)   # the state is left as it was
        ou, op = oadj.step_vjp(d, oc, scheme, u, p, dt, lu, lp)
        eu, ep = rel(gu, ou), rel(gp, op)
        print(f"{scheme}: vjp rel err u {eu:.2e} phi {ep:.2e}", flush=True)
        assert eu < 1e-11 and ep < 1e-11


def test_adjoint_gradient_matches_the_oracle_and_finite_differences(g):
    """The checkpointed device gradient of J = k + <phi^2> after 12 forced IF-RK4 steps: equal to the
    oracle's to 1e-10, and to fourth-order central differences of device runs along 4 random
    solenoidal directions to 1e-6."""
    n, dt, steps, every = 32, 2e-3, 12, 5
    sim, u, p, oc = make_sim(g, n, np.float64, forcing=0.1, nu=0.02, seed=43)
    d = dns.Domain(n)
    with sim:
        J, gu, gp = sim.adjoint_gradient(dt, steps, every)
        Jo, ou, op = oadj.gradient(d, oc, "rk4", u, p, dt, steps, every)
        assert J == pytest.approx(Jo, rel=1e-12)
        print(f"adjoint gradient vs oracle: u {rel(gu, ou):.2e} phi {rel(gp, op):.2e}", flush=True)
        assert rel(gu, ou) < 1e-10 and rel(gp, op) < 1e-10
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float64)
    cond = g.FlowConditions(viscosity=0.02, schmidt=1.0,
                            velocity_forcing=g.ConstantPower(dom, injection_rate=0.1, lowest_shell=1, highest_shell=2))

    def J_of(u0):
        with g.Simulation(dom, conditions=cond, state=g.ResolvedFields(u0, p), mean_gradient=1.0) as s:
            for _ in range(steps):
                s.step_fields(dt)
            r = s.reductions()
            return r["kinetic_energy"] + r["scalar_variance"]

    for k in range(4):
        du = oadj.solenoidal_direction(d, 60 + k)
        e = 0.1   # |delta| = 1 against |u_hat| ~ N^3: the linear regime
        fd = (-J_of(u + 2 * e * du) + 8 * J_of(u + e * du) - 8 * J_of(u - e * du) + J_of(u - 2 * e * du)) / (12 * e)
        an = oadj.inner(d, gu, du)
        print(f"direction {k}: adjoint {an:.12e} fd {fd:.12e} rel {abs(an - fd) / abs(fd):.2e}", flush=True)
        assert abs(an - fd) <= 1e-6 * abs(fd)


def test_fp32_bulk_adjoint_gradient_against_the_fp64_oracle(g):
    n, dt, steps = 32, 2e-3, 6
    sim, u, p, oc = make_sim(g, n, np.float32, forcing=0.1, nu=0.02, seed=47)
    d = dns.Domain(n)
    with sim:
        J, gu, gp = sim.adjoint_gradient(dt, steps, 3)
    Jo, ou, op = oadj.gradient(d, oc, "rk4", u, p, dt, steps, 3)
    print(f"FP32-bulk adjoint: J {abs(J - Jo) / Jo:.2e}, grad u {rel(gu, ou):.2e}, phi {rel(gp, op):.2e}", flush=True)
    assert J == pytest.approx(Jo, rel=1e-5)
    assert rel(gu, ou) < 1e-4 and rel(gp, op) < 1e-4


def test_adjoint_gradient_equals_the_dual_derivative_of_the_oracle(g):
    """<grad, v> of the device gradient against the oracle's forward-mode derivative along v, at
    32^3 over 20 forced steps with a ragged last checkpoint segment (6, 6, 6, 2)."""
    n, dt, steps, every = 32, 2e-3, 20, 6
    sim, u, p, oc = make_sim(g, n, np.float64, forcing=0.1, nu=0.02, seed=53)
    d = dns.Domain(n)
    with sim:
        J, gu, gp = sim.adjoint_gradient(dt, steps, every)
    rng = np.random.default_rng(54)
    for k in range(4):
        du = oadj.solenoidal_direction(d, 70 + k)
        dp = None
        if k % 2:
            dp = dns.rfft3(rng.standard_normal((n, n, n))) * d.mask
            dp /= math.sqrt(oadj.inner(d, dp, dp))
        Jd, dJ = dual.directional_derivative(d, oc, "rk4", u, p, dt, steps, du, dp)
        an = oadj.inner(d, gu, du) + (oadj.inner(d, gp, dp) if dp is not None else 0.0)
        print(f"direction {k}: adjoint {an:.12e} dual {dJ:.12e} rel {abs(an - dJ) / abs(dJ):.2e}", flush=True)
        assert J == pytest.approx(Jd, rel=1e-12)
        assert abs(an - dJ) <= 1e-10 * abs(dJ)


def _forced_run(g, n, dtype, seed=2031):
    """Forced isotropic turbulence: E(k) ~ k^4 exp(-2 (k/4)^2) with k = 0.5, ConstantPower eps = 0.1
    on shells 1..2, nu for k_max eta = 1.5, a scalar with mean gradient G = 1 and no initial
    fluctuation."""
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=dtype)
    cond = g.FlowConditions(viscosity=g.predicted_viscosity(dom, 0.1, kmax_eta=1.5), schmidt=1.0,
                            velocity_forcing=g.ConstantPower(dom, injection_rate=0.1, lowest_shell=1, highest_shell=2))
    u = g.random_phase_velocity(n, seed=seed)
    p = np.zeros((n, n, n // 2 + 1), dtype=np.complex128)
    return g.Simulation(dom, conditions=cond, state=g.ResolvedFields(u, p), mean_gradient=1.0), u, p


def test_adjoint_at_128_over_200_steps_against_central_differences(g):
    """The adjoint workload: 128^3, 200 forced IF-RK4 steps, checkpoints every 20.  FP64: the
    gradient against fourth-order central differences of device runs along 16 seeded solenoidal
    directions, each within 1e-6 relative.  FP32-bulk: the gradient against the FP64 one."""
    n, dt, steps, every = 128, 2e-3, 200, 20
    d = dns.Domain(n)
    sim, u, p = _forced_run(g, n, np.float64)
    with sim:
        t0 = time.perf_counter()
        J, gu, gp = sim.adjoint_gradient(dt, steps, every)
        print(f"FP64 gradient: J {J:.15e} in {time.perf_counter() - t0:.2f} s", flush=True)

        def J_of(u0):
            sim.set_fields(g.ResolvedFields(u0, p))
            for _ in range(steps):
                sim.step_fields(dt, graph=True)
            r = sim.reductions()
            return r["kinetic_energy"] + r["scalar_variance"]

        assert J_of(u) == pytest.approx(J, rel=1e-13)
        # |e v|_w = 200 against |u_hat|_w ~ N^3: a 1e-4 relative perturbation, far above the
        # trajectory's round-off and deep in the range where the O(e^4) truncation is negligible
        e = 200.0
        worst = 0.0
        for k in range(16):
            du = oadj.solenoidal_direction(d, 1000 + k)
            jp2, jp1, jm1, jm2 = (J_of(u + s * e * du) for s in (2, 1, -1, -2))
            fd4 = (-jp2 + 8 * jp1 - 8 * jm1 + jm2) / (12 * e)
            fd2 = (jp1 - jm1) / (2 * e)
            an = oadj.inner(d, gu, du)
            r4 = abs(an - fd4) / abs(fd4)
            worst = max(worst, r4)
            print(f"direction {k:2d}: adjoint {an:.12e} fd4 {fd4:.12e} rel {r4:.2e} (fd2 rel "
                  f"{abs(an - fd2) / abs(fd2):.2e})", flush=True)
        print(f"worst of 16: {worst:.2e}", flush=True)
        assert worst <= 1e-6
    sim32, _, _ = _forced_run(g, n, np.float32)
    with sim32:
        t0 = time.perf_counter()
        J32, gu32, gp32 = sim32.adjoint_gradient(dt, steps, every)
        print(f"FP32-bulk gradient in {time.perf_counter() - t0:.2f} s: J rel {abs(J32 - J) / J:.2e}, grad u rel "
              f"{rel(gu32, gu):.2e}, grad phi rel {rel(gp32, gp):.2e}", flush=True)
    assert J32 == pytest.approx(J, rel=1e-5)
    assert rel(gu32, gu) < 1e-3


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_revolve_checkpointing_gives_the_same_bits(g, dtype):
    """Binomial (revolve) checkpointing with 1, 2, 3 and 13 state slots over 12 forced steps gives
    the uniformly checkpointed gradient and objective bit for bit (the same kernels on the same
    states, only the recomputation order differs), and leaves the run at its initial state."""
    n, dt, steps = 32, 2e-3, 12
    sim, u, p, oc = make_sim(g, n, dtype, forcing=0.1, nu=0.02, seed=59)
    with sim:
        before = sim.state().velocity_spectral
        J, gu, gp = sim.adjoint_gradient(dt, steps, 5)
        for slots in (1, 2, 3, 13):
            Jr, gur, gpr = sim.adjoint_gradient(dt, steps, snapshots=slots)
            print(f"{np.dtype(dtype).name} revolve with {slots} slots: J {Jr!r} (uniform {J!r})", flush=True)
            assert Jr == J
            assert np.array_equal(gur, gu) and np.array_equal(gpr, gp)
        assert np.array_equal(sim.state().velocity_spectral, before)
        with pytest.raises(ValueError, match="snapshots must be at least 1"):
            sim.adjoint_gradient(dt, steps, snapshots=0)
