"""The spin-up observables through the C ABI on an MI355X: the derivative skewness and the integral
length against the oracle, and the warm-up driver on a forced run (src/Observables.jl:465-581,
src/SimulationRun.jl:1560-1660)."""
import math
import warnings

import numpy as np
import pytest

from oracle import dns
from oracle import observables as obs

from test_gpu_parity import L, make_sim, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [32, 48, 64])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-11), (np.float32, 2e-4)])
def test_skewness_and_integral_length_match_the_oracle(g, n, dtype, tol):
    sim, u, p, oc = make_sim(g, n, dtype, forcing=0.1)
    d = dns.Domain(n)
    with sim:
        for _ in range(20):          # away from the random-phase start, where S is near 0
            sim.advance(5e-3)
        st = sim.state()
        s_dev = sim.derivative_skewness()
        s_or = obs.derivative_skewness(d, st.velocity_spectral)
        kappa, E = dns.energy_spectrum(d, st.velocity_spectral)
        l_or = obs.integral_length(kappa, E)
        print(f"n={n} {np.dtype(dtype).name}: S device {s_dev:.10f} oracle {s_or:.10f}; L11 {sim.integral_length():.10f} "
              f"oracle {l_or:.10f}", flush=True)
        assert abs(s_dev - s_or) <= tol * max(abs(s_or), 1e-3)
        assert sim.integral_length() == pytest.approx(l_or, rel=tol)


def test_skewness_is_odd_zero_on_the_closed_form_field_and_leaves_the_run_unchanged(g):
    n, dt = 32, 2e-3
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float64)
    cond = g.FlowConditions(viscosity=0.02, schmidt=1.0)
    x = np.arange(n) * L / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    uc = g.truncate_to_band(dns.rfft3(np.stack([0.7 * np.sin(Y), -1.3 * np.sin(Z), 0.4 * np.sin(X)])), dom)
    with g.Simulation(dom, conditions=cond, state=g.ResolvedFields(uc, None)) as sim:
        assert abs(sim.derivative_skewness()) < 1e-12
        assert sim.integral_length() == pytest.approx(3 * math.pi / 4, rel=1e-12)
    a, u, p, oc = make_sim(g, n, np.float64, forcing=0.1)
    b, *_ = make_sim(g, n, np.float64, forcing=0.1)
    with a, b:
        a.advance(dt)
        b.advance(dt)
        s = a.derivative_skewness()             # reuses the pencil buffers, then redeposits
        a.advance(dt)
        b.advance(dt)
        assert np.array_equal(a.state().velocity_spectral, b.state().velocity_spectral)
        neg = a.state()
        a.set_fields(g.ResolvedFields(-neg.velocity_spectral, neg.scalar_spectral))
        b_state = b.state()
        assert s != 0
        a2 = a.derivative_skewness()
        a.set_fields(g.ResolvedFields(b_state.velocity_spectral, b_state.scalar_spectral))
        assert a2 == pytest.approx(-a.derivative_skewness(), rel=1e-12)


def test_warmup_spins_a_forced_run_up_to_a_flat_dissipation(g):
    """warmup() = run(StationarityCriterion(...)): the eddy-turnover clock advances by dt eps / k;
    the run stops at a flat dissipation (and scalar variance, forced by the mean gradient), or at
    the cap; with the skewness gate the skewness is sampled on the way."""
    n = 32
    sim, *_ = make_sim(g, n, np.float32, forcing=0.1, nu=0.02)
    with sim:
        crit = g.StationarityCriterion(min_turnovers=0.5, window_turnovers=0.5, tolerance=0.1, max_turnovers=4,
                                       gate_skewness=True, skewness_floor=0.05)
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            sim.run(crit, dt=1e-2)
        m = crit.monitor
        print(f"warm-up: {sim.step} steps, {crit.elapsed_turn:.2f} turnovers, converged {m.converged}, "
              f"{len(m.turn)} samples, last S {m.skew[-1]:.3f}", flush=True)
        assert sim.step > 0 and crit.elapsed_turn >= 0.5
        assert m.converged or crit.elapsed_turn >= 4
        assert m.gate_scalar                      # the mean gradient drives the scalar
        assert any(s != 0 for s in m.skew)
        with pytest.raises(ValueError, match="must return a Bool"):
            sim.run(lambda s: 1, dt=1e-2)


@pytest.mark.parametrize("n", [32, 48])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 2e-5)])
def test_below_scale_matches_the_oracle(g, n, dtype, tol):
    """kinetic_energy_and_dissipation_below_scale on the device (k_below_scale) against the oracle,
    every side l = L/N .. L; the bar is relative to k and eps (the small sides are differences)."""
    from oracle import observables as obs
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=2 * np.pi, dealiasing=g.TwoThirds(), element_type=dtype)
    u = g.random_phase_velocity(n, seed=17)
    with g.Simulation(dom, conditions=g.FlowConditions(viscosity=0.01, schmidt=1.0), state=g.ResolvedFields(u, None)) as sim:
        got = sim.kinetic_energy_and_dissipation_below_scale()
        st = sim.state().velocity_spectral
    d = dns.Domain(n)
    sc, kl, el = obs.below_scale(d, st, 0.01)
    assert np.allclose(got["scales"], sc, rtol=1e-15)
    assert np.max(np.abs(got["kinetic_energy"] - kl)) <= tol * kl[-1]
    assert np.max(np.abs(got["dissipation"] - el)) <= tol * el[-1]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_skewness_of_a_zero_field_is_zero(g, dtype):
    """test/regressions/04-edge-cases-04.jl: the strain-free branch gives 0, not 0/0."""
    n = 16
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=2 * np.pi, dealiasing=g.TwoThirds(), element_type=dtype)
    u = np.zeros((3, n, n, n // 2 + 1), dtype=np.complex128)
    with g.Simulation(dom, conditions=g.FlowConditions(viscosity=0.02, schmidt=1.0),
                      state=g.ResolvedFields(u, None)) as sim:
        assert sim.derivative_skewness() == 0.0
