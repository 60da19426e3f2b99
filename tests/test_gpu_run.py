"""The run loop through the C ABI on an MI355X: the adaptive step, the blow-up guards and the
epoch protocol of the tracer trail (src/SimulationRun.jl:1051-1096,1155-1229; src/Cache.jl:291-421)."""
import math

import numpy as np
import pytest

from oracle import dns

from test_gpu_parity import L, make_sim, rel

pytestmark = pytest.mark.gpu


def test_adaptive_advance_follows_the_oracle_with_the_same_step_rule(g):
    """advance() with no dt takes the CFL step from the device's max|u| each iteration; the oracle
    taking the same rule from its own deposit lands on the same times and the same state."""
    n = 32
    sim, u, p, oc = make_sim(g, n, np.float64, forcing=0.1, nu=0.02)
    d = dns.Domain(n)
    uo, po, dep = u, p, dns.deposit(d, u, p)
    t = 0.0
    with sim:
        for _ in range(5):
            dt = g.cfl_bound(dep.max_speed, L / n, sim.cfl, max(oc.viscosity, oc.scalar_diffusivity), L, sim.max_dt)
            assert sim.cfl_timestep() == pytest.approx(dt, rel=1e-12)
            sim.advance()
            uo, po, dep = dns.step(d, oc, "rk4", uo, po, dt, dep)
            t += dt
        assert sim.step == 5
        assert sim.time == pytest.approx(t, rel=1e-12)
        st = sim.state()
        assert rel(st.velocity_spectral, uo) < 1e-11
        assert rel(st.scalar_spectral, po) < 1e-11


def test_blow_up_guards_stop_the_run(g):
    """A state scaled to 1e30 overflows the FP32 products: the step's deposit reports a non-finite
    max|u|, advance raises the blow-up error after the deposit, and the CFL rule refuses to pick a
    step.  No tracers (their cell index would read a non-finite position)."""
    n = 32
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float32)
    u = g.random_phase_velocity(n, seed=5) * 1e30
    p = g.random_phase_scalar(n, seed=6)
    with g.Simulation(dom, conditions=g.FlowConditions(viscosity=0.02, schmidt=1.0),
                      state=g.ResolvedFields(u, p), mean_gradient=1.0) as sim:
        with pytest.raises(g._lib.BlowUpError, match="non-finite velocity at step 1"):
            sim.advance(1e-3)
        assert not math.isfinite(sim.max_speed())
        with pytest.raises(RuntimeError, match="blown up"):
            sim.cfl_timestep()
        assert sim.step == 0 and sim.time == 0.0   # the clock does not advance past a blow-up


def test_tracer_trail_epochs_throw_or_serve(g):
    """advect needs the fields exactly one step ahead of the positions and the trail's newest sample
    taken at those positions; a sample time must move forward."""
    n, dt = 32, 2e-3
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float64)
    sim, *_ = make_sim(g, n, np.float64, tracers=g.Tracers(dom, number_of_particles=64, seed=9))
    with sim:
        with pytest.raises(RuntimeError, match="one step ahead"):
            sim.advect(dt)                  # the fields have not moved yet
        sim.step_fields(dt)
        sim.advect(dt)
        with pytest.raises(RuntimeError, match="one step ahead"):
            sim.advect(dt)                  # the positions already moved with these fields
        with pytest.raises(ValueError, match="later than the previous"):
            sim.deposit_tracers(0.0)
        sim.deposit_tracers(dt)
        with pytest.raises(RuntimeError, match="one step ahead"):
            sim.advect(dt)                  # a fresh sample, but the fields have not stepped again
        sim.step_fields(dt)
        sim.advect(dt)
        sim.deposit_tracers(2 * dt)
        v, s = sim.tracer_samples()
        assert np.all(np.isfinite(v)) and np.all(np.isfinite(s))
