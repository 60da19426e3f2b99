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


def test_advance_refuses_a_parted_run_before_anything_moves(g):
    """src/SimulationRun.jl:1177-1182 (trail_between_iterations, src/Cache.jl:342-349): after a
    field step taken alone, or a push never closed by a deposit, advance() raises before the fields
    move: the state, the positions and the clock are left bit for bit as they were."""
    n, dt = 32, 2e-3
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float64)
    for parting in ("fields", "push"):
        sim, *_ = make_sim(g, n, np.float64, tracers=g.Tracers(dom, number_of_particles=64, seed=9))
        with sim:
            sim.advance(dt)
            sim.step_fields(dt)
            if parting == "push":
                sim.advect(dt)
            st0, x0 = sim.state(), sim.tracer_positions()
            match = "taken 1 step" if parting == "fields" else "never taken"
            with pytest.raises(RuntimeError, match=match):
                sim.advance(dt)
            st1, x1 = sim.state(), sim.tracer_positions()
            assert np.array_equal(st0.velocity_spectral, st1.velocity_spectral)
            assert np.array_equal(st0.scalar_spectral, st1.scalar_spectral)
            assert np.array_equal(x0, x1)
            assert sim.step == 1 and sim.time == dt


@pytest.mark.parametrize("dtype,band_tol", [(np.float64, 1e-12), (np.float32, 2e-6)])
def test_velocity_slice_heun_with_band_rescaling_and_an_inert_scalar(g, dtype, band_tol):
    """test/velocity_slice.jl: forced 32^3 HIT by IF-Heun, 300 steps at cfl 0.5, VarianceRescaling
    holding the band's kinetic energy at 0.6 (shells 1-2), a zero passive scalar.  The energy stays
    finite and bounded with no drift between the second and third hundred steps, the state stays
    solenoidal, the scalar stays exactly zero, and the band holds its target after every step."""
    n = 32
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=2 * math.pi, dealiasing=g.TwoThirds(), element_type=dtype)
    vf = g.VarianceRescaling(dom, target=0.6, over="band", lowest_shell=1, highest_shell=2)
    u = g.random_phase_velocity(n, energy=0.6, peak=2.0, seed=42)
    p = np.zeros((n, n, n // 2 + 1), dtype=np.complex128)
    energies, worst_band = [], 0.0
    with g.Simulation(dom, conditions=g.FlowConditions(viscosity=5e-2, schmidt=1.0, velocity_forcing=vf),
                      state=g.ResolvedFields(u, p), field_integration_scheme=g.IFHeun) as sim:
        dx = dom.grid_spacing
        for _ in range(300):
            r = sim.reductions()
            energies.append(r["kinetic_energy"])
            sim.advance(0.5 * dx / max(sim.max_speed(), 1e-6))
            worst_band = max(worst_band, abs(sim.reductions()["band_energy"] / 0.6 - 1))
        st = sim.state()
    assert all(math.isfinite(e) for e in energies)
    assert 0.1 < energies[-1] < 10.0
    early, late = float(np.mean(energies[100:200])), float(np.mean(energies[200:300]))
    assert abs(late - early) / early < 1.0
    m = np.fft.fftfreq(n, 1.0 / n)
    kz, ky, kx = m[:, None, None], m[None, :, None], np.arange(n // 2 + 1)[None, None, :]
    uh = st.velocity_spectral
    div = kx * uh[0] + ky * uh[1] + kz * uh[2]
    assert np.linalg.norm(div) / np.linalg.norm(uh) < 1e-8 * (1 if dtype is np.float64 else 100)
    assert not np.any(st.scalar_spectral)
    print(f"{np.dtype(dtype).name}: k {energies[0]:.4f} -> {energies[-1]:.4f}, band worst {worst_band:.2e}", flush=True)
    assert worst_band < band_tol


def test_an_unforced_decaying_run_keeps_a_bounded_step_and_a_meaningful_clock(g):
    """test/regressions/02-r5-free-modules-02.jl: as an unforced flow decays, max|u| falls and the
    advective bound grows without limit, but the resolution bound 1 / (nu kappa_min^2) caps the
    step; the clock stays finite and strictly increasing, and the energy never grows."""
    n, nu = 16, 0.05
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=2 * math.pi, dealiasing=g.TwoThirds(),
                        element_type=np.float64)
    cap = 1.0 / nu
    with g.Simulation(dom, conditions=g.FlowConditions(viscosity=nu, schmidt=1.0),
                      state=g.ResolvedFields(g.random_phase_velocity(n, seed=7), None),
                      field_integration_scheme=g.IFLowStorageRK3) as sim:
        energies, times, steps = [sim.reductions()["kinetic_energy"]], [sim.time], []
        for _ in range(400):
            dt = sim.cfl_timestep()
            steps.append(dt)
            sim.advance(dt)
            energies.append(sim.reductions()["kinetic_energy"])
            times.append(sim.time)
    print(f"k {energies[0]:.3e} -> {energies[-1]:.3e}, dt {steps[0]:.3e} -> {steps[-1]:.3e}, t {times[-1]:.3f}",
          flush=True)
    assert all(0 < s <= cap * (1 + 1e-15) for s in steps)
    assert all(math.isfinite(t) for t in times) and all(b > a for a, b in zip(times, times[1:]))
    assert all(b <= a * (1 + 1e-12) for a, b in zip(energies, energies[1:]))
    assert energies[-1] < 1e-3 * energies[0]
    assert steps[-1] == pytest.approx(cap, rel=1e-12)     # the resolution bound has taken over


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_an_adaptive_run_replays_one_captured_step_bitwise(g, dtype):
    """advance() with the CFL step replays the captured step (one capture for every dt, the dt
    written on the device before each replay): 15 adaptive iterations equal the kernel-by-kernel
    run bit for bit, in state, tracers and clock, under IF-RK4 and IF-SSPRK3 (whose substeps are
    dt/2)."""
    n = 32
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=dtype)
    for scheme in (g.IFLowStorageRK4, g.IFSSPRK3):
        runs = []
        for graph in (True, False):
            sim, *_ = make_sim(g, n, dtype, scheme=scheme, forcing=0.1,
                               tracers=g.Tracers(dom, number_of_particles=128, seed=3))
            sim.use_graph = graph
            runs.append(sim)
        a, b = runs
        with a, b:
            dts = []
            for _ in range(15):
                a.advance()
                b.advance()
                dts.append(a.time)
            assert len(set(np.diff([0.0] + dts))) > 5          # the step really changed
            assert a.time == b.time and a.step == b.step
            sa, sb = a.state(), b.state()
            assert np.array_equal(sa.velocity_spectral, sb.velocity_spectral)
            assert np.array_equal(sa.scalar_spectral, sb.scalar_spectral)
            assert np.array_equal(a.tracer_positions(), b.tracer_positions())
