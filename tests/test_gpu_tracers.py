"""The multistep tracer schemes on the device: AB(k), ABM(k) and Heun = ABM(1) over the run's
velocity trail, the trail itself, its time derivative and the start-up bound.

The oracle's Adams step (oracle/tracers.py) follows src/ParticlePropagator.jl:287-372 and the
trail src/Particles.jl:404-631.  Bars: positions within 1e-11 (FP64) and 5e-4 (FP32) after 10 steps of
uneven length, trail samples to 1e-11 / 5e-5 relative; the order of accuracy on a trajectory known
in closed form.
"""
import math

import numpy as np
import pytest

from oracle import dns
from oracle import tracers as trc

pytestmark = pytest.mark.gpu
L = 2 * math.pi


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _scheme(g, kind, k):
    return g.AB(k) if kind == "ab" else g.ABM(k)


@pytest.mark.parametrize("kind,k", [("ab", 4), ("ab", 2), ("abm", 3), ("abm", 1)])
@pytest.mark.parametrize("dtype,tol,stol", [(np.float64, 1e-11, 1e-11), (np.float32, 5e-4, 5e-5)])
def test_multistep_tracers_match_the_oracle(g, kind, k, dtype, tol, stol):
    """A forced 32^3 flow with 400 tracers, 10 steps of uneven dt: positions, the trail's samples
    and times, and the Lagrangian acceleration against the oracle's loop."""
    n, m = 32, 400
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=dtype)
    pos = trc.uniform_positions(m, L, 11).T.astype(dtype).astype(np.float64)
    u, p = g.random_phase_velocity(n, seed=51), g.random_phase_scalar(n, seed=52)
    vf = g.ConstantPower(dom, injection_rate=0.1, lowest_shell=1, highest_shell=2)
    cond = g.FlowConditions(viscosity=0.02, schmidt=1.0, velocity_forcing=vf)
    oc = dns.Conditions(viscosity=0.02, schmidt=1.0, forcing=(float(dtype(0.1)), 1, 2), mean_gradient=1.0,
                        band_eps=float(np.finfo(dtype).eps))
    d = dns.Domain(n)
    cap = k if kind == "ab" else k + 1
    dts = [2e-3 * (1 + 0.3 * math.sin(i)) for i in range(10)]
    with g.Simulation(dom, conditions=cond, state=g.ResolvedFields(u, p), mean_gradient=1.0,
                      tracers=g.Tracers(dom, positions=pos), particle_integration_scheme=_scheme(g, kind, k)) as sim:
        uo, po, dep = u, p, dns.deposit(d, u, p)
        X = pos.T.copy()
        samples, times = [trc.interpolate(dep.u, X, L)], [0.0]
        for step, dt in enumerate(dts, 1):
            uo, po, dep = dns.step(d, oc, "rk4", uo, po, dt, dep)
            full = len(samples) >= k
            J = k if full else len(samples)
            X = trc.adams_advect(X, samples[:J], times[:J], dep.u, dt, L, corrected=(not full) or kind == "abm")
            samples = [trc.interpolate(dep.u, X, L)] + samples[: cap - 1]
            times = [times[0] + dt] + times[: cap - 1]
            sim.advance(dt)
            if step in (1, k, 10):
                got = sim.tracer_positions()
                err = np.max(np.abs(got - X.T))
                print(f"{kind}({k}) {np.dtype(dtype).name} step {step}: positions {err:.2e}", flush=True)
                assert err < tol
        t_dev = sim.trail_times()
        assert len(t_dev) == len(times) == cap
        assert np.allclose(t_dev, times, rtol=0, atol=1e-15)
        for off in range(cap):
            assert rel(sim.trail_sample(-off), samples[off].T) < stol
        if cap > 1:
            acc = sim.time_derivative(1)
            want = trc.time_derivative(samples, times, 1)
            assert rel(acc, want.T) < (1e-9 if dtype is np.float64 else 0.05)


@pytest.mark.parametrize("kind,k,order", [("ab", 2, 2), ("ab", 3, 3), ("abm", 2, 3)])
def test_multistep_order_on_the_decaying_shear_wave(g, kind, k, order):
    """u_x = A exp(-nu t) sin y is an exact solution the IF step keeps exactly, and a tracer at a grid
    value of y (where Catmull-Rom is exact) moves as x(t) = x0 + A sin y0 (1 - exp(-nu t)) / nu:
    the error falls as dt^order (the start-up steps ABM(1), ABM(2) leave an O(dt^3) part)."""
    n, nu, A, T = 32, 1.0, 0.8, 0.5
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float64)
    x = np.arange(n) * L / n
    uu = np.zeros((3, n, n, n))
    uu[0] = A * np.sin(x)[None, :, None]
    uh = g.truncate_to_band(dns.rfft3(uu), dom)
    y0 = 5 * L / n
    pos = np.array([[0.3, 1.1, 2.9], [y0, y0, y0], [0.5, 1.0, 1.5]])
    want = pos[0] + A * math.sin(y0) * (1 - math.exp(-nu * T)) / nu
    errs = []
    for steps in (10, 20, 40):
        with g.Simulation(dom, conditions=g.FlowConditions(viscosity=nu, schmidt=1.0), state=g.ResolvedFields(uh, None),
                          tracers=g.Tracers(dom, positions=pos), particle_integration_scheme=_scheme(g, kind, k)) as sim:
            for _ in range(steps):
                sim.advance(T / steps)
            got = sim.tracer_positions()
        errs.append(float(np.max(np.abs(got[0] - want))))
        assert np.max(np.abs(got[1] - y0)) <= 1e-14
    print(f"{kind}({k}) errors {errs}", flush=True)
    r = 2.0 ** order
    assert 0.7 * r < errs[0] / errs[1] < 1.4 * r and 0.7 * r < errs[1] / errs[2] < 1.4 * r


@pytest.mark.parametrize("kind,k", [("ab", 3), ("abm", 1)])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-11), (np.float32, 5e-4)])
def test_unwrapped_positions_keep_the_winding(g, kind, k, dtype, tol):
    """src/ParticlePropagator.jl:299,313 and src/Particles.jl:171-175: the unwrapped record starts
    at the positions and takes every push's terms without the wrap.  A mean flow (+2 along x, -2
    along z) carries tracers seeded by the faces across them; the record matches the oracle's, its
    winding is a whole number of boxes, and reduced to the box it is the wrapped position."""
    n, m = 32, 200
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=dtype)
    rng = np.random.default_rng(3)
    pos = np.stack([rng.uniform(L - 0.15, L, m), rng.uniform(0, L, m), rng.uniform(0, 0.15, m)])
    pos = pos.astype(dtype).astype(np.float64)
    u = g.random_phase_velocity(n, seed=53)
    u[0, 0, 0, 0] += 2.0 * n ** 3
    u[2, 0, 0, 0] -= 2.0 * n ** 3
    oc = dns.Conditions(viscosity=0.02, schmidt=1.0)
    d = dns.Domain(n)
    cap = k if kind == "ab" else k + 1
    dts = [1e-2 * (1 + 0.2 * math.cos(i)) for i in range(10)]
    with g.Simulation(dom, conditions=g.FlowConditions(viscosity=0.02, schmidt=1.0), state=g.ResolvedFields(u, None),
                      tracers=g.Tracers(dom, positions=pos, unwrapped=True),
                      particle_integration_scheme=_scheme(g, kind, k)) as sim:
        assert np.array_equal(sim.unwrapped_positions(), sim.tracer_positions())
        uo, dep = u, dns.deposit(d, u, None)
        X, U = pos.T.copy(), pos.T.copy()
        samples, times = [trc.interpolate(dep.u, X, L)], [0.0]
        for dt in dts:
            uo, _, dep = dns.step(d, oc, "rk4", uo, None, dt, dep)
            full = len(samples) >= k
            J = k if full else len(samples)
            X, U = trc.adams_advect(X, samples[:J], times[:J], dep.u, dt, L,
                                    corrected=(not full) or kind == "abm", unwrapped=U)
            samples = [trc.interpolate(dep.u, X, L)] + samples[: cap - 1]
            times = [times[0] + dt] + times[: cap - 1]
            sim.advance(dt)
        x, uw = sim.tracer_positions(), g.unwrapped_positions(sim)
    crossed = int(np.count_nonzero((uw < 0) | (uw >= L)))
    print(f"{kind}({k}) {np.dtype(dtype).name}: unwrapped {np.max(np.abs(uw - U.T)):.2e}, "
          f"{crossed} crossings", flush=True)
    assert crossed > m                                   # most tracers crossed the x and the z faces
    assert np.max(np.abs(uw - U.T)) < tol
    assert np.max(np.abs(x - X.T)) < tol
    winding = (uw - x) / L
    assert np.max(np.abs(winding - np.round(winding))) < (1e-12 if dtype is np.float64 else 1e-5)
    assert np.all(np.round(winding[0]) >= 0) and np.all(np.round(winding[2]) <= 0)


def test_unwrapped_positions_are_opt_in(g):
    n = 16
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float64)
    with g.Simulation(dom, conditions=g.FlowConditions(viscosity=0.02, schmidt=1.0),
                      state=g.ResolvedFields(g.random_phase_velocity(n, seed=1), None),
                      tracers=g.Tracers(dom, number_of_particles=10)) as sim:
        with pytest.raises(ValueError, match="no unwrapped positions"):
            g.unwrapped_positions(sim)
        buf = np.zeros((3, 10))
        with pytest.raises(Exception, match="no unwrapped positions"):   # the C-ABI refuses it too
            sim._call("hdns_tracers_unwrapped_download", g._lib.dptr(buf))


def test_shortened_first_steps_start_small_and_grow(g):
    """src/SimulationRun.jl:1099-1129 under AB(4): the first step is tau (dt/tau)^(4/3) with tau the
    Kolmogorov time of the initial state; each later one at most twice the last, up to the CFL step."""
    n = 32
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float64)
    cond = g.FlowConditions(viscosity=0.02, schmidt=1.0)
    with g.Simulation(dom, conditions=cond, state=g.ResolvedFields(g.random_phase_velocity(n, seed=61), None),
                      tracers=g.Tracers(dom, number_of_particles=64), particle_integration_scheme=g.AB(4),
                      shortened_first_steps=True) as sim:
        eps = sim.reductions()["dissipation"]
        tau = math.sqrt(0.02 / eps)
        dts = []
        for _ in range(8):
            plain = g.cfl_bound(sim.max_speed(), dom.grid_spacing, sim.cfl, 0.02, L)
            dt = sim.cfl_timestep()
            if not dts:
                assert dt == pytest.approx(tau * (plain / tau) ** (4 / 3), rel=1e-12)
                assert dt < plain
            else:
                assert dt == pytest.approx(min(2 * dts[-1], plain), rel=1e-12)
            dts.append(dt)
            sim.advance()
        assert dts[-1] == pytest.approx(g.cfl_bound(sim.max_speed(), dom.grid_spacing, sim.cfl, 0.02, L), rel=0.2)


def test_trail_guards(g):
    n = 16
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float64)
    with g.Simulation(dom, conditions=g.FlowConditions(viscosity=0.02, schmidt=1.0),
                      state=g.ResolvedFields(g.random_phase_velocity(n, seed=1), None),
                      tracers=g.Tracers(dom, number_of_particles=10), particle_integration_scheme=g.AB(3)) as sim:
        assert sim.trail_times() == [0.0]
        with pytest.raises(Exception, match="order 1 needs at least 2"):
            sim.time_derivative(1)
        with pytest.raises(Exception, match="trail"):
            sim.trail_sample(-1)
        sim.step_fields(1e-3)
        sim.advect(1e-3)
        with pytest.raises(Exception):   # pushed twice without a sample in between
            sim._call("hdns_advect", 1e-3)


def _tracked(g, mode, scheme, dtype, steps=6, m=300, n=16):
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=dtype)
    u, p = g.random_phase_velocity(n, seed=3), g.random_phase_scalar(n, seed=4)
    sim = g.Simulation(dom, conditions=g.FlowConditions(viscosity=0.05, schmidt=1.0), state=g.ResolvedFields(u, p),
                       mean_gradient=1.0, tracers=g.Tracers(dom, number_of_particles=m, seed=5, unwrapped=True),
                       particle_integration_scheme=scheme, traversal_mode=mode)
    for _ in range(steps):
        sim.advance(0.01)
    return sim


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_the_fused_push_and_sample_is_bitwise_the_two_calls(g, dtype):
    """advance() pushes and samples in one launch (hdns_advect_deposit); a run that makes the two
    calls itself (step_fields, rebuild, advect, deposit_tracers), and one whose gathers run on the
    capped launch (128-thread blocks, two per CU), are bitwise the same in positions, unwrapped
    positions, every trail sample and time, and the scalar samples.  AB(3) fills its ring of three,
    so the sample's slot is one the push reads; 1000 tracers make several blocks."""
    n, m, dt = 32, 1000, 0.01
    for scheme in (g.Heun(), g.AB(3), g.ABM(2)):
        runs = []
        for fused, capped in ((True, 0), (False, 0), (True, 1)):
            dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=dtype)
            u, p = g.random_phase_velocity(n, seed=3), g.random_phase_scalar(n, seed=4)
            sim = g.Simulation(dom, conditions=g.FlowConditions(viscosity=0.05, schmidt=1.0),
                               state=g.ResolvedFields(u, p), mean_gradient=1.0,
                               tracers=g.Tracers(dom, number_of_particles=m, seed=5, unwrapped=True),
                               particle_integration_scheme=scheme)
            assert sim.tracer_schedule(capped) == capped
            for i in range(7):
                if fused:
                    sim.advance(dt)
                else:
                    sim.step_fields(dt)
                    sim.rebuild_traversal()
                    sim.advect(dt)
                    sim.deposit_tracers((i + 1) * dt)
            runs.append(sim)
        a = runs[0]
        for b in runs[1:]:
            assert np.array_equal(a.tracer_positions(), b.tracer_positions())
            assert np.array_equal(a.unwrapped_positions(), b.unwrapped_positions())
            assert a.trail_times() == b.trail_times()
            for j in range(len(a.trail_times())):
                assert np.array_equal(a.trail_sample(-j), b.trail_sample(-j))
            (va, sa), (vb, sb) = a.tracer_samples(), b.tracer_samples()
            assert np.array_equal(va, vb) and np.array_equal(sa, sb)
        for r in runs:
            r.close()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_the_traversal_order_is_a_value_preserving_accelerator(g, dtype):
    """After test/particles.jl:389-424: a run with the traversal order, rebuilt every iteration or
    left stale under manual, is bitwise the run without it (positions, unwrapped positions, every
    trail sample and time); a rebuilt order equals a freshly built one from the same positions and
    sorts the tracers by tile, and within a tile by index."""
    n, m = 16, 300
    for scheme in (g.Heun(), g.AB(3), g.ABM(3)):
        off = _tracked(g, "none", scheme, dtype)
        with off:
            with pytest.raises(ValueError, match="traversal_mode = :none"):
                off.rebuild_traversal()
            x_off, w_off, t_off = off.tracer_positions(), off.unwrapped_positions(), off.trail_times()
            trail_off = [off.trail_sample(-j) for j in range(len(t_off))]
            for mode in ("automatic", "manual"):
                with _tracked(g, mode, scheme, dtype) as on:
                    assert np.array_equal(on.tracer_positions(), x_off)
                    assert np.array_equal(on.unwrapped_positions(), w_off)
                    assert on.trail_times() == t_off
                    for j, v in enumerate(trail_off):
                        assert np.array_equal(on.trail_sample(-j), v)
    with _tracked(g, "manual", g.Heun(), dtype) as man:
        man.rebuild_traversal()
        rebuilt = man.traversal_order()
        x = man.tracer_positions()
        dom = man.domain
        with g.Simulation(dom, conditions=man.conditions, state=man.state(), mean_gradient=1.0,
                          tracers=g.Tracers(dom, positions=x)) as fresh:
            assert np.array_equal(fresh.tracer_positions(), x)
            assert np.array_equal(rebuilt, fresh.traversal_order())
        assert np.array_equal(np.sort(rebuilt), np.arange(m))
        keys = trc.cell_keys(x.T, n, L, dtype)
        i, j, k = keys % n, keys // n % n, keys // (n * n)
        tiles = (k // 8 * (n // 8) + j // 8) * (n // 8) + i // 8
        assert np.array_equal(rebuilt, np.argsort(tiles, kind="stable"))
