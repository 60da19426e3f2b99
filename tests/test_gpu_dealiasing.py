"""HouLi and PhaseShift dealiasing on the device against the FP64 oracle, through the C ABI.

HouLi (src/SimulationDomain.jl:76-138,365-376) keeps every mode and filters the nonlinear terms
by exp(-s (|k|/k_Ny)^p); PhaseShift (src/SimulationDomain.jl:140-176, src/Generator.jl:502-519,
668-704) averages the products on the grid and on the grid displaced by half a cell, in the
ball |k| <= (2 sqrt 2 / 3) k_Ny.  The oracle's versions are pinned in tests/test_oracle.py (the
2N zero-padded products).  Bars: FP64 one step 1e-12; FP32-bulk 10 steps 1e-5 of the FP64 oracle.
"""
import math

import numpy as np
import pytest

from oracle import dns
from oracle import tracers as trc

from test_gpu_parity import L, rel

pytestmark = pytest.mark.gpu

KINDS = {"hou_li": lambda g: g.HouLi(), "phase_shift": lambda g: g.PhaseShift()}


def _filled(d, seed, energy=0.5):
    """A random solenoidal, Hermitian velocity and scalar filling the strategy's retained modes
    with a spectrum falling off as exp(-(|m|/(N/8))^2) (HouLi: out to Nyquist and the corners)."""
    rng = np.random.default_rng(seed)
    env = np.exp(-(d.m2 / (d.n / 8.0) ** 2)) * d.mask
    shape = (3,) + d.m2.shape
    u = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * env
    u[:, 0, 0, 0] = 0
    u = dns.reality_projection(d, dns.leray(d, u))
    p = (rng.standard_normal(d.m2.shape) + 1j * rng.standard_normal(d.m2.shape)) * env
    p[0, 0, 0] = 0
    p = dns.reality_projection(d, p[None])[0]
    e = 0.5 * dns.weighted_inner(d, u)
    return u * math.sqrt(energy / e), p * math.sqrt(energy / dns.weighted_inner(d, p))


def _sim(g, kind, n, dtype, u, p, scheme, forcing=0.1, nu=0.02, tracers=None, particle=None, dl=None):
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=dl or KINDS[kind](g), element_type=dtype)
    vf = g.ConstantPower(dom, injection_rate=forcing, lowest_shell=1, highest_shell=2) if forcing else g.NoForcing()
    return g.Simulation(dom, conditions=g.FlowConditions(viscosity=nu, schmidt=1.0, velocity_forcing=vf),
                        state=g.ResolvedFields(u, p), mean_gradient=1.0 if p is not None else 0.0,
                        field_integration_scheme=scheme, tracers=tracers, particle_integration_scheme=particle)


@pytest.mark.parametrize("kind", ["hou_li", "phase_shift"])
@pytest.mark.parametrize("n", [32, 48, 64])
@pytest.mark.parametrize("scheme", ["rk4", "ssprk3"])
def test_one_fp64_step_matches_the_oracle(g, kind, n, scheme):
    d = dns.Domain(n, dealiasing=kind)
    u, p = _filled(d, 100 + n)
    oc = dns.Conditions(viscosity=0.02, schmidt=1.0, forcing=(0.1, 1, 2), mean_gradient=1.0)
    sch = g.IFLowStorageRK4 if scheme == "rk4" else g.IFSSPRK3
    with _sim(g, kind, n, np.float64, u, p, sch) as sim:
        mx, my, mz, mask = sim.frame_tables()
        assert np.array_equal(mask.astype(bool), d.mask)
        dep = dns.deposit(d, u, p)
        assert rel(sim.physical(0), dep.u) < 1e-12 and rel(sim.physical(1), dep.phi) < 1e-12
        sim.advance(1e-3)
        uo, po, dep = dns.step(d, oc, scheme, u, p, 1e-3)
        st = sim.state()
        eu, ep = rel(st.velocity_spectral, uo), rel(st.scalar_spectral, po)
        print(f"{kind} {n}^3 {scheme}: u {eu:.2e} phi {ep:.2e}", flush=True)
        assert eu < 1e-12 and ep < 1e-12
        assert rel(sim.physical(0), dep.u) < 1e-12
        red, want = sim.reductions(), dns.reductions(d, oc, uo, po)
        for k in want:
            assert red[k] == pytest.approx(want[k], rel=1e-12), k


@pytest.mark.parametrize("kind", ["hou_li", "phase_shift"])
def test_fp32_bulk_follows_the_fp64_oracle_over_10_steps(g, kind):
    n, dt = 64, 1e-3
    d = dns.Domain(n, dealiasing=kind)
    u, p = _filled(d, 7)
    oc = dns.Conditions(viscosity=0.02, schmidt=1.0, forcing=(float(np.float32(0.1)), 1, 2), mean_gradient=1.0,
                        band_eps=float(np.finfo(np.float32).eps))
    uo, po, dep = u, p, None
    d32 = dns.Domain(n, dealiasing=kind, element_type=np.float32)   # the FP32 run's filter strength
    with _sim(g, kind, n, np.float32, u, p, g.IFLowStorageRK4) as sim:
        for s in range(1, 11):
            sim.advance(dt)
            uo, po, dep = dns.step(d32, oc, "rk4", uo, po, dt, dep)
        st = sim.state()
        eu, ep = rel(st.velocity_spectral, uo), rel(st.scalar_spectral, po)
        print(f"{kind} FP32-bulk after 10 steps: u {eu:.2e} phi {ep:.2e}", flush=True)
        assert eu < 1e-5 and ep < 1e-5


def test_hou_li_keeps_the_nyquist_planes_real_at_128(g):
    """At 128^3 under HouLi the state carries the kx = N/2 plane: after steps it is exactly
    Hermitian there as on kx = 0 (the reality projection of both self-conjugate planes)."""
    n = 128
    d = dns.Domain(n, dealiasing="hou_li")
    u, p = _filled(d, 5)
    with _sim(g, "hou_li", n, np.float64, u, p, g.IFSSPRK3) as sim:
        for _ in range(3):
            sim.advance(1e-3)
        st = sim.state()
    neg = (-np.arange(n)) % n
    for f in (st.velocity_spectral, st.scalar_spectral):
        for plane in (0, n // 2):
            s = f[..., plane]
            assert np.array_equal(s, np.conj(s[..., neg, :][..., :, neg]))
        assert np.max(np.abs(f[..., n // 2])) > 0   # the plane is populated


def test_the_published_hou_li_configuration_at_128(g):
    """docs/src/validation.md:29-39: 128^3 FP64, HouLi(order 36, strength 63.463), IFSSPRK3,
    ConstantPower eps = 1 on shells 1-2, nu = 9.032e-3, dt = 2e-3, no scalar, 10^4 AB(4) tracers.
    One step against the oracle, then five more with the tracers' Adams start-up against the
    oracle's trail."""
    n, dt, m = 128, 2e-3, 10_000
    hl = dict(filter_order=36, filter_strength=63.463)
    d = dns.Domain(n, dealiasing="hou_li", **hl)
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.HouLi(order=36, strength=63.463),
                        element_type=np.float64)
    nu = g.predicted_viscosity(dom, 1.0, kmax_eta=1.5)
    assert nu == pytest.approx(9.032e-3, rel=1e-3)
    u = g.random_phase_velocity(n, seed=2026)
    X = g.uniform_positions(m, L, 2028)
    oc = dns.Conditions(viscosity=nu, schmidt=1.0, forcing=(1.0, 1, 2))
    with _sim(g, "hou_li", n, np.float64, u, None, g.IFSSPRK3, forcing=1.0, nu=nu,
              tracers=g.Tracers(dom, positions=X), particle=g.AB(4), dl=g.HouLi(order=36, strength=63.463)) as sim:
        dep = dns.deposit(d, u, None)
        samples, times = [trc.interpolate(dep.u, X.T, L)], [0.0]
        uo, Xo = u, X.T.copy()
        for s in range(1, 7):
            sim.advance(dt)
            uo, _, dep = dns.step(d, oc, "ssprk3", uo, None, dt, dep)
            Xo = trc.adams_advect(Xo, samples, times, dep.u, dt, L, corrected=len(samples) < 4)
            samples.insert(0, trc.interpolate(dep.u, Xo, L))
            times.insert(0, s * dt)
            del samples[4:], times[4:]
            if s in (1, 6):
                eu = rel(sim.state().velocity_spectral, uo)
                dx = np.abs((sim.tracer_positions() - Xo.T + L / 2) % L - L / 2).max()
                v, _ = sim.tracer_samples()
                print(f"published HouLi configuration, step {s}: u {eu:.2e} |dX| {dx:.2e} v {rel(v, samples[0].T):.2e}",
                      flush=True)
                assert eu < 1e-12 and dx < 1e-12 and rel(v, samples[0].T) < 1e-12


def test_phase_shift_refuses_a_state_outside_its_ball(g):
    n = 32
    d = dns.Domain(n, dealiasing="phase_shift")
    u, p = _filled(d, 3)
    u[0, 0, 0, n // 2] = 1.0   # the axis Nyquist mode lies outside sqrt(2) N / 3
    with pytest.raises(ValueError, match="phase-shift dealiasing ball"):
        _sim(g, "phase_shift", n, np.float64, u, p, g.IFLowStorageRK4)


@pytest.mark.parametrize("kind", ["hou_li", "phase_shift"])
def test_ensemble_boxes_under_each_strategy_equal_their_standalone_runs(g, kind):
    members = []
    for b, n in enumerate((32, 48)):
        d = dns.Domain(n, dealiasing=kind)
        u, p = _filled(d, 40 + b)
        dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=KINDS[kind](g), element_type=np.float32)
        members.append(g.EnsembleMember(dom, g.FlowConditions(viscosity=0.01 * (b + 1), schmidt=1.0),
                                        g.ResolvedFields(u, p)))
    with g.Ensemble(members, field_integration_scheme=g.IFLowStorageRK3, mean_gradient=1.0) as ens:
        for s in range(3):
            ens.step(1e-3, graph=s != 1)
        for b, mb in enumerate(members):
            with g.Simulation(mb.domain, conditions=mb.conditions, state=mb.state,
                              field_integration_scheme=g.IFLowStorageRK3, mean_gradient=1.0) as alone:
                for _ in range(3):
                    alone.advance(1e-3)
                assert np.array_equal(alone.state().velocity_spectral, ens.box(b).state().velocity_spectral), b


def test_the_diagnostics_built_for_two_thirds_refuse_the_wide_strategies(g):
    n = 32
    d = dns.Domain(n, dealiasing="hou_li")
    u, p = _filled(d, 9)
    with _sim(g, "hou_li", n, np.float64, u, p, g.IFLowStorageRK4) as sim:
        with pytest.raises(NotImplementedError, match="TwoThirds"):
            sim.adjoint_gradient(1e-3, 2, 1)
        with pytest.raises(NotImplementedError, match="TwoThirds"):
            sim.derivative_skewness()
        k, e = sim.energy_spectrum()
        assert np.sum(e) * (2 * math.pi / L) * 2 == pytest.approx(2 * sim.reductions()["kinetic_energy"], rel=1e-12)


@pytest.mark.parametrize("kind", ["hou_li", "phase_shift"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_the_autotuner_serves_the_wide_strategies(g, kind, dtype):
    """The autotuner times every schedule of a HouLi or PhaseShift run, whose z forward kernels read
    the filter table or the displaced products, leaves its state bitwise unchanged, and the tuned run
    (IFSSPRK3 with AB(4) tracers, the published configuration's schemes) follows the untuned one to
    the rounding of the schedules."""
    n, dt = 32, 2e-3
    d = dns.Domain(n, dealiasing=kind)
    u, p = _filled(d, 71)
    runs = []
    for tune in (True, False):
        tr = g.Tracers(g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=KINDS[kind](g),
                                     element_type=dtype), number_of_particles=500, seed=13)
        sim = _sim(g, kind, n, dtype, u, p, g.IFSSPRK3, forcing=1.0, tracers=tr, particle=g.AB(4))
        sim.advance(dt)
        if tune:
            before = sim.state()
            sim.autotune(reps=1)
            after = sim.state()
            assert np.array_equal(before.velocity_spectral, after.velocity_spectral)
            assert np.array_equal(before.scalar_spectral, after.scalar_spectral)
        for _ in range(3):
            sim.advance(dt)
        runs.append((sim.state(), sim.tracer_positions()))
        sim.close()
    (sa, xa), (sb, xb) = runs
    bar = 1e-12 if dtype == np.float64 else 1e-5
    print(f"{kind} {np.dtype(dtype).name}: tuned against untuned u {rel(sa.velocity_spectral, sb.velocity_spectral):.2e}",
          flush=True)
    assert rel(sa.velocity_spectral, sb.velocity_spectral) < bar
    assert rel(sa.scalar_spectral, sb.scalar_spectral) < bar
    assert np.abs((xa - xb + L / 2) % L - L / 2).max() < (1e-12 if dtype == np.float64 else 1e-4)
