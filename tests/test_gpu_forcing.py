"""The feedback and rescaling velocity forcings on the device, against the FP64 oracle.

VarianceController (src/Forcing.jl:722-871, :1041-1065) rides the per-stage amplitude that
ConstantPower uses, with its own numerator; VarianceRescaling (src/Forcing.jl:467-597) is the
post-step projection of the band.  Bars: FP64 to 1e-11 and FP32-bulk to 1e-5 relative after 10
steps, as for the other forcings; the held quantities exactly (to roundoff) where the mathematics
fixes them.
"""
import math

import numpy as np
import pytest

from oracle import dns

pytestmark = pytest.mark.gpu
L = 2 * math.pi


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _forcing(g, dom, dtype, kind, **kw):
    """The library's forcing and the oracle's Conditions keywords for it."""
    if kind == "controller":
        f = g.VarianceController(dom, target=kw.get("target", 1.5), gain=kw.get("gain", 1.0),
                                 turnover_time=kw.get("turnover_time", 0.8), highest_shell=kw.get("highest", 2))
        return f, {"controller": (f.target, f.gain, f.turnover_time, 1, f.highest_shell)}
    f = g.VarianceRescaling(dom, target=kw.get("target", 0.3), over=kw.get("over", "total"),
                            highest_shell=kw.get("highest", 2))
    return f, {"rescaling": (f.target, f.over, 1, f.highest_shell)}


def _run(g, n, dtype, kind, scheme="rk4", steps=10, dt=2e-3, check=(1, 10), scalar=True, **kw):
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=dtype)
    f, okw = _forcing(g, dom, dtype, kind, **kw)
    cond = g.FlowConditions(viscosity=0.02, schmidt=1.0, velocity_forcing=f)
    u = g.random_phase_velocity(n, seed=41)
    p = g.random_phase_scalar(n, seed=42) if scalar else None
    oc = dns.Conditions(viscosity=0.02, schmidt=1.0, mean_gradient=1.0 if scalar else 0.0,
                        band_eps=float(np.finfo(dtype).eps), **okw)
    sch = {"rk4": g.IFLowStorageRK4, "rk3": g.IFLowStorageRK3, "ssprk3": g.IFSSPRK3}[scheme]
    d = dns.Domain(n)
    uo, po, dep = u, p, None
    out = []
    with g.Simulation(dom, conditions=cond, state=g.ResolvedFields(u, p), mean_gradient=1.0 if scalar else 0.0,
                      field_integration_scheme=sch) as sim:
        for s in range(1, steps + 1):
            sim.advance(dt)
            uo, po, dep = dns.step(d, oc, scheme, uo, po, dt, dep)
            if s in check:
                st = sim.state()
                eu = rel(st.velocity_spectral, uo)
                ep = rel(st.scalar_spectral, po) if scalar else 0.0
                print(f"{kind} {kw} {scheme} {np.dtype(dtype).name} step {s}: u {eu:.2e} phi {ep:.2e}", flush=True)
                out.append((s, eu, ep, sim.reductions(), dns.reductions(d, oc, uo, po)))
    return out


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-11), (np.float32, 1e-5)])
@pytest.mark.parametrize("highest", [2, math.inf])
def test_variance_controller_matches_the_oracle(g, dtype, tol, highest):
    for s, eu, ep, red, ored in _run(g, 32, dtype, "controller", highest=highest):
        assert eu < tol and ep < tol
        assert red["kinetic_energy"] == pytest.approx(ored["kinetic_energy"], rel=10 * tol)


def test_variance_controller_relaxes_k_exponentially_on_the_device(g):
    """At N = 32 no product mode aliases into the ball, so the dealiased nonlinear term carries no
    energy and dk/dt = -G (k - k_inf)/t_l exactly: k(t) = k_inf + (k0 - k_inf) exp(-G t/t_l).  The
    bar is IF-RK4's own truncation error at this dt (the oracle reaches 9.5e-10; 1.6e-8 at dt 2e-3)."""
    n, dt, steps, gain, tl, target = 32, 1e-3, 100, 2.0, 0.8, 1.5
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float64)
    f = g.VarianceController(dom, target=target, gain=gain, turnover_time=tl)
    u = g.random_phase_velocity(n, seed=43)
    with g.Simulation(dom, conditions=g.FlowConditions(viscosity=0.05, schmidt=1.0, velocity_forcing=f),
                      state=g.ResolvedFields(u, None)) as sim:
        k0 = sim.reductions()["kinetic_energy"]
        for s in range(1, steps + 1):
            sim.advance(dt)
            want = target + (k0 - target) * math.exp(-gain * s * dt / tl)
            assert sim.reductions()["kinetic_energy"] == pytest.approx(want, rel=3e-9), s


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-11), (np.float32, 1e-5)])
@pytest.mark.parametrize("over,highest", [("total", 2), ("band", 2), ("total", math.inf)])
def test_variance_rescaling_matches_the_oracle(g, dtype, tol, over, highest):
    """The seeded field holds k = 0.5, 0.06 of it on shells 1-2."""
    target = {("total", 2): 0.55, ("band", 2): 0.1, ("total", math.inf): 0.3}[(over, highest)]
    for s, eu, ep, red, ored in _run(g, 32, dtype, "rescaling", over=over, highest=highest, target=target):
        assert eu < tol and ep < tol
        if over == "total":
            assert red["kinetic_energy"] == pytest.approx(target, rel=10 * tol)


def test_variance_rescaling_zeroes_the_band_when_the_rest_exceeds_the_target(g):
    """over total with k - k_band = 0.44 above the target 0.3: the band is zeroed, the closest
    reachable state (src/Forcing.jl:514-517), and stays empty."""
    for s, eu, ep, red, ored in _run(g, 32, np.float64, "rescaling", over="total", target=0.3, steps=3,
                                     check=(1, 3)):
        assert eu < 1e-11
        assert red["band_energy"] <= 1e-30


def test_variance_rescaling_holds_the_whole_field_after_every_step(g):
    """over total: k equals the target after every step, under IFSSPRK3 and without the scalar."""
    for s, eu, ep, red, ored in _run(g, 48, np.float64, "rescaling", scheme="ssprk3", steps=5,
                                     check=(1, 2, 3, 4, 5), scalar=False, target=0.52):
        assert eu < 1e-11
        assert red["kinetic_energy"] == pytest.approx(0.52, rel=1e-13)


def test_variance_rescaling_leaves_an_empty_band_alone(g):
    """The shear wave u_x = 0.5 sin 3z (shell 3) carries no nonlinear term, so shells 1-2 stay empty
    up to roundoff: the band is not divided by ~0 and the run equals the unforced run bit for bit."""
    n = 32
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float64)
    x = np.arange(n) * L / n
    uu = np.zeros((3, n, n, n))
    uu[0] = 0.5 * np.sin(3 * x)[:, None, None]
    u = g.truncate_to_band(dns.rfft3(uu), dom)
    states = []
    for f in (g.NoForcing(), g.VarianceRescaling(dom, target=0.5, over="band")):
        with g.Simulation(dom, conditions=g.FlowConditions(viscosity=0.02, schmidt=1.0, velocity_forcing=f),
                          state=g.ResolvedFields(u, None)) as sim:
            for _ in range(3):
                sim.advance(2e-3)
            states.append(sim.state().velocity_spectral)
    assert np.array_equal(states[0], states[1])


def test_variance_rescaling_at_512_in_fp32_holds_the_target(g):
    """The rescale kernels at the headline size: k at the target to FP32 roundoff after 3 steps."""
    n = 512
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float32)
    f = g.VarianceRescaling(dom, target=0.55, over="total", highest_shell=2)
    u = g.random_phase_velocity(n, seed=45)
    with g.Simulation(dom, conditions=g.FlowConditions(viscosity=2e-3, schmidt=1.0, velocity_forcing=f),
                      state=g.ResolvedFields(u, None)) as sim:
        for _ in range(3):
            sim.advance(1e-3)
            assert sim.reductions()["kinetic_energy"] == pytest.approx(0.55, rel=2e-6)


@pytest.mark.parametrize("kind", ["controller", "rescaling"])
def test_the_adjoint_refuses_the_variance_forcings(g, kind):
    n = 16
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float64)
    f, _ = _forcing(g, dom, np.float64, kind)
    with g.Simulation(dom, conditions=g.FlowConditions(viscosity=0.02, schmidt=1.0, velocity_forcing=f),
                      state=g.ResolvedFields(g.random_phase_velocity(n, seed=1), None)) as sim:
        with pytest.raises(Exception, match="NoForcing and ConstantPower"):
            sim.adjoint_gradient(1e-3, 2)


# --- scalar forcings ---------------------------------------------------------------------------

def _scalar_forcing(g, dom, kind):
    """The library's scalar forcing and the oracle's tuple for it."""
    if kind == "constant_power":
        f = g.ConstantPower(dom, injection_rate=0.05, lowest_shell=1, highest_shell=3)
        return f, ("constant_power", f.injection_rate, 1, 3)
    if kind == "linear":
        f = g.Linear(dom, rate=0.5, lowest_shell=1, highest_shell=2)
        return f, ("linear", f.rate, 1, 2)
    if kind == "controller":
        f = g.VarianceController(dom, target=0.9, gain=2.0, turnover_time=0.6, highest_shell=math.inf)
        return f, ("controller", f.target, f.gain, f.turnover_time, 1, math.inf)
    over = kind.split("_")[1]   # the seeded scalar holds <phi^2> = 0.5, 0.058 of it on shells 1-2
    f = g.VarianceRescaling(dom, target=0.1 if over == "band" else 0.55, over=over, lowest_shell=1, highest_shell=2)
    return f, ("rescaling", f.target, over, 1, 2)


@pytest.mark.parametrize("kind", ["constant_power", "linear", "controller", "rescaling_band", "rescaling_total"])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-11), (np.float32, 1e-5)])
@pytest.mark.parametrize("scheme", ["rk4", "ssprk3"])
def test_scalar_forcings_match_the_oracle(g, kind, dtype, tol, scheme):
    """Each scalar forcing (variance factor 1, diffusivity nu/Sc) beside a ConstantPower velocity
    forcing and a mean gradient, 10 steps against the oracle, checked after 1 and 10."""
    n, dt = 32, 2e-3
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=dtype)
    sf, osf = _scalar_forcing(g, dom, kind)
    vf = g.ConstantPower(dom, injection_rate=0.1)
    cond = g.FlowConditions(viscosity=0.02, schmidt=0.7, velocity_forcing=vf, scalar_forcing=sf)
    u, p = g.random_phase_velocity(n, seed=81), g.random_phase_scalar(n, seed=82)
    oc = dns.Conditions(viscosity=0.02, schmidt=0.7, mean_gradient=0.5, forcing=(vf.injection_rate, 1, 2),
                        scalar_forcing=osf, band_eps=float(np.finfo(dtype).eps))
    d = dns.Domain(n)
    sch = {"rk4": g.IFLowStorageRK4, "ssprk3": g.IFSSPRK3}[scheme]
    uo, po, dep = u, p, None
    with g.Simulation(dom, conditions=cond, state=g.ResolvedFields(u, p), mean_gradient=0.5,
                      field_integration_scheme=sch) as sim:
        for s in range(1, 11):
            sim.advance(dt)
            uo, po, dep = dns.step(d, oc, scheme, uo, po, dt, dep)
            if s in (1, 10):
                st = sim.state()
                eu, ep = rel(st.velocity_spectral, uo), rel(st.scalar_spectral, po)
                print(f"scalar {kind} {scheme} {np.dtype(dtype).name} step {s}: u {eu:.2e} phi {ep:.2e}", flush=True)
                assert eu < tol and ep < tol
        if kind == "rescaling_total":
            assert sim.reductions()["scalar_variance"] == pytest.approx(0.55, rel=1e-12 if dtype is np.float64 else 1e-5)


def test_scalar_linear_forcing_grows_its_band_exactly_on_the_device(g):
    """u = 0: phi = 0.6 (sin y + 0.5 cos 3z) evolves as exp((A - kappa) t) on shell 1 and
    exp(-9 kappa t) on shell 3, under every field scheme."""
    n, nu, sc, A, dt, steps = 32, 0.05, 0.5, 0.4, 1e-2, 20
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float64)
    x = np.arange(n) * L / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    ph = g.truncate_to_band(dns.rfft3(0.6 * (np.sin(Y) + 0.5 * np.cos(3 * Z))), dom)
    d = dns.Domain(n)
    kap = nu / sc
    want = ph * np.where(d.shell_band(1, 2), np.exp((A - kap * d.k2) * steps * dt), np.exp(-kap * d.k2 * steps * dt))
    for scheme in (g.IFLowStorageRK3, g.IFLowStorageRK4, g.IFSSPRK3, g.IFHeun):
        cond = g.FlowConditions(viscosity=nu, schmidt=sc, scalar_forcing=g.Linear(dom, rate=A))
        with g.Simulation(dom, conditions=cond, state=g.ResolvedFields(np.zeros((3,) + ph.shape, complex), ph),
                          field_integration_scheme=scheme) as sim:
            for _ in range(steps):
                sim.advance(dt)
            got = sim.state().scalar_spectral
        assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want)), scheme.__name__


def test_a_scalar_forcing_needs_the_scalar_and_the_adjoint_refuses_it(g):
    n = 16
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float64)
    cond = g.FlowConditions(viscosity=0.02, schmidt=1.0, scalar_forcing=g.ConstantPower(dom, injection_rate=0.1))
    with pytest.raises(ValueError, match="carries the scalar"):
        g.Simulation(dom, conditions=cond, state=g.ResolvedFields(g.random_phase_velocity(n, seed=1), None))
    with g.Simulation(dom, conditions=cond, state=g.ResolvedFields(g.random_phase_velocity(n, seed=1),
                                                                   g.random_phase_scalar(n, seed=2))) as sim:
        with pytest.raises(Exception, match="scalar forcing"):
            sim.adjoint_gradient(1e-3, 2)


# --- StochasticOU ------------------------------------------------------------------------------

@pytest.mark.parametrize("field", ["velocity", "scalar"])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-11), (np.float32, 1e-5)])
@pytest.mark.parametrize("scheme", ["rk4", "ssprk3"])
def test_stochastic_ou_matches_the_oracle_on_its_own_draws(g, field, dtype, tol, scheme):
    """The band force the host's process holds each step (src/Forcing.jl:1395-1405), added at every
    stage on the device, against the oracle given the same force: 10 steps of uneven dt."""
    n = 32
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=dtype)
    ou = g.StochasticOU(dom, forcing_variance=0.5, correlation_time=0.05, lowest_shell=1, highest_shell=2, seed=99)
    kw = {"velocity_forcing": ou} if field == "velocity" else {"scalar_forcing": ou}
    cond = g.FlowConditions(viscosity=0.02, schmidt=0.8, **kw)
    u, p = g.random_phase_velocity(n, seed=91), g.random_phase_scalar(n, seed=92)
    d = dns.Domain(n)
    uo, po, dep = u, p, None
    sch = {"rk4": g.IFLowStorageRK4, "ssprk3": g.IFSSPRK3}[scheme]
    with g.Simulation(dom, conditions=cond, state=g.ResolvedFields(u, p), mean_gradient=0.5,
                      field_integration_scheme=sch) as sim:
        for s in range(1, 11):
            dt = 2e-3 * (1 + 0.3 * math.sin(s))
            sim.advance(dt)
            src = sim.ou[0 if field == "velocity" else 1].full_source(n)
            okw = {"velocity_source": src} if field == "velocity" else {"scalar_source": src}
            oc = dns.Conditions(viscosity=0.02, schmidt=0.8, mean_gradient=0.5, band_eps=float(np.finfo(dtype).eps), **okw)
            uo, po, dep = dns.step(d, oc, scheme, uo, po, dt, dep)
            if s in (1, 10):
                st = sim.state()
                eu, ep = rel(st.velocity_spectral, uo), rel(st.scalar_spectral, po)
                print(f"ou {field} {scheme} {np.dtype(dtype).name} step {s}: u {eu:.2e} phi {ep:.2e}", flush=True)
                assert eu < tol and ep < tol


def test_stochastic_ou_injects_its_force_into_a_fluid_at_rest(g):
    """From rest, with nu = 0 in the band's short window, the first step moves the band modes by
    dt f to first order and by nothing else: the device's increment over one small step, divided
    by dt, is the uploaded force (to O(dt))."""
    n, dt = 32, 1e-6
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float64)
    ou = g.StochasticOU(dom, forcing_variance=1.0, correlation_time=10.0, seed=5)
    zero = np.zeros((3, n, n, n // 2 + 1), dtype=np.complex128)
    with g.Simulation(dom, conditions=g.FlowConditions(viscosity=1e-3, schmidt=1.0, velocity_forcing=ou),
                      state=g.ResolvedFields(zero, None)) as sim:
        sim.advance(dt)
        got = sim.state().velocity_spectral / dt
        want = sim.ou[0].full_source(n)
    assert rel(got, want) < 1e-4
