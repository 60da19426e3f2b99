"""Box scaling on the device (test/box_scaling.jl): a run on a box of side L = 0.02 is the run on
L = 2 pi with every quantity rescaled by its dimension.  With sigma = L / 2 pi, lengths and
velocities scale by sigma, the kinetic energy, the dissipation, the viscosity and the injection
rates by sigma^2, wavenumbers by 1 / sigma, the mean scalar gradient by 1 / sigma, and time and the
scalar not at all.  On the 2 pi box the fundamental 2 pi / L is 1, so a dropped or doubled k0
anywhere (the integrating factor, the Leray projection, the curl, the dissipation, a forcing, the
CFL bound, the tracer stencil) passes every other test; here it fails by orders of magnitude.
The oracle is the scaling itself.  Bar: 1e-13 relative, as the reference asks (FP64).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
LS = 0.02
SIGMA = LS / (2 * math.pi)
N = 16


def rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(np.asarray(a))))


def norm_rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(a)))


def _domain(g, side):
    return g.FixedDomain(grid_points_per_axis=N, side_length=side, dealiasing=g.TwoThirds(), element_type=np.float64)


def _forcings(g, name, dom, s2):
    """The case's velocity and scalar forcings on a box whose sigma^2 is s2 (1 on the 2 pi box)."""
    if name == "NoForcing":
        return None, None
    if name == "ConstantPower":
        return g.ConstantPower(dom, injection_rate=0.3 * s2), None
    if name == "Linear":
        return g.Linear(dom, rate=0.2), None
    if name == "VarianceRescaling":
        return g.VarianceRescaling(dom, target=1.0 * s2, over="total"), None
    if name == "VarianceController":
        return g.VarianceController(dom, target=1.5 * s2, turnover_time=1.0), None
    return (g.ConstantPower(dom, injection_rate=0.3 * s2),
            g.VarianceRescaling(dom, target=1.0, over="total"))


@pytest.mark.parametrize("name", ["NoForcing", "ConstantPower", "Linear", "VarianceRescaling", "VarianceController",
                                  "ConstantPower with a scalar"])
def test_fixed_steps_scale_with_the_box(g, name):
    """test/box_scaling.jl:45-72: four fixed steps of each forcing (and a forced scalar under a
    mean gradient) on both boxes."""
    scalar = name.endswith("scalar")
    runs = []
    for side, s in ((2 * math.pi, 1.0), (LS, SIGMA)):
        dom = _domain(g, side)
        vf, sf = _forcings(g, name, dom, s * s)
        cond = g.FlowConditions(viscosity=0.02 * s * s, schmidt=1.0, velocity_forcing=vf, scalar_forcing=sf)
        u = g.random_phase_velocity(N, side_length=side, energy=0.5 * s * s, seed=81)
        p = g.random_phase_scalar(N, side_length=side, variance=0.5, seed=82) if scalar else None
        with g.Simulation(dom, conditions=cond, state=g.ResolvedFields(u, p),
                          mean_gradient=(0.8 / s if scalar else 0.0)) as sim:
            for _ in range(4):
                sim.advance(0.002)
            st = sim.state()
            kappa, e = sim.energy_spectrum()
            # test/validation.jl:26-35: the shells are 2 pi / L wide, and sum E dkappa is k
            dk = 2 * math.pi / side
            assert kappa[1] == pytest.approx(dk, rel=1e-15)
            assert float(np.sum(e)) * dk == pytest.approx(sim.reductions()["kinetic_energy"], rel=1e-12)
            runs.append((st.velocity_spectral, st.scalar_spectral, sim.reductions(), e))
    (uR, pR, rR, eR), (uL, pL, rL, eL) = runs
    assert rel(eR, eL / SIGMA ** 3) < 1e-13   # E(kappa) dkappa is an energy: E scales by sigma^3
    eu = rel(uR, uL / SIGMA)
    print(f"{name}: velocity {eu:.2e}", flush=True)
    assert eu < 1e-13
    assert rL["kinetic_energy"] == pytest.approx(SIGMA ** 2 * rR["kinetic_energy"], rel=1e-13)
    assert rL["dissipation"] == pytest.approx(SIGMA ** 2 * rR["dissipation"], rel=1e-13)
    if scalar:
        ep = rel(pR, pL)
        print(f"{name}: scalar {ep:.2e}", flush=True)
        assert ep < 1e-13
        assert rL["scalar_variance"] == pytest.approx(rR["scalar_variance"], rel=1e-13)


def test_adaptive_steps_a_stochastic_forcing_and_tracers_scale_with_the_box(g):
    """test/box_scaling.jl:73-101: eight CFL steps under StochasticOU with AB(3) tracers and their
    unwrapped record; the clock agrees, since the CFL step is box-free."""
    rng = np.random.default_rng(9)
    positions = rng.uniform(0, 2 * math.pi, (3, 20))
    out = []
    for side, s in ((2 * math.pi, 1.0), (LS, SIGMA)):
        dom = _domain(g, side)
        vf = g.StochasticOU(dom, forcing_variance=0.5 * s * s, correlation_time=0.7, seed=5)
        cond = g.FlowConditions(viscosity=0.02 * s * s, schmidt=1.0, velocity_forcing=vf)
        u = g.random_phase_velocity(N, side_length=side, energy=0.5 * s * s, seed=83)
        with g.Simulation(dom, conditions=cond, state=g.ResolvedFields(u, None),
                          tracers=g.Tracers(dom, positions=positions * s, unwrapped=True),
                          particle_integration_scheme=g.AB(3)) as sim:
            for _ in range(8):
                sim.advance()
            out.append((sim.time, sim.state().velocity_spectral, sim.tracer_positions(), sim.unwrapped_positions(),
                        sim.reductions()))
    (tR, uR, xR, wR, rR), (tL, uL, xL, wL, rL) = out
    print(f"clock {tR} / {tL}, velocity {rel(uR, uL / SIGMA):.2e}, positions {norm_rel(xR, xL / SIGMA):.2e}, "
          f"unwrapped {norm_rel(wR, wL / SIGMA):.2e}", flush=True)
    assert tL == pytest.approx(tR, rel=1e-13)
    assert rel(uR, uL / SIGMA) < 1e-13
    assert norm_rel(xR, xL / SIGMA) < 1e-13
    assert norm_rel(wR, wL / SIGMA) < 1e-13
    assert rL["kinetic_energy"] == pytest.approx(SIGMA ** 2 * rR["kinetic_energy"], rel=1e-13)
    assert rL["dissipation"] == pytest.approx(SIGMA ** 2 * rR["dissipation"], rel=1e-13)
