"""Budgets and identities on the device (test/forcing.jl: "ConstantPower injects exactly its
injection rate"; test/validation.jl: "dissipation -- the vorticity and Dirichlet routes").

The 2/3-truncated nonlinear terms move energy and scalar variance between retained modes without
creating any, so over one short step
    dk/dt = P - eps,     d<phi^2>/dt = P_phi - eps_phi
hold to the step's truncation error, with the time integrals of the dissipations taken by the
trapezoid rule (error dt^2 eps'' / 12).  ConstantPower sets its amplitude from each stage's own
band, so P enters exactly; a band counted once instead of twice, or the amplitude of the previous
stage, fails by order one.  The dissipation the run reports, nu <u, |k|^2 u> (the Dirichlet route),
equals nu <|omega|^2> from the deposited physical vorticity on any solenoidal state.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
L = 2 * math.pi


@pytest.mark.parametrize("scheme", ["IFLowStorageRK4", "IFSSPRK3", "IFLowStorageRK3"])
def test_constant_power_injects_exactly_its_rate(g, scheme):
    n, P, Pphi, dt = 32, 0.1, 0.07, 1e-3
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float64)
    cond = g.FlowConditions(viscosity=0.02, schmidt=0.8, velocity_forcing=g.ConstantPower(dom, injection_rate=P),
                            scalar_forcing=g.ConstantPower(dom, injection_rate=Pphi, lowest_shell=2, highest_shell=3))
    with g.Simulation(dom, conditions=cond, field_integration_scheme=getattr(g, scheme),
                      state=g.ResolvedFields(g.random_phase_velocity(n, seed=96), g.random_phase_scalar(n, seed=97)),
                      mean_gradient=0.0) as sim:
        r0 = sim.reductions()
        sim.advance(dt)
        r1 = sim.reductions()
    dk = (r1["kinetic_energy"] - r0["kinetic_energy"]) / dt
    dv = (r1["scalar_variance"] - r0["scalar_variance"]) / dt
    res_k = dk - (P - 0.5 * (r0["dissipation"] + r1["dissipation"]))
    res_v = dv - (Pphi - 0.5 * (r0["scalar_dissipation"] + r1["scalar_dissipation"]))
    print(f"{scheme}: dk/dt {dk:.6f} residual {res_k:.2e}; d<phi^2>/dt {dv:.6f} residual {res_v:.2e}", flush=True)
    assert abs(res_k) < 1e-4 * P
    assert abs(res_v) < 1e-4 * Pphi


@pytest.mark.parametrize("n", [32, 48])
def test_dissipation_by_the_dirichlet_route_equals_nu_omega_squared(g, n):
    nu = 0.02
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float64)
    cond = g.FlowConditions(viscosity=nu, schmidt=1.0, velocity_forcing=g.ConstantPower(dom, injection_rate=0.1))
    with g.Simulation(dom, conditions=cond, state=g.ResolvedFields(g.random_phase_velocity(n, seed=98), None),
                      keep_vorticity=True) as sim:
        sim.advance(2e-3)
        w = sim.physical(g._lib.FIELD_VORTICITY)
        eps = sim.reductions()["dissipation"]
    direct = nu * float(np.mean(np.sum(w * w, axis=0)))
    print(f"{n}: Dirichlet {eps!r}, vorticity {direct!r}", flush=True)
    assert direct == pytest.approx(eps, rel=1e-12)
