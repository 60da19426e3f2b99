"""Properties of the field schemes on the device that need no oracle (test/integrators.jl:
"temporal convergence order"; test/scalar.jl: "diffusion exactness (u = 0)", "mean conservation
(conservative form)"):
  * each scheme's error against a 256-step run falls as dt^order over 8, 16, 32 steps;
  * with u = 0 every step is the integrating factor alone, so the scalar decays exactly as
    exp(-kappa k^2 t) (the mean-gradient source -G u_z vanishes with u);
  * the scalar's mean, the k = 0 coefficient, is conserved to the bit: the flux divergence has
    no k = 0 part and the integrating factor there is exp(0) = 1.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
L = 2 * math.pi
SCHEMES = [("IFLowStorageRK3", 3), ("IFLowStorageRK4", 4), ("IFSSPRK3", 3), ("IFHeun", 2)]


def _dom(g, n, dtype=np.float64):
    return g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=dtype)


@pytest.mark.parametrize("scheme,order", SCHEMES)
def test_field_schemes_converge_at_their_order(g, scheme, order):
    n, T = 32, 0.4
    dom = _dom(g, n)
    u, p = g.random_phase_velocity(n, seed=91), g.random_phase_scalar(n, seed=92)

    def run(steps):
        with g.Simulation(dom, conditions=g.FlowConditions(viscosity=0.05, schmidt=1.0),
                          state=g.ResolvedFields(u, p), field_integration_scheme=getattr(g, scheme),
                          mean_gradient=1.0) as sim:
            for _ in range(steps):
                sim.advance(T / steps)
            st = sim.state()
            return np.concatenate([st.velocity_spectral.ravel(), st.scalar_spectral.ravel()])

    fine = run(256)
    errs = [float(np.max(np.abs(run(s) - fine))) for s in (8, 16, 32)]
    ratios = [errs[0] / errs[1], errs[1] / errs[2]]
    print(f"{scheme}: errors {errs}, ratios {ratios}", flush=True)
    r = 2.0 ** order
    assert all(0.7 * r < q < 1.4 * r for q in ratios)


@pytest.mark.parametrize("scheme", [s for s, _ in SCHEMES])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-14), (np.float32, 2e-6)])
def test_pure_diffusion_is_exact(g, scheme, dtype, tol):
    n, kappa = 32, 0.05 / 0.7
    dom = _dom(g, n, dtype)
    p = g.random_phase_scalar(n, seed=93)
    p0 = p.astype(np.complex64).astype(np.complex128) if dtype is np.float32 else p
    u = np.zeros((3, n, n, n // 2 + 1), dtype=np.complex128)
    dts = [0.01, 0.013, 0.008, 0.011, 0.009]
    with g.Simulation(dom, conditions=g.FlowConditions(viscosity=0.05, schmidt=0.7), state=g.ResolvedFields(u, p),
                      field_integration_scheme=getattr(g, scheme), mean_gradient=0.7) as sim:
        for dt in dts:
            sim.advance(dt)
        st = sim.state()
    m = np.fft.fftfreq(n, 1.0 / n)
    k2 = m[:, None, None] ** 2 + m[None, :, None] ** 2 + np.arange(n // 2 + 1)[None, None, :] ** 2
    want = p0 * np.exp(-kappa * k2 * sum(dts))
    err = float(np.max(np.abs(st.scalar_spectral - want)) / np.max(np.abs(want)))
    print(f"{scheme} {np.dtype(dtype).name}: {err:.2e}", flush=True)
    assert err < tol
    assert not np.any(st.velocity_spectral)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_the_scalar_mean_is_conserved_to_the_bit(g, dtype):
    n = 32
    dom = _dom(g, n, dtype)
    p = g.random_phase_scalar(n, seed=94)
    p[0, 0, 0] = 0.3 * n ** 3
    vf = g.ConstantPower(dom, injection_rate=0.1)
    with g.Simulation(dom, conditions=g.FlowConditions(viscosity=0.02, schmidt=1.0, velocity_forcing=vf),
                      state=g.ResolvedFields(g.random_phase_velocity(n, seed=95), p), mean_gradient=1.0) as sim:
        for _ in range(10):
            sim.advance(0.005)
        st = sim.state()
    held = complex(np.complex64(p[0, 0, 0])) if dtype is np.float32 else p[0, 0, 0]   # the upload's rounding
    assert st.scalar_spectral[0, 0, 0] == held
    assert np.max(np.abs(st.scalar_spectral[1:4, 1:4, 1:4])) > 0   # the rest of the field moved


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_a_resting_flow_stays_at_rest_under_the_zero_speed_step(g, dtype):
    """test/regressions/03-numerics-and-types-03-cfl-zero-speed.jl: at max|u| = 0 the adaptive
    step falls back to cfl dx; a resting unforced flow is a fixed point of every advance, the clock
    moves, and the scalar's decay over an interval does not depend on how it is split."""
    n, nu = 16, 1e-3
    dom = _dom(g, n, dtype)
    u = np.zeros((3, n, n, n // 2 + 1), dtype=np.complex128)
    p = g.random_phase_scalar(n, seed=99)
    runs = []
    for splits in (1, 2):
        with g.Simulation(dom, conditions=g.FlowConditions(viscosity=nu, schmidt=1.0), state=g.ResolvedFields(u, p),
                          field_integration_scheme=g.IFLowStorageRK3) as sim:
            assert sim.max_speed() == 0
            dt = sim.cfl_timestep()
            assert dt == pytest.approx(sim.cfl * dom.grid_spacing, rel=1e-15)
            for _ in range(20 // splits):
                sim.advance(dt * splits)
            st = sim.state()
            assert not np.any(st.velocity_spectral)
            assert sim.reductions()["kinetic_energy"] == 0
            assert math.isfinite(sim.time) and sim.time == pytest.approx(20 * dt, rel=1e-12)
            runs.append(st.scalar_spectral)
    err = float(np.max(np.abs(runs[0] - runs[1])) / np.max(np.abs(runs[1])))
    # FP32: 20 steps x 3 stage factors round up to 60 x 2^-24 = 3.6e-6 apart; a step-dependent
    # decay would differ at order one
    assert err < (1e-14 if dtype is np.float64 else 5e-6)
