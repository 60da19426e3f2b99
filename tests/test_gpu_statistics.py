"""Long runs on an MI355X, compared statistically: once chaos has decorrelated the FP32-bulk and
FP64 trajectories of the same forced flow, their time-averaged spectra and energy budget must still
agree, and the ConstantPower injection must balance the mean dissipation; over a shorter span the
FP64 device run follows the FP64 oracle's trajectory itself."""
import math

import numpy as np
import pytest

from oracle import dns

from test_gpu_parity import L, make_sim, rel

pytestmark = pytest.mark.gpu


def _forced(g, n, dtype, seed=2026):
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=dtype)
    cond = g.FlowConditions(viscosity=g.predicted_viscosity(dom, 0.1, kmax_eta=1.5), schmidt=1.0,
                            velocity_forcing=g.ConstantPower(dom, injection_rate=0.1, lowest_shell=1, highest_shell=2))
    u = g.random_phase_velocity(n, seed=seed)
    p = np.zeros((n, n, n // 2 + 1), dtype=np.complex128)
    return g.Simulation(dom, conditions=cond, state=g.ResolvedFields(u, p), mean_gradient=1.0)


@pytest.mark.timeout(600)
def test_long_run_statistics_fp32_bulk_against_fp64(g):
    """64^3, eps = 0.1 on shells 1-2, k_max eta = 1.5, mean-gradient scalar, IF-RK4 at dt = 5e-3
    for 12,000 steps (T = 60, about a dozen large-eddy times k/eps), statistics over the second half
    sampled every 100 steps.  A third run, FP64 from the initial state perturbed by 1e-6, measures
    the sampling noise of these time averages: the FP32-bulk run must agree with the FP64 run as
    closely as two FP64 realizations do (the scalar under a mean gradient decorrelates slowly, so
    its averages carry most of that noise)."""
    n, dt, steps, every = 64, 5e-3, 12000, 100
    a, b, c = _forced(g, n, np.float32), _forced(g, n, np.float64), _forced(g, n, np.float64)
    u = c.state()
    c.set_fields(g.ResolvedFields(u.velocity_spectral * (1 + 1e-6), u.scalar_spectral))
    keys = ("kinetic_energy", "dissipation", "scalar_variance", "scalar_dissipation")
    runs = (a, b, c)
    acc = [[] for _ in runs]
    with a, b, c:
        for s in range(1, steps + 1):
            for sim in runs:
                sim.step_fields(dt, graph=True)
            if s % every == 0 and s > steps // 2:
                for i, sim in enumerate(runs):
                    r = sim.reductions()
                    acc[i].append((np.array([r[k] for k in keys]), sim.energy_spectrum(0)[1], sim.energy_spectrum(1)[1]))
        late = rel(a.state().velocity_spectral, b.state().velocity_spectral)
    means = [tuple(np.mean(x, axis=0) for x in zip(*ac)) for ac in acc]
    (ra, eua, epa), (rb, eub, epb), (rc, euc, epc) = means
    shells = np.arange(1, 20)

    def diff(x, y):
        return np.abs(x - y) / np.abs(y)

    print(f"trajectories at the end: relative distance {late:.2e} (decorrelated)", flush=True)
    for i, k in enumerate(keys):
        print(f"<{k}>: FP32 {ra[i]:.5e} FP64 {rb[i]:.5e} FP64' {rc[i]:.5e} | FP32-FP64 {diff(ra[i], rb[i]):.3f} "
              f"FP64'-FP64 {diff(rc[i], rb[i]):.3f}", flush=True)
    du, dun = diff(eua[shells], eub[shells]), diff(euc[shells], eub[shells])
    dp, dpn = diff(epa[shells], epb[shells]), diff(epc[shells], epb[shells])
    print("E(k) FP32-FP64:", " ".join(f"{v:.3f}" for v in du), flush=True)
    print("E(k) FP64'-FP64:", " ".join(f"{v:.3f}" for v in dun), flush=True)
    print("E_phi(k) FP32-FP64:", " ".join(f"{v:.3f}" for v in dp), flush=True)
    print("E_phi(k) FP64'-FP64:", " ".join(f"{v:.3f}" for v in dpn), flush=True)
    assert late > 1e-3                                  # the comparison is statistical, not pointwise
    # the velocity: tight, and the forcing's power dissipated on average
    assert np.all(diff(ra[:2], rb[:2]) < 0.05)
    assert np.mean(du) < 0.06 and np.max(du) < 0.15
    for r_ in (ra, rb, rc):
        assert abs(r_[1] - 0.1) < 0.03 * 0.1
    # the scalar: within the spread of two FP64 realizations (with a floor for an unlucky pair)
    assert np.all(diff(ra[2:], rb[2:]) <= np.maximum(2 * diff(rc[2:], rb[2:]), 0.1))
    assert np.mean(dp) <= max(2 * np.mean(dpn), 0.1)


def test_fp64_device_follows_the_oracle_trajectory_over_200_steps(g):
    """32^3 forced, 200 IF-RK4 steps: the FP64 device trajectory stays on the oracle's (round-off
    growth only)."""
    n, dt, steps = 32, 5e-3, 200
    sim, u, p, oc = make_sim(g, n, np.float64, forcing=0.1, nu=0.02, seed=11)
    d = dns.Domain(n)
    uo, po, dep = u, p, None
    with sim:
        for s in range(steps):
            sim.step_fields(dt)
            uo, po, dep = dns.step(d, oc, "rk4", uo, po, dt, dep)
        st = sim.state()
        eu, ep = rel(st.velocity_spectral, uo), rel(st.scalar_spectral, po)
    print(f"after {steps} steps: u {eu:.2e} phi {ep:.2e}", flush=True)
    assert eu < 1e-9 and ep < 1e-9
