"""Host-side rows of the run API on the CPU: the viscosity rule, the CFL ceiling and the adaptive
step's three bounds, the reality check of the adoption door and the band truncation
(src/SimulationRun.jl:647-655,1051-1096,1236-1248; src/Propagator.jl:895-919;
src/SimulationDomain.jl:520-537)."""
import math

import numpy as np
import pytest

from oracle import dns

L = 2 * math.pi


@pytest.mark.parametrize("n,nu", [(64, 1.347e-2), (128, 5.346e-3), (256, 2.121e-3), (512, 8.419e-4),
                                  (1024, 3.341e-4)])
def test_predicted_viscosity_gives_kmax_eta_of_one_and_a_half(g, n, nu):
    """nu = eps^(1/3) eta^(4/3) with eta = 1.5 / kappa_max and kappa_max = (2/3)(N/2)(2 pi / L):
    at eps = 0.1 the values quoted for the configurations, and k_max eta = 1.5 back from nu."""
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float64)
    got = g.predicted_viscosity(dom, 0.1, kmax_eta=1.5)
    assert got == pytest.approx(nu, rel=5e-4)
    kmax = (2 / 3) * (n / 2)
    eta = (got ** 3 / 0.1) ** 0.25
    assert kmax * eta == pytest.approx(1.5, rel=1e-12)


def test_recommended_cfl_is_the_base_ceiling_under_two_thirds(g):
    dom = g.FixedDomain(grid_points_per_axis=32, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float64)
    assert g.recommended_cfl(g.IFLowStorageRK4, dom) == 0.96
    assert g.recommended_cfl(g.IFLowStorageRK3, dom) == 0.5
    assert g.recommended_cfl(g.IFSSPRK3, dom) == 0.5   # src/Propagator.jl:911, the RK3 ceiling
    assert g.recommended_cfl(g.IFHeun, dom) == 0.25    # src/Propagator.jl:906


def test_cfl_bound_takes_the_smallest_of_its_three_bounds(g):
    dx = L / 64
    # advective binds in a turbulent flow
    assert g.cfl_bound(2.0, dx, 0.96, 1e-2, L) == pytest.approx(0.96 * dx / 2.0, rel=1e-15)
    # at rest the advective bound is cfl dx, and the resolution bound 1 / (D kappa_min^2) can bind
    assert g.cfl_bound(0.0, dx, 0.96, 1e-2, L) == pytest.approx(0.96 * dx, rel=1e-15)
    assert g.cfl_bound(0.0, dx, 0.96, 20.0, L) == pytest.approx(1 / (20.0 * 1.0 ** 2), rel=1e-15)
    assert g.cfl_bound(1e-9, dx, 0.96, 10.0, 4 * math.pi) == pytest.approx(1 / (10.0 * 0.5 ** 2), rel=1e-15)
    # the caller's ceiling
    assert g.cfl_bound(2.0, dx, 0.96, 1e-2, L, max_dt=1e-4) == 1e-4
    # a blown-up flow is an error, not a step
    for bad in (math.nan, math.inf):
        with pytest.raises(RuntimeError, match="blown up"):
            g.cfl_bound(bad, dx, 0.96, 1e-2, L, step=7)


def test_reality_check_of_the_adoption_door(g):
    """A Hermitian kx = 0 plane passes; an O(1) anti-Hermitian pair, or a NaN, is refused; a
    round-off disagreement is not a violation."""
    from gbg_amd.api import _check_reality

    n = 16
    d = dns.Domain(n)
    rng = np.random.default_rng(4)
    f = dns.reality_projection(d, dns.rfft3(rng.standard_normal((n, n, n))) * d.mask)
    _check_reality(f, "scalar", np.float64)
    near = f.copy()
    near[1, 2, 0] *= 1 + 1e-12
    _check_reality(near, "scalar", np.float64)
    bad = f.copy()
    bad[1, 2, 0] += 1j * np.abs(f).max()
    with pytest.raises(ValueError, match="not the spectrum of a real"):
        _check_reality(bad, "scalar", np.float64)
    nan = f.copy()
    nan[3, 3, 3] = np.nan
    with pytest.raises(ValueError):
        _check_reality(nan, "scalar", np.float64)


def test_truncate_to_band_keeps_exactly_the_closed_ball(g):
    for n in (16, 48):
        dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float64)
        f = np.ones((3, n, n, n // 2 + 1), dtype=np.complex128)
        g.truncate_to_band(f, dom)
        assert np.array_equal(f[0] != 0, dns.Domain(n).mask)


def test_integral_length_of_a_measured_spectrum(g):
    """src/Observables.jl:567-581 through the API's spectrum form, after
    test/observables_analytic.jl:71-73."""
    assert g.integral_length(([0.0, 1.0, 2.0], [5.0, 2.0, 1.0])) == pytest.approx(
        3 * math.pi / 4 * (2.0 / 1.0 + 1.0 / 2.0) / (2.0 + 1.0), rel=1e-14)
    with pytest.raises(ValueError, match="no energy"):
        g.integral_length(([0.0, 1.0], [5.0, 0.0]))


def test_stationarity_monitor_flatness_and_gates(g):
    """The warm-up detector's core on synthetic signals (src/SimulationRun.jl:1433-1477): a constant
    dissipation is flat once min_turnovers have passed and the window is full; a drifting one is
    not; the skewness gate needs a plateau of size at least skewness_floor."""
    from gbg_amd.api import StationarityMonitor

    def feed(m, f, upto, skew=lambda t: 0.0):
        t, out = 0.0, False
        while t <= upto + 1e-12:
            out = m.push_and_test(t, f(t), 1.0, skew(t))
            t += 0.05
        return out

    kw = dict(min_turnovers=5, window_turnovers=4, tolerance=0.05, gate_scalar=False, gate_skewness=False,
              skewness_floor=0.1)
    assert not feed(StationarityMonitor(**kw), lambda t: 0.1, 4.9)       # before min_turnovers
    assert feed(StationarityMonitor(**kw), lambda t: 0.1, 6.0)
    assert not feed(StationarityMonitor(**kw), lambda t: 0.1 * (1 + 0.1 * t), 8.0)   # drifts 5 % per half-window
    gated = dict(kw, gate_skewness=True)
    assert not feed(StationarityMonitor(**gated), lambda t: 0.1, 8.0, skew=lambda t: 0.02)   # below the floor
    assert feed(StationarityMonitor(**gated), lambda t: 0.1, 8.0, skew=lambda t: -0.45)
    # the flatness test itself: two half-window means against tolerance * |mean|
    S = StationarityMonitor
    assert S.is_flat([0, 1, 2, 3], [1.0, 1.0, 1.04, 1.04], 3.0, 4.0, 0.05)
    assert not S.is_flat([0, 1, 2, 3], [1.0, 1.0, 1.2, 1.2], 3.0, 4.0, 0.05)
    assert S.is_flat([0, 1, 2, 3], [0.0, 0.0, 0.0, 0.0], 3.0, 4.0, 0.05)    # identically zero is flat
    assert not S.is_flat([2, 3], [1.0, 1.0], 3.0, 4.0, 0.05)                # one half empty


def test_variance_controller_default_turnover_time_follows_the_placeholder_model(g):
    """t_l = 2 k_inf / (3 eps_inf), eps_inf = (k_inf / k~)^(3/2) / L, k~ piecewise linear in
    1/highest_shell through (0, 0.47), (1/4, 0.56), (1/2, 0.72) (src/Forcing.jl:852-862, :887, :945-953)."""
    dom = g.FixedDomain(grid_points_per_axis=32, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float64)
    for hi, kt in ((2, 0.72), (4, 0.56), (math.inf, 0.47), (3, 0.56 + 0.16 * (1 / 3 - 0.25) / 0.25)):
        f = g.VarianceController(dom, target=1.5, highest_shell=hi)
        eps = (1.5 / kt) ** 1.5 / L
        assert f.turnover_time == pytest.approx(2 * 1.5 / (3 * eps), rel=1e-12)
        assert f.turnover_time_derived
    # test/forcing.jl:369-371: a band the model does not cover needs the time stated
    with pytest.raises(ValueError, match="turnover_time"):
        g.VarianceController(dom, target=1.5, lowest_shell=2, highest_shell=3)
    assert g.VarianceController(dom, target=1.5, lowest_shell=2, highest_shell=3, turnover_time=1.0).turnover_time == 1.0
    # test/forcing.jl:407-408 and the target guard
    for bad in ({"gain": -1.0}, {"gain": math.inf}, {"target": 0.0}):
        kw = {"target": 1.5, **bad}
        with pytest.raises(ValueError):
            g.VarianceController(dom, **kw)


def test_variance_rescaling_guards(g):
    dom = g.FixedDomain(grid_points_per_axis=16, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float32)
    with pytest.raises(ValueError, match="over"):
        g.VarianceRescaling(dom, target=0.5, over="shell")
    with pytest.raises(ValueError, match="non-negative"):
        g.VarianceRescaling(dom, target=-1.0, over="band")
    f = g.VarianceRescaling(dom, target=0.1, over="total", highest_shell=math.inf)
    assert f.target == float(np.float32(0.1)) and f.highest_shell == math.inf


def test_particle_scheme_guards_and_their_trail_capacities(g):
    """src/ParticlePropagator.jl:106-115, :225-249: ABM(k) k = 1..7 keeps k + 1 samples and is of
    order k + 1; AB(k) k = 1..8 keeps k and is of order k; Heun is ABM(1)."""
    for k in (0, 8, 2.0):
        with pytest.raises(ValueError, match="ABM"):
            g.ABM(k)
    for k in (0, 9):
        with pytest.raises(ValueError, match="AB"):
            g.AB(k)
    assert (g.ABM(3).trail_capacity, g.ABM(3).order, g.ABM(3).corrects) == (4, 4, True)
    assert (g.AB(4).trail_capacity, g.AB(4).order, g.AB(4).corrects) == (4, 4, False)
    h = g.Heun()
    assert (h.steps, h.trail_capacity, h.order, repr(h)) == (1, 2, 2, "Heun()")


def test_startup_bound(g):
    """src/SimulationRun.jl:1099-1129: absent for one-sample schemes; tau (dt/tau)^(p/3) with one
    sample in the trail; then the growth factor (2 through order 5, 1.5 through 7, 1.25 at 8) times
    the last step."""
    from gbg_amd.api import startup_bound
    nu, eps, dt = 0.01, 0.1, 0.02
    tau = math.sqrt(nu / eps)
    assert startup_bound(g.Heun(), [0.0], dt, nu, eps) == math.inf
    assert startup_bound(g.AB(4), [0.0], dt, nu, eps) == pytest.approx(tau * (dt / tau) ** (4 / 3), rel=1e-15)
    assert startup_bound(g.ABM(5), [0.0], dt, nu, eps) == pytest.approx(tau * (dt / tau) ** 2, rel=1e-15)
    assert startup_bound(g.AB(4), [0.3, 0.2], dt, nu, eps) == pytest.approx(0.2, rel=1e-12)
    assert startup_bound(g.AB(7), [0.3, 0.2], dt, nu, eps) == pytest.approx(0.15, rel=1e-12)
    assert startup_bound(g.AB(8), [0.3, 0.2], dt, nu, eps) == pytest.approx(0.125, rel=1e-12)
    assert startup_bound(g.AB(4), [0.0], dt, nu, 0.0) == math.inf       # a flow at rest


def test_checkpoint_store_round_trips_groups_attributes_and_datasets(g, tmp_path):
    """The container of the checkpoint schema: nested groups, JSON attributes (inf included),
    complex and real datasets; loading executes nothing from the file (allow_pickle=False)."""
    from gbg_amd import checkpoint as ck
    s = ck.Store()
    top = ck._tagged(s.root, "simulation", "Simulation")
    top.attrs.update(time=0.25, step=3, schedule=[1, 0, 0, 0, 1])
    f = ck._tagged(top, "velocity_forcing", "Linear")
    f.attrs.update(rate=0.3, lowest_shell=1, highest_shell=math.inf)
    z = (np.arange(12) + 1j * np.arange(12)).reshape(3, 4)
    st = ck._tagged(top, "state", "ResolvedFields")
    st["velocity_spectral"] = z
    s.save(tmp_path / "c.npz")
    r = ck.Store.load(tmp_path / "c.npz")
    top2 = r.root.group("simulation", "t")
    assert ck._require_tag(top2, "t", "Simulation")["schedule"] == [1, 0, 0, 0, 1]
    assert top2.group("velocity_forcing", "t").attrs["highest_shell"] == math.inf
    assert np.array_equal(top2.group("state", "t")["velocity_spectral"], z)
    with pytest.raises(ValueError, match="holds a"):
        ck._require_tag(top2.group("state", "t"), "t", "Simulation")
    with pytest.raises(ValueError, match="no object"):
        r.root.group("missing", "read_simulation")
    with pytest.raises(ValueError, match="newer"):
        top2.attrs["version"] = 99
        ck._require_tag(top2, "t", "Simulation")


def _ou(g, order, **kw):
    from gbg_amd.stochastic import OUProcess
    dom = g.FixedDomain(grid_points_per_axis=32, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float64)
    f = g.StochasticOU(dom, forcing_variance=kw.get("var", 0.3), correlation_time=kw.get("tl", 0.5),
                       lowest_shell=1, highest_shell=kw.get("hi", 2), seed=kw.get("seed", 12345))
    return f, OUProcess(f, order)


def test_stochastic_ou_draws_are_transverse_real_band_limited_and_keyed(g):
    """src/Forcing.jl:1360-1385: each increment is a real field's spectrum cut to the band and, for
    the velocity, transverse to kappa; the stream is a pure function of (seed, draw, field)."""
    f, p = _ou(g, 1)
    nc = f.companion_points
    xi = p.increment(7)
    m = np.fft.fftfreq(nc, 1.0 / nc)
    kz, ky, kx = m[:, None, None], m[None, :, None], np.arange(nc // 2 + 1)[None, None, :]
    assert np.max(np.abs(kx * xi[0] + ky * xi[1] + kz * xi[2])) < 1e-12 * np.max(np.abs(xi))
    assert np.all(xi[:, ~f.band] == 0)
    plane = xi[..., 0]                                       # kx = 0: xi(-k) = conj xi(k)
    neg = (-np.arange(nc)) % nc
    assert np.max(np.abs(plane - np.conj(plane[:, neg][:, :, neg]))) < 1e-12
    assert np.array_equal(p.increment(7), xi)
    assert not np.array_equal(p.increment(8), xi)
    _, s = _ou(g, 0)
    assert not np.allclose(s.increment(7), xi[0])           # the scalar draws its own stream


def test_stochastic_ou_transition_is_exact_and_stationary(g):
    """a <- alpha a + sqrt(sigma^2 (1 - alpha^2)) xi exactly, and the physical mean square of the
    force averages sigma^2 over the stream (2000 draws, 3 % bar)."""
    f, p = _ou(g, 1, var=0.3, tl=0.5)
    before, draws = p.process.copy(), p.draws
    p.advance(0.05)
    alpha = math.exp(-0.05 / 0.5)
    want = alpha * before + math.sqrt(0.3 * (1 - alpha ** 2)) * p.increment(draws)
    assert np.max(np.abs(p.process - want)) <= 1e-14 * np.max(np.abs(want))
    nc = f.companion_points
    w = np.full(nc // 2 + 1, 2.0)
    w[0] = 1.0
    ms = [float(np.sum(w * np.abs(p.increment(j)) ** 2)) / nc ** 6 for j in range(2000)]
    assert np.mean(ms) == pytest.approx(1.0, rel=0.03)


def test_stochastic_ou_guards(g):
    dom = g.FixedDomain(grid_points_per_axis=16, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float64)
    with pytest.raises(ValueError, match="finite `highest_shell`"):
        g.StochasticOU(dom, forcing_variance=0.1, correlation_time=0.5, highest_shell=math.inf)
    with pytest.raises(ValueError, match="dealias cutoff"):
        g.StochasticOU(dom, forcing_variance=0.1, correlation_time=0.5, highest_shell=6)
    with pytest.raises(ValueError, match="correlation_time"):
        g.StochasticOU(dom, forcing_variance=0.1, correlation_time=0.0)
    with pytest.raises(ValueError, match="forcing_variance"):
        g.StochasticOU(dom, forcing_variance=-1.0, correlation_time=0.5)


def test_derived_scales_closed_forms_and_guards(g):
    """src/Observables.jl:611-760: Re_box = eps^(1/3) L^(4/3) / nu, Re_lambda = sqrt(20/3) k /
    sqrt(eps nu) = u' lambda / nu, lambda = sqrt(10 nu k / eps), eta = (nu^3/eps)^(1/4),
    tau_eta = sqrt(nu/eps), u_eta = (eps nu)^(1/4), and eta = u_eta tau_eta, eta u_eta / nu = 1."""
    k, eps, nu, Lb = 1.5, 0.1, 1e-3, 2 * math.pi
    lam = g.taylor_length(k, eps, nu)
    up = math.sqrt(2 * k / 3)
    assert g.reynolds_taylor(k, eps, nu) == pytest.approx(up * lam / nu, rel=1e-14)
    assert lam == pytest.approx(math.sqrt(15 * nu * up ** 2 / eps), rel=1e-14)
    assert g.reynolds_box(Lb, eps, nu) == pytest.approx((eps * Lb ** 4) ** (1 / 3) / nu, rel=1e-14)
    eta, tau, ue = g.kolmogorov_length(eps, nu), g.kolmogorov_time(eps, nu), g.kolmogorov_velocity(eps, nu)
    assert eta == pytest.approx(ue * tau, rel=1e-14) and eta * ue / nu == pytest.approx(1.0, rel=1e-14)
    with pytest.raises(ValueError, match="dissipation"):
        g.kolmogorov_length(0.0, nu)
    with pytest.raises(ValueError, match="viscosity"):
        g.reynolds_taylor(k, eps, -1.0)
    assert g.reynolds_box(Lb, 0.0, nu) == 0.0


def test_the_seed_is_the_same_field_on_every_box(g):
    """test/box_scaling.jl:34-43: the random-phase seed on a box of side 0.02 is the 2 pi seed with
    the velocity scaled by sigma = 0.02 / 2 pi (energy by sigma^2) and the scalar unchanged."""
    sigma = 0.02 / (2 * math.pi)
    u = g.random_phase_velocity(16, side_length=2 * math.pi, energy=1.5, seed=4)
    us = g.random_phase_velocity(16, side_length=0.02, energy=1.5 * sigma ** 2, seed=4)
    assert np.max(np.abs(u - us / sigma)) / np.max(np.abs(u)) < 1e-13
    p = g.random_phase_scalar(16, side_length=2 * math.pi, seed=5)
    ps = g.random_phase_scalar(16, side_length=0.02, seed=5)
    assert np.max(np.abs(p - ps)) / np.max(np.abs(p)) < 1e-13


@pytest.mark.parametrize("order", [1, 0])
def test_stochastic_ou_mean_square_spread_shares_and_lag(g, order):
    """test/forcing.jl:StochasticOU: over draws ten correlation times apart the force's physical mean
    square averages sigma^2 (2 %) and spreads as a chi^2 over D real degrees of freedom, sqrt(2/D)
    (10 %; D = 2 modes for the velocity after the transverse projection, modes for the scalar), which
    an increment normalised by its own realized norm fails.  At dt = T_L / 10 every band mode holds
    the same share (20 %) and the lag-one correlation is exp(-dt / T_L) (0.01)."""
    f, p = _ou(g, order, var=0.4, tl=0.5)
    nc = f.companion_points
    w = np.full(nc // 2 + 1, 2.0)
    w[0] = 1.0

    def mean_square(a):
        return float(np.sum(w * np.abs(a) ** 2)) / nc ** 6

    samples = []
    for _ in range(4000):
        p.advance(5.0)
        samples.append(mean_square(p.process))
    avg = float(np.mean(samples))
    assert abs(avg / 0.4 - 1) < 0.02
    dof = (2 if order == 1 else 1) * f.modes
    spread = float(np.std(samples, ddof=1)) / avg
    assert abs(spread / math.sqrt(2 / dof) - 1) < 0.1, (spread, math.sqrt(2 / dof))
    alpha = math.exp(-0.05 / 0.5)
    per_mode = np.zeros(p.process.shape[-3:])
    lagged = power = 0.0
    prev = p.process.copy()
    for _ in range(20000):
        p.advance(0.05)
        a = p.process
        per_mode += np.sum(np.abs(a) ** 2, axis=0) if order == 1 else np.abs(a) ** 2
        lagged += float(np.real(np.sum(a * np.conj(prev))))
        power += float(np.sum(np.abs(prev) ** 2))
        prev = a.copy()
    forced = per_mode[f.band]
    assert np.all(per_mode[~f.band] == 0)
    assert forced.max() / forced.min() - 1 < 0.2
    assert abs(lagged / power - alpha) < 0.01


def test_stochastic_ou_every_seed_bit_reaches_the_stream(g):
    """test/forcing.jl: a negative seed builds, draws and differs from its magnitude, and a seed's
    high bits change the stream."""
    dom = g.FixedDomain(grid_points_per_axis=16, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float64)

    def draw(seed):
        return g.api.OUProcess(g.StochasticOU(dom, forcing_variance=0.3, correlation_time=0.4, seed=seed), 1).process

    assert np.array_equal(draw(7), draw(7))
    assert not np.array_equal(draw(7), draw(8))
    assert not np.array_equal(draw(-7), draw(7))
    assert not np.array_equal(draw(1), draw(1 + (1 << 40)))
    assert not np.array_equal(draw(1), draw(1 - (1 << 63)))


def test_a_seed_on_a_grid_with_an_empty_ball_is_refused(g):
    """test/regressions/05-edge-cases-05.jl: a grid whose dealias ball holds no nonzero mode must
    not silently seed a zero field; the requested energy is realized otherwise."""
    with pytest.raises(ValueError, match="no nonzero mode"):
        g.random_phase_velocity(2, energy=0.5)
    with pytest.raises(ValueError, match="no nonzero mode"):
        g.random_phase_scalar(2, variance=0.5)
    u = g.random_phase_velocity(16, energy=0.37, seed=3)
    assert g.seeding.parseval_energy(u, 16) == pytest.approx(0.37, rel=1e-14)


def test_integral_length_ignores_a_uniform_mean_flow(g):
    """test/regressions/04-r5-free-modules-04.jl: the kappa = 0 shell is out of both sums, so a
    mean-flow-only field is refused rather than given L11 = 0, and adding a mean flow leaves L11
    as it was.  L11 of energy on one shell kappa is 3 pi / (4 kappa)."""
    kappa = np.arange(6.0)
    with pytest.raises(ValueError, match="no energy"):
        g.integral_length((kappa, np.array([0.7, 0, 0, 0, 0, 0])))
    e = np.array([0.0, 0.2, 0.5, 0.1, 0.05, 0.0])
    plain = g.integral_length((kappa, e))
    e_mean = e.copy()
    e_mean[0] = 3.0
    assert g.integral_length((kappa, e_mean)) == plain
    assert g.integral_length((kappa, np.array([0, 0, 1.0, 0, 0, 0]))) == pytest.approx(3 * math.pi / 8, rel=1e-15)


@pytest.mark.parametrize("bad", [0.0, -1.0, math.inf, math.nan])
def test_predicted_viscosity_refuses_a_bad_resolution_criterion_by_name(g, bad):
    """test/regressions/04-r4-free-hot-path-04.jl: kappa_max eta must be finite and positive."""
    dom = g.FixedDomain(grid_points_per_axis=16, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float64)
    with pytest.raises(ValueError, match="kmax_eta"):
        g.predicted_viscosity(dom, 0.1, kmax_eta=bad)
    with pytest.raises(ValueError, match="dissipation"):
        g.predicted_viscosity(dom, bad if bad != 0.0 else -0.0, kmax_eta=1.5)


def _amplification_2n(scheme, z):
    """R(z) of a 2N scheme on y' = z y over one unit step: h <- a_i h + z q, q <- q + b_i h."""
    a, b, _ = dns.scheme_coefficients(scheme)
    q, h = 1.0 + 0j, 0j
    for ai, bi in zip(a, b):
        h = ai * h + z * q
        q = q + bi * h
    return q


def test_the_cfl_ceiling_keeps_the_advective_rate_inside_the_stability_limit(g):
    """test/observables_analytic.jl: C kappa_max dx is the advective rate times the step.  For the
    three-stage third-order schemes it stays below sqrt(3), the imaginary-axis limit they share; for
    IF-RK4 below the Carpenter-Kennedy scheme's own limit, found here from its coefficients.  At
    that rate the amplification factor of the scheme's own coefficients is at most 1."""
    dom = g.FixedDomain(grid_points_per_axis=64, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float64)
    dx = dom.grid_spacing

    def rate(scheme):
        return g.recommended_cfl(scheme, dom) * dom.max_retained_wavenumber() * dx

    def ssp3(z):
        return 1 + z + z * z / 2 + z ** 3 / 6

    def limit(amp):
        lo, hi = 0.0, 4.0
        for _ in range(60):
            mid = 0.5 * (lo + hi)
            ok = all(abs(amp(1j * y)) <= 1 + 1e-12 for y in np.linspace(0, mid, 400))
            lo, hi = (mid, hi) if ok else (lo, mid)
        return lo

    rk3 = limit(lambda z: _amplification_2n("rk3", z))
    assert rk3 == pytest.approx(math.sqrt(3), rel=1e-6)          # every three-stage third-order scheme
    assert limit(ssp3) == pytest.approx(math.sqrt(3), rel=1e-6)
    for scheme, lim in ((g.IFLowStorageRK3, rk3), (g.IFSSPRK3, math.sqrt(3)),
                        (g.IFLowStorageRK4, limit(lambda z: _amplification_2n("rk4", z)))):
        assert rate(scheme) < lim, scheme
    assert abs(_amplification_2n("rk3", 1j * rate(g.IFLowStorageRK3))) <= 1
    assert abs(_amplification_2n("rk4", 1j * rate(g.IFLowStorageRK4))) <= 1


def test_a_stored_trail_sample_that_does_not_fit_is_refused_by_number(g, tmp_path):
    """read_simulation refuses a trail sample whose shape is not (3, M) before any run is built
    (ext/GBGHomogeneousTurbulenceHDF5Ext.jl:268-283); a fitting trail comes back newest first."""
    from gbg_amd import checkpoint as ck
    s = ck.Store()
    tr = ck._tagged(s.root, "velocity_at_particles_trail", "ParticleTrail")
    tr.attrs["times"] = [0.0, 0.1]
    a, b = np.arange(12.0).reshape(3, 4), -np.arange(12.0).reshape(3, 4)
    tr["sample1"], tr["sample2"] = a, b
    s.save(tmp_path / "t.npz")
    trail = ck.Store.load(tmp_path / "t.npz").root.group("velocity_at_particles_trail", "t")
    got = ck.stored_trail_samples(trail, 4)
    assert got.shape == (2, 3, 4) and np.array_equal(got[0], b) and np.array_equal(got[1], a)
    with pytest.raises(ValueError, match="stored sample 1 does not fit"):
        ck.stored_trail_samples(trail, 5)
    tr["sample2"] = np.zeros((4, 3))   # an (M, 3) array must not be read as (3, M)
    s.save(tmp_path / "u.npz")
    trail = ck.Store.load(tmp_path / "u.npz").root.group("velocity_at_particles_trail", "t")
    with pytest.raises(ValueError, match="stored sample 2 does not fit"):
        ck.stored_trail_samples(trail, 4)


def test_a_fixed_dt_under_a_shortened_multistep_start_is_refused(g):
    """src/SimulationRun.jl:1158-1161: a fixed dt and shortened_first_steps are two rules for the
    same steps; a multistep particle scheme refuses the pair before anything moves."""
    sim = g.Simulation.__new__(g.Simulation)   # the guard runs before any device call
    sim.shortened_first_steps = True
    sim.particle_integration_scheme = g.AB(3)
    with pytest.raises(ValueError, match="shortened_first_steps = true"):
        sim.advance(0.01)
    sim.particle_integration_scheme = g.ABM(2)
    with pytest.raises(ValueError, match="drop `dt`"):
        sim.advance(0.01)


@pytest.mark.parametrize("n", [64, 48])
def test_the_seeded_spectrum_follows_k4_exp_gaussian_with_half_unit_energy(g, n):
    """The benchmark seed: every mode off the kx = 0 and Nyquist planes carries |u_hat|^2 in exact
    proportion to k^2 exp(-2 (k/4)^2), i.e. E(k) ~ k^4 exp(-2 (k/4)^2) once the 4 pi k^2 modes of a
    shell are summed; the field is solenoidal, zero-mean, inside the closed 2/3 ball, and its kinetic
    energy is 1/2.  The shell spectrum, its log-slope and its peak follow the same law."""
    u = g.random_phase_velocity(n, seed=2026)
    d = dns.Domain(n)
    shape = (n, n, n // 2 + 1)                                     # [z][y][x]
    mx = np.broadcast_to(d.mx[None, None, :], shape)
    my = np.broadcast_to(d.my[None, :, None], shape)
    mz = np.broadcast_to(d.mz[:, None, None], shape)
    k2 = (mx * mx + my * my + mz * mz).astype(np.float64)
    e_mode = np.sum(np.abs(u) ** 2, axis=0)
    assert np.all(e_mode[~d.mask] == 0) and e_mode[0, 0, 0] == 0
    div = mx * u[0] + my * u[1] + mz * u[2]
    assert np.max(np.abs(div)) < 1e-12 * np.max(np.abs(u))
    law = k2 * np.exp(-2.0 * k2 / 16.0)
    interior = d.mask & (mx > 0) & (mx < n // 2)
    ratio = e_mode[interior] / law[interior]
    assert np.max(np.abs(ratio / ratio[0] - 1)) < 1e-12           # exact, mode by mode
    from gbg_amd.seeding import parseval_energy
    assert parseval_energy(u, n) == pytest.approx(0.5, rel=1e-14)
    # the shell spectrum E(kappa_n), shells n - 1/2 <= |m| < n + 1/2, Parseval weights
    w = np.where((mx == 0) | (mx == n // 2), 1.0, 2.0)
    shell = np.floor(np.sqrt(k2) + 0.5).astype(int)
    E = np.bincount(shell.ravel(), weights=(0.5 * w * e_mode).ravel()) / float(n) ** 6
    assert np.sum(E) == pytest.approx(0.5, rel=1e-13)
    # against the law integrated over each shell, C * int_{n-1/2}^{n+1/2} k^4 exp(-2 (k/4)^2) dk: the
    # per-shell ratio carries only the lattice's mode count (a shell holds about, not exactly,
    # 4 pi n^2 modes), so a fit of ln(E/I) = c + a ln n + b (n/4)^2 is flat: a and b near 0.  A
    # k^2 law would give a = -2, a peak at 3 would give b = 0.78.
    from scipy.integrate import quad
    kk = np.arange(1, 2 * n // 7)                                 # whole shells inside the ball
    law_k = lambda k: k ** 4 * math.exp(-2.0 * (k / 4.0) ** 2)
    integ = np.array([quad(law_k, k - 0.5, k + 0.5)[0] for k in kk])
    ratio = E[kk] / integ
    assert np.max(np.abs(ratio / np.exp(np.mean(np.log(ratio))) - 1)) < 0.4
    fit = np.linalg.lstsq(np.stack([np.ones(len(kk)), np.log(kk), (kk / 4.0) ** 2], 1), np.log(ratio), rcond=None)[0]
    assert abs(fit[1]) < 0.3 and abs(fit[2]) < 0.05
    assert int(np.argmax(E)) == 4                                 # the peak at k_p = 4


def test_the_dealiasing_strategies_fix_the_bands_the_cfl_and_the_viscosity(g):
    """src/SimulationDomain.jl:181-195,306-457 and src/Propagator.jl:919: the retained and resolved
    Nyquist fractions of TwoThirds (2/3, 2/3), HouLi (1, 4/5) and PhaseShift (2 sqrt2/3 both); the
    CFL ceiling scaled by (2/3)/f_ret; the viscosity from the resolved band.  The published HouLi
    comparison (docs/src/validation.md:29-46): 128^3, eps = 1, k_max eta = 1.5 on the resolved
    edge 0.8 k_Ny = 51.2 gives nu = 9.032e-3."""
    two, hl, ps = g.TwoThirds(), g.HouLi(), g.PhaseShift()
    doms = {k: g.FixedDomain(grid_points_per_axis=128, side_length=L, dealiasing=v, element_type=np.float64)
            for k, v in (("two", two), ("hl", hl), ("ps", ps))}
    assert doms["two"].max_retained_wavenumber() == pytest.approx(128 / 3, rel=1e-15)
    assert doms["hl"].max_retained_wavenumber() == 64.0 and doms["hl"].max_resolved_wavenumber() == pytest.approx(51.2)
    assert doms["ps"].max_retained_wavenumber() == pytest.approx(2 * math.sqrt(2) / 3 * 64, rel=1e-15)
    assert g.recommended_cfl(g.IFLowStorageRK4, doms["two"]) == 0.96
    assert g.recommended_cfl(g.IFSSPRK3, doms["hl"]) == pytest.approx(0.5 * 2 / 3, rel=1e-15)
    assert g.recommended_cfl(g.IFSSPRK3, doms["ps"]) == pytest.approx(0.5 / math.sqrt(2), rel=1e-15)
    nu = g.predicted_viscosity(doms["hl"], 1.0, kmax_eta=1.5)
    assert nu == pytest.approx(9.032e-3, rel=1e-3)
    assert nu == pytest.approx((1.5 / 51.2) ** (4 / 3), rel=1e-12)
    assert hl.strength_for(np.float64) == pytest.approx(36.04365338911715, rel=1e-15)
    assert hl.strength_for(np.float32) == pytest.approx(15.942385152878742, rel=1e-15)
    # s = 36 (1 - 2/N)^-36 at N = 128 matches TurTLE's (N/2 - 1) normalisation: 63.463
    assert 36 * (1 - 2 / 128) ** -36 == pytest.approx(63.463, abs=5e-4)
    for bad in (0, -3, 2.5):
        with pytest.raises(ValueError, match="order"):
            g.HouLi(order=bad)
    with pytest.raises(NotImplementedError, match="HouLi and PhaseShift"):
        g.FixedDomain(grid_points_per_axis=512, side_length=L, dealiasing=hl, element_type=np.float32)


def test_truncate_to_band_under_each_strategy(g):
    """truncate_to_band! applies the dealias kernel (src/SimulationDomain.jl:520-537): the closed
    balls 9|m|^2 <= N^2 (TwoThirds) and <= 2 N^2 (PhaseShift), and the HouLi filter's multiply."""
    for dl, kind in ((g.TwoThirds(), "two_thirds"), (g.PhaseShift(), "phase_shift"), (g.HouLi(order=8, strength=3.0), "hou_li")):
        dom = g.FixedDomain(grid_points_per_axis=32, side_length=L, dealiasing=dl, element_type=np.float64)
        f32 = np.ones((32, 32, 17), dtype=complex)
        got = g.truncate_to_band(f32, dom)
        od = dns.Domain(32, dealiasing=kind, filter_order=8, filter_strength=3.0)
        assert np.allclose(got, od.dealias, rtol=1e-15, atol=0)
