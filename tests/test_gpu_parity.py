"""Parity of the HIP step with the FP64 oracle, through the C ABI, on an MI355X.

Bars: integer tables and cell keys bit-exact; FP64 fields to 1e-12 relative; FP32-bulk fields to
the FP32 transform floor (about 1e-6 relative per step), and the velocity within 1e-5 relative
of the FP64 oracle after 10 IF-RK4 steps.
"""
import math
import time

import numpy as np
import pytest

from oracle import dns
from oracle import tracers as trc

pytestmark = pytest.mark.gpu
L = 2 * math.pi


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def make_sim(g, n, dtype, *, scheme=None, forcing=None, G=1.0, nu=0.02, seed=7, tracers=None,
             keep_vorticity=False, scalar=True):
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=dtype)
    u = g.random_phase_velocity(n, seed=seed)
    p = g.random_phase_scalar(n, seed=seed + 1) if scalar else None
    vf = g.ConstantPower(dom, injection_rate=forcing, lowest_shell=1, highest_shell=2) if forcing else g.NoForcing()
    cond = g.FlowConditions(viscosity=nu, schmidt=1.0, velocity_forcing=vf)
    sim = g.Simulation(dom, conditions=cond, state=g.ResolvedFields(u, p), mean_gradient=G if scalar else 0.0,
                       field_integration_scheme=scheme or g.IFLowStorageRK4, tracers=tracers,
                       keep_vorticity=keep_vorticity)
    ocond = dns.Conditions(viscosity=nu, schmidt=1.0, forcing=(float(dtype(forcing)), 1, 2) if forcing else None,
                           mean_gradient=G if scalar else 0.0, band_eps=float(np.finfo(dtype).eps))
    return sim, u, p, ocond


@pytest.mark.parametrize("n", [16, 48, 64])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_frame_tables_bit_exact(g, n, dtype):
    sim, *_ = make_sim(g, n, dtype)
    with sim:
        mx, my, mz, mask = sim.frame_tables()
    d = dns.Domain(n)
    assert np.array_equal(mx, d.mx) and np.array_equal(my, d.my) and np.array_equal(mz, d.mz)
    assert np.array_equal(mask.astype(bool), d.mask)


@pytest.mark.parametrize("n", [32, 48])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 3e-6)])
def test_deposit_matches_the_oracle(g, n, dtype, tol):
    sim, u, p, oc = make_sim(g, n, dtype, keep_vorticity=True, forcing=0.1)
    d = dns.Domain(n)
    with sim:
        dep = dns.deposit(d, u, p)
        assert rel(sim.physical(0), dep.u) < tol
        assert rel(sim.physical(2), dep.w) < tol
        assert rel(sim.physical(1), dep.phi) < tol
        assert sim.max_speed() == pytest.approx(dep.max_speed, rel=tol)
        r = sim.reductions()
        want = dns.reductions(d, oc, u, p)
        for k in ("kinetic_energy", "dissipation", "scalar_variance", "scalar_dissipation"):
            assert r[k] == pytest.approx(want[k], rel=10 * tol if dtype is np.float32 else 1e-13), k
        band = 0.5 * dns.weighted_inner(d, u, d.shell_band(1, 2))
        assert r["band_energy"] == pytest.approx(band, rel=10 * tol if dtype is np.float32 else 1e-13)
        kap, e = sim.energy_spectrum()
        _, eo = dns.energy_spectrum(d, u)
        assert np.max(np.abs(e[: len(eo)] - eo)) <= 10 * tol * np.max(eo) + 1e-300


def _scheme(g, name):
    return {"rk3": g.IFLowStorageRK3, "rk4": g.IFLowStorageRK4, "ssprk3": g.IFSSPRK3, "heun": g.IFHeun}[name]


@pytest.mark.parametrize("n", [32, 48, 96])
@pytest.mark.parametrize("scheme", ["rk3", "rk4", "ssprk3", "heun"])
@pytest.mark.parametrize("forcing", [None, 0.1])
def test_one_step_fp64_matches_the_oracle(g, n, scheme, forcing):
    """48 and 96 carry radix-3 passes, and 3 | N puts the boundary shell 9|m|^2 = N^2 inside the
    ball (test/regressions/r6-inclusive-ball-exact.jl)."""
    sch = _scheme(g, scheme)
    sim, u, p, oc = make_sim(g, n, np.float64, scheme=sch, forcing=forcing)
    d = dns.Domain(n)
    with sim:
        sim.advance(2e-3)
        got = sim.state()
        uo, po, dep = dns.step(d, oc, scheme, u, p, 2e-3)
        assert rel(got.velocity_spectral, uo) < 1e-12
        assert rel(got.scalar_spectral, po) < 1e-12
        assert sim.max_speed() == pytest.approx(dep.max_speed, rel=1e-12)
        assert rel(sim.physical(0), dep.u) < 1e-12


def test_ten_steps_fp64_and_fp32_bulk_against_the_fp64_oracle(g):
    """The north-star accuracy bar: FP32-bulk velocity within 1e-5 of the FP64 oracle after 10 steps."""
    n, steps, dt = 64, 10, 2e-3
    d = dns.Domain(n)
    sim64, u, p, oc = make_sim(g, n, np.float64, forcing=0.1, nu=0.01)
    sim32, *_ = make_sim(g, n, np.float32, forcing=0.1, nu=0.01)
    uo, po, dep = u, p, None
    errs = []
    with sim64, sim32:
        for s in range(steps):
            uo, po, dep = dns.step(d, oc, "rk4", uo, po, dt, dep)
            sim64.advance(dt)
            sim32.advance(dt)
            s32 = sim32.state()
            errs.append((rel(s32.velocity_spectral, uo), rel(s32.scalar_spectral, po)))
        s64 = sim64.state()
        print("FP32-bulk rel. error per step (u, phi):", ["%.2e/%.2e" % e for e in errs], flush=True)
        assert rel(s64.velocity_spectral, uo) < 1e-11
        assert rel(s64.scalar_spectral, po) < 1e-11
        assert errs[-1][0] <= 1e-5
        assert errs[-1][1] <= 1e-5
        _, e32 = sim32.energy_spectrum()
        _, eo = dns.energy_spectrum(d, uo)
        live = eo > 1e-12 * eo.max()
        assert np.max(np.abs(e32[: len(eo)][live] - eo[live]) / eo[live]) < 1e-4


def test_upload_rejects_content_outside_the_ball(g):
    n = 16
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float64)
    u = g.random_phase_velocity(n, seed=1)
    u[0, 0, 0, 7] = 1e-3                                  # mode (7, 0, 0): 9*49 > 256
    with pytest.raises(ValueError, match="outside the 2/3"):
        g.Simulation(dom, conditions=g.FlowConditions(viscosity=0.01, schmidt=1.0), state=g.ResolvedFields(u, None))


def test_no_scalar_run_matches_the_oracle(g):
    n = 32
    sim, u, _, oc = make_sim(g, n, np.float64, scalar=False, forcing=0.1)
    d = dns.Domain(n)
    with sim:
        sim.advance(2e-3)
        uo, _, _ = dns.step(d, oc, "rk4", u, None, 2e-3)
        assert rel(sim.state().velocity_spectral, uo) < 1e-12


@pytest.mark.parametrize("n", [32, 48])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-11), (np.float32, 5e-5)])
def test_tracers_match_the_oracle(g, n, dtype, tol):
    m, dt = 500, 2e-3
    pos = trc.uniform_positions(m, L, 5).T.copy()          # [3, M]
    pos = pos.astype(dtype).astype(np.float64)
    sim, u, p, oc = make_sim(g, n, dtype, tracers=g.Tracers(
        g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=dtype),
        positions=pos))
    d = dns.Domain(n)
    with sim:
        dep = dns.deposit(d, u, p)
        v, s = sim.tracer_samples()
        assert rel(v, trc.interpolate(dep.u, pos.T, L).T) < tol
        assert rel(s, trc.interpolate(dep.phi, pos.T, L)) < tol
        X = pos.T.copy()
        vo = trc.interpolate(dep.u, X, L)
        uo, po = u, p
        for step in range(1, 11):
            uo, po, dep = dns.step(d, oc, "rk4", uo, po, dt, dep)
            X, vo = trc.heun_advect(X, vo, dep.u, dt, L)
            sim.advance(dt)
            if step in (1, 10):   # positions and sampled u, phi after 1 and 10 steps
                got = sim.tracer_positions()
                assert np.max(np.abs(got - X.T)) < 10 * tol
                v, s = sim.tracer_samples()
                assert rel(v, vo.T) < 10 * tol
                assert rel(s, trc.interpolate(dep.phi, X, L)) < 10 * tol
        # cell keys: the device's integer contract, bit-exact from the device's own positions
        keys = sim.tracer_cells()
        assert np.array_equal(keys, trc.cell_keys(got.T, n, L, dtype))
        sim._call("hdns_rebuild_traversal")
        order = sim.traversal_order()
        assert np.array_equal(np.sort(order), np.arange(m))
        tiles = (keys // (n * n) // 8 * (n // 8) + (keys // n) % n // 8) * (n // 8) + keys % n // 8
        # by tile, and within a tile by index: the stable order of the tile keys
        assert np.array_equal(order, np.argsort(tiles, kind="stable"))


def test_a_step_at_256_keeps_the_state_solenoidal_real_and_energy_consistent(g):
    n = 256
    sim, u, p, oc = make_sim(g, n, np.float32, forcing=0.1, nu=2.121e-3)
    d = dns.Domain(n)
    with sim:
        e_spec = sim.reductions()["kinetic_energy"]
        uphys = sim.physical(0)
        e_phys = 0.5 * float(np.mean(np.sum(uphys.astype(np.float64) ** 2, axis=0)))
        assert e_phys == pytest.approx(e_spec, rel=1e-5)
        assert e_spec == pytest.approx(0.5, rel=1e-6)
        sim.advance(1e-3)
        st = sim.state().velocity_spectral
        assert dns.divergence_residual(d, st) < 1e-6
        assert np.array_equal(dns.reality_projection(d, st), st)
        assert np.all(st[:, 0, 0, 0] == 0)


@pytest.mark.parametrize("scalar", [True, False])
def test_fp32_bulk_at_512_agrees_with_the_fp64_step(g, scalar):
    """At 512^3 the FP32 x pass runs on one wavefront per row pair with the data in registers; the
    FP64 context runs the LDS transforms that the oracle pins at the smaller sizes.  One IF-RK4 step
    of each from the same state agrees to the FP32 transform floor, spectrally and physically."""
    n, dt = 512, 1e-3
    a, *_ = make_sim(g, n, np.float32, forcing=0.1, nu=8.42e-4, scalar=scalar)
    b, *_ = make_sim(g, n, np.float64, forcing=0.1, nu=8.42e-4, scalar=scalar)
    with a, b:
        if not scalar:   # the FP32 run on the y passes' second radix plan (16.8.4)
            a.schedule((1, 0, 0, 0, 1))
            a._call("hdns_deposit")
        for k in ("kinetic_energy", "band_energy"):
            assert a.reductions()[k] == pytest.approx(b.reductions()[k], rel=1e-6)
        assert rel(a.physical(0), b.physical(0)) < 2e-6
        u0 = b.state().velocity_spectral
        a.advance(dt)
        b.advance(dt)
        assert a.max_speed() == pytest.approx(b.max_speed(), rel=2e-6)
        sa, sb = a.state(), b.state()
        err = rel(sa.velocity_spectral, sb.velocity_spectral)
        # the step's increment (nonlinear, forcing and viscous terms) against the FP32 rounding of u
        derr = rel(sa.velocity_spectral - u0, sb.velocity_spectral - u0)
        print(f"n=512 scalar={scalar}: velocity rel err {err:.3e}, increment rel err {derr:.3e}", flush=True)
        assert err < 3e-6
        assert derr < 2e-3
        if scalar:
            assert rel(sa.scalar_spectral, sb.scalar_spectral) < 3e-6
        assert rel(a.physical(0), b.physical(0)) < 3e-6


@pytest.mark.parametrize("n,dtype,tol", [(512, np.float32, 1e-6), (256, np.float32, 1e-6), (256, np.float64, 1e-14)])
@pytest.mark.parametrize("scalar", [True, False])
def test_x_pass_fetch_schedules_agree(g, n, dtype, tol, scalar):
    """The register-resident x pass (FP32 at 512: row pairs of 512 points in radix-8 passes; FP32
    and FP64 at 256: row pairs of 256 points in radix-4 passes) reads each component's rows either from their copy staged in
    LDS by direct-to-LDS loads (x schedule 0) or straight from HBM (x schedule 1).  The values are
    the same; the two builds may contract different multiply-adds, so two steps, the physical
    fields and max|u| agree to rounding (a misplaced row would be an O(1) error)."""
    dt = 1e-3
    a, *_ = make_sim(g, n, dtype, forcing=0.1, nu=8.42e-4, scalar=scalar)
    b, *_ = make_sim(g, n, dtype, forcing=0.1, nu=8.42e-4, scalar=scalar)
    with a, b:
        a.schedule((0, 0, 0, 0, 0))
        b.schedule((0, 0, 0, 1, 0))
        for _ in range(2):
            a.advance(dt)
            b.advance(dt)
        sa, sb = a.state(), b.state()
        eu = rel(sa.velocity_spectral, sb.velocity_spectral)
        ep = rel(sa.scalar_spectral, sb.scalar_spectral) if scalar else 0.0
        ex = rel(a.physical(0), b.physical(0))
        print(f"x fetch schedules at {n} {np.dtype(dtype).name}, scalar={scalar}: u {eu:.2e} phi {ep:.2e} "
              f"physical {ex:.2e}", flush=True)
        assert eu < tol and ep < tol and ex < tol
        assert a.max_speed() == pytest.approx(b.max_speed(), rel=tol)


def test_graph_captured_step_replays_the_eager_step_bitwise(g):
    """Replays interleaved with eager work and host reads; every replay equals the eager step."""
    n, dt = 32, 2e-3
    eager, *_ = make_sim(g, n, np.float32, forcing=0.1)
    graph, *_ = make_sim(g, n, np.float32, forcing=0.1)
    with eager, graph:
        for _ in range(4):
            eager.step_fields(dt)
            graph.step_fields(dt, graph=True)
            a, b = eager.state(), graph.state()
            assert np.array_equal(a.velocity_spectral, b.velocity_spectral)
            assert np.array_equal(a.scalar_spectral, b.scalar_spectral)
            assert eager.max_speed() == graph.max_speed()
            assert eager.reductions() == graph.reductions()
        # a new dt re-captures
        eager.step_fields(dt / 2)
        graph.step_fields(dt / 2, graph=True)
        assert np.array_equal(eager.state().velocity_spectral, graph.state().velocity_spectral)


def test_autotune_leaves_the_state_bitwise_unchanged_and_schedules_agree(g):
    n, dt = 64, 2e-3
    sim, *_ = make_sim(g, n, np.float32, forcing=0.1)
    ref, *_ = make_sim(g, n, np.float32, forcing=0.1)
    with sim, ref:
        before = sim.state().velocity_spectral.copy()
        report = sim.autotune(reps=2)
        assert set(report) == {"y_passes", "z_forward", "z_inverse", "x_pass", "y_radix_plan"}
        assert np.array_equal(sim.state().velocity_spectral, before)
        # the chosen z tiling sets the block order of the FP64 partial sums, so equal to rounding
        ra, rb = sim.reductions(), ref.reductions()
        assert ra.keys() == rb.keys()
        for key in ra:
            assert abs(ra[key] - rb[key]) <= 1e-13 * abs(rb[key]), key
        for sched in [(0, 0, 0, 0, 0), (1, 1, 1, 1, 1), (0, 1, 0, 1, 0)]:
            sim.schedule(sched)
            assert sim.schedule() == sched
        sim.schedule((0, 0, 0, 0, 0))
        ref.schedule((1, 1, 1, 1, 1))
        sim.step_fields(dt)
        ref.step_fields(dt)
        a, b = sim.state(), ref.state()
        assert rel(a.velocity_spectral, b.velocity_spectral) < 1e-6
        assert rel(a.scalar_spectral, b.scalar_spectral) < 1e-6


@pytest.mark.parametrize("n,dtype", [(16, np.float64), (48, np.float64), (64, np.float32), (96, np.float64),
                                     (128, np.float32), (256, np.float32)])
@pytest.mark.parametrize("sched", [(1, 0, 2, 0, 0), (1, 2, 0, 0, 0)], ids=["z_inverse_wide", "z_forward_streamed"])
def test_z_pass_staging_schedules_are_bitwise_schedule_0(g, n, dtype, sched):
    """The z passes' third schedules move the same values through LDS: the z inverse writes each
    transformed tile back 16 bytes per lane, the z forward's scalar half streams each flux tile
    into LDS by direct-to-LDS loads while the previous one transforms.  So one forced step with the
    scalar (and the deposit it ends with) is bitwise schedule 0's; and it matches the oracle."""
    sim, u, p, oc = make_sim(g, n, dtype, forcing=0.1)
    ref, *_ = make_sim(g, n, dtype, forcing=0.1)
    with sim, ref:
        sim.schedule(sched)
        ref.schedule((1, 0, 0, 0, 0))
        for s in (sim, ref):
            s._call("hdns_deposit")
            s.advance(2e-3)
        a, b = sim.state(), ref.state()
        assert np.array_equal(a.velocity_spectral, b.velocity_spectral)
        assert np.array_equal(a.scalar_spectral, b.scalar_spectral)
        assert sim.reductions() == ref.reductions()
        if dtype == np.float64:
            uo, po, _ = dns.step(dns.Domain(n), oc, "rk4", u, p, 2e-3)
            assert rel(a.velocity_spectral, uo) < 1e-12
            assert rel(a.scalar_spectral, po) < 1e-12


@pytest.mark.parametrize("n", [16, 32, 48, 64, 96, 128, 256])
def test_z_inverse_on_column_pairs_matches_schedule_0_and_the_oracle(g, n):
    """The z inverse's fourth schedule (FP32: two adjacent columns per transform element, the
    FP64 transforms' shape): one forced FP32 step with the scalar agrees with schedule 0's to FP32
    rounding (the FP64 partial sums add the modes in another order) and with the FP64 oracle to
    the FP32 bar; its deposit's reductions agree with schedule 0's to 1e-12."""
    sim, u, p, oc = make_sim(g, n, np.float32, forcing=0.1)
    ref, *_ = make_sim(g, n, np.float32, forcing=0.1)
    with sim, ref:
        sim.schedule((1, 0, 3, 0, 0))
        ref.schedule((1, 0, 0, 0, 0))
        for s in (sim, ref):
            s._call("hdns_deposit")
        ra, rb = sim.reductions(), ref.reductions()
        for key in ra:
            assert abs(ra[key] - rb[key]) <= 1e-12 * abs(rb[key]) + 1e-300, key
        for s in (sim, ref):
            s.advance(2e-3)
        a, b = sim.state(), ref.state()
        assert rel(a.velocity_spectral, b.velocity_spectral) < 1e-6
        assert rel(a.scalar_spectral, b.scalar_spectral) < 1e-6
        uo, po, _ = dns.step(dns.Domain(n), oc, "rk4", u, p, 2e-3)
        assert rel(a.velocity_spectral, uo) < 3e-6
        assert rel(a.scalar_spectral, po) < 3e-6


@pytest.mark.parametrize("n", [16, 32, 48, 64, 96, 128])
def test_y_radix_plan_led_by_radix_16_matches_the_oracle(g, n):
    """The y passes' second radix plan (16.2, 16.4, 16.3, 16.8, 16.2.3 ..., 2.8 at 16) in one FP64
    forced step with the scalar, against the oracle."""
    sim, u, p, oc = make_sim(g, n, np.float64, forcing=0.1)
    with sim:
        sim.schedule((1, 0, 0, 0, 1))
        sim._call("hdns_deposit")
        sim.advance(2e-3)
        uo, po, _ = dns.step(dns.Domain(n), oc, "rk4", u, p, 2e-3)
        st = sim.state()
        assert rel(st.velocity_spectral, uo) < 1e-12
        assert rel(st.scalar_spectral, po) < 1e-12


def test_y_radix_plans_agree_at_256(g):
    """16.16 (one LDS exchange) against 8.8.4 at 256^3, FP64, one forced step."""
    a, *_ = make_sim(g, 256, np.float64, forcing=0.1)
    b, *_ = make_sim(g, 256, np.float64, forcing=0.1)
    with a, b:
        a.schedule((1, 0, 0, 0, 0))
        b.schedule((1, 0, 0, 0, 1))
        b._call("hdns_deposit")
        a.advance(1e-3)
        b.advance(1e-3)
        assert rel(b.state().velocity_spectral, a.state().velocity_spectral) < 1e-13


def _member(g, n, dtype, nu, seed, m=0, unwrapped=False):
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=dtype)
    cond = g.FlowConditions(viscosity=nu, schmidt=1.0)
    state = g.ResolvedFields(g.random_phase_velocity(n, seed=seed), g.random_phase_scalar(n, seed=seed + 1))
    tr = g.Tracers(dom, number_of_particles=m, seed=seed + 2, unwrapped=unwrapped) if m else None
    return g.EnsembleMember(dom, cond, state, tr)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_ensemble_boxes_equal_their_standalone_runs_bitwise(g, dtype):
    """Uneven sizes (3 | 48 included) and viscosities in one ensemble, IF-RK3, graph-captured: each
    box's state, reductions, max|u| and tracers equal a standalone run of that box bit for bit."""
    sizes = [32, 48, 32, 64, 48]
    nus = [0.005, 0.012, 0.02, 0.008, 0.05]
    m, dt, steps = 64, 2e-3, 3
    members = [_member(g, n, dtype, nu, 100 + 7 * b, m) for b, (n, nu) in enumerate(zip(sizes, nus))]
    with g.Ensemble(members, field_integration_scheme=g.IFLowStorageRK3, mean_gradient=1.0) as ens:
        for s in range(steps):
            ens.step(dt, graph=s != 1)   # graph replays around one eager step
            ens.advance_tracers(dt)
        for b, mb in enumerate(members):
            alone = g.Simulation(mb.domain, conditions=mb.conditions, state=mb.state, tracers=mb.tracers,
                                 field_integration_scheme=g.IFLowStorageRK3, mean_gradient=1.0)
            with alone:
                for _ in range(steps):
                    alone.advance(dt, check=False)
                box = ens.box(b)
                a, e = alone.state(), box.state()
                print(f"box {b} ({sizes[b]}^3, nu {nus[b]}): max|du| "
                      f"{np.max(np.abs(a.velocity_spectral - e.velocity_spectral)):.3e}", flush=True)
                assert np.array_equal(a.velocity_spectral, e.velocity_spectral), b
                assert np.array_equal(a.scalar_spectral, e.scalar_spectral), b
                assert alone.reductions() == box.reductions(), b
                assert alone.max_speed() == box.max_speed(), b
                assert np.array_equal(alone.tracer_positions(), box.tracer_positions()), b
                va, sa = alone.tracer_samples()
                ve, se = box.tracer_samples()
                assert np.array_equal(va, ve) and np.array_equal(sa, se), b


def test_ensemble_tracers_under_ab3_equal_their_standalone_runs_bitwise(g):
    """The multistep push over each batch's trail: 4 steps of AB(3) in a mixed ensemble equal the
    standalone runs bit for bit, positions, the unwrapped record and the whole trail."""
    sizes, nus = [32, 48, 32], [0.01, 0.02, 0.03]
    m, dt, steps = 64, 2e-3, 4
    members = [_member(g, n, np.float32, nu, 300 + b, m, unwrapped=True) for b, (n, nu) in enumerate(zip(sizes, nus))]
    with g.Ensemble(members, field_integration_scheme=g.IFLowStorageRK3, mean_gradient=1.0,
                    particle_integration_scheme=g.AB(3)) as ens:
        for s in range(steps):
            ens.step(dt)
            ens.advance_tracers(dt)
        for b, mb in enumerate(members):
            with g.Simulation(mb.domain, conditions=mb.conditions, state=mb.state, tracers=mb.tracers,
                              field_integration_scheme=g.IFLowStorageRK3, mean_gradient=1.0,
                              particle_integration_scheme=g.AB(3)) as alone:
                for _ in range(steps):
                    alone.advance(dt, check=False)
                box = ens.box(b)
                assert np.array_equal(alone.tracer_positions(), box.tracer_positions()), b
                assert np.array_equal(alone.unwrapped_positions(), box.unwrapped_positions()), b
                va, _ = alone.tracer_samples()
                ve, _ = box.tracer_samples()
                assert np.array_equal(va, ve), b


def test_ensemble_box_at_48_matches_the_fp64_oracle(g):
    """A 48^3 box of a mixed ensemble (FP64, IF-RK3) against the oracle after two steps."""
    members = [_member(g, 32, np.float64, 0.01, 11), _member(g, 48, np.float64, 0.02, 12)]
    d = dns.Domain(48)
    oc = dns.Conditions(viscosity=0.02, schmidt=1.0, forcing=None, mean_gradient=1.0, band_eps=float(np.finfo(np.float64).eps))
    uo, po = members[1].state.velocity_spectral, members[1].state.scalar_spectral
    dep = None
    with g.Ensemble(members, field_integration_scheme=g.IFLowStorageRK3, mean_gradient=1.0) as ens:
        for _ in range(2):
            ens.step(2e-3)
            uo, po, dep = dns.step(d, oc, "rk3", uo, po, 2e-3, dep)
        got = ens.box(1).state()
        assert rel(got.velocity_spectral, uo) < 1e-12
        assert rel(got.scalar_spectral, po) < 1e-12


def test_state_distance_matches_host_norms(g):
    """hdns_state_distance: the RMS difference of two runs' states and the second's RMS, against the
    same norms of the downloaded spectra (Parseval weights over the half spectrum)."""
    n = 32
    a, *_ = make_sim(g, n, np.float32, forcing=0.1)
    b, *_ = make_sim(g, n, np.float64, forcing=0.1)
    w = np.full(n // 2 + 1, 2.0)
    w[0] = w[-1] = 1.0

    def rms(f):
        return math.sqrt(float(np.sum(w * np.abs(f) ** 2)) / float(n) ** 6)

    with a, b:
        for _ in range(3):
            a.advance(2e-3)
            b.advance(2e-3)
        r = a.distance(b)
        sa, sb = a.state(), b.state()
        assert r["velocity_rms"] == pytest.approx(rms(sb.velocity_spectral), rel=1e-12)
        assert r["velocity_rms_difference"] == pytest.approx(rms(sa.velocity_spectral - sb.velocity_spectral), rel=1e-9)
        assert r["scalar_rms"] == pytest.approx(rms(sb.scalar_spectral), rel=1e-12)
        assert r["scalar_rms_difference"] == pytest.approx(rms(sa.scalar_spectral - sb.scalar_spectral), rel=1e-9)
        assert 0 < r["velocity_relative"] < 1e-5
        assert b.distance(b)["velocity_rms_difference"] == 0.0


def _spectrum_error(a, b, which):
    _, ea = a.energy_spectrum(which)
    _, eb = b.energy_spectrum(which)
    live = eb > 1e-12 * eb.max()
    return float(np.max(np.abs(ea[live] - eb[live]) / eb[live]))


@pytest.mark.timeout(900)
def test_fp32_bulk_against_fp64_at_1024_over_10_steps(g):
    """1024^3 forced isotropic turbulence with a mean-gradient scalar (zero initial fluctuation),
    IF-RK4, 2^24 Heun tracers from seed 2028: the FP32-bulk run against an FP64 run of the same flow
    on the same GPU, per step and per field (state distance), per reduction and per spectrum, and
    for the tracers' positions and samples.  Bars after 10 steps: velocity and scalar within 1e-5
    relative; positions within 2e-5 (a few FP32 roundings of 2 pi) and samples within 1e-5."""
    n, dt, steps, m = 1024, 5e-4, 10, 1 << 24
    t0 = time.perf_counter()
    u = g.random_phase_velocity(n, seed=2026)
    p = np.zeros((n, n, n // 2 + 1), dtype=np.complex128)
    pos = g.uniform_positions(m, L, 2028, np.float32).astype(np.float64)   # one set for both runs
    print(f"n=1024: seeded in {time.perf_counter() - t0:.1f} s", flush=True)
    runs = []
    for dtype in (np.float32, np.float64):
        dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=dtype)
        cond = g.FlowConditions(viscosity=g.predicted_viscosity(dom, 0.1, kmax_eta=1.5), schmidt=1.0,
                                velocity_forcing=g.ConstantPower(dom, injection_rate=0.1, lowest_shell=1,
                                                                 highest_shell=2))
        runs.append(g.Simulation(dom, conditions=cond, state=g.ResolvedFields(u, p), mean_gradient=1.0,
                                 tracers=g.Tracers(dom, positions=pos)))
        print(f"n=1024 {np.dtype(dtype).name}: context {runs[-1].memory_bytes() / 2 ** 30:.1f} GiB, "
              f"adopted at {time.perf_counter() - t0:.1f} s", flush=True)
    del u, p
    a, b = runs
    a.schedule((1, 0, 0, 0, 1))   # the FP32 run on the y passes' second radix plan (16.8.8)
    a._call("hdns_deposit")
    with a, b:
        keys = ("kinetic_energy", "dissipation", "scalar_variance", "scalar_dissipation")
        for s in range(1, steps + 1):
            a.advance(dt)
            b.advance(dt)
            r = a.distance(b)
            ra, rb = a.reductions(), b.reductions()
            red = " ".join(f"{k} {abs(ra[k] - rb[k]) / abs(rb[k]):.1e}" for k in keys)
            print(f"step {s:2d}: u {r['velocity_relative']:.2e} phi {r['scalar_relative']:.2e} | {red} | "
                  f"E(k) {_spectrum_error(a, b, 0):.1e} E_phi(k) {_spectrum_error(a, b, 1):.1e}", flush=True)
        assert r["velocity_relative"] <= 1e-5
        assert r["scalar_relative"] <= 1e-5
        dx = np.abs((a.tracer_positions() - b.tracer_positions() + L / 2) % L - L / 2).max()
        (va, sa), (vb, sb) = a.tracer_samples(), b.tracer_samples()
        print(f"2^24 tracers after {steps} steps: |dX| {dx:.2e}, v {rel(va, vb):.2e}, phi {rel(sa, sb):.2e}", flush=True)
        assert dx < 2e-5 and rel(va, vb) < 1e-5 and rel(sa, sb) < 1e-5


@pytest.mark.parametrize("name", ["dns16_rk4_forced", "dns16_rk3_unforced"])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-11), (np.float32, 1e-5)])
def test_device_reproduces_the_golden_fixtures(g, name, dtype, tol):
    """The committed oracle fixtures (tests/golden): states after 1 and 10 steps, reductions, max|u|,
    and for the forced case the tracers after 3 Heun steps."""
    from test_golden import GOLDEN, load_case

    d, cond, scheme, dt, z, unpack = load_case(name)
    n = d.n
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=dtype)
    vf = g.ConstantPower(dom, injection_rate=cond.forcing[0], lowest_shell=cond.forcing[1],
                         highest_shell=cond.forcing[2]) if cond.forcing else g.NoForcing()
    fc = g.FlowConditions(viscosity=cond.viscosity, schmidt=cond.schmidt, velocity_forcing=vf)
    t = np.load(GOLDEN / "tracers16_rk4_forced.npz") if name == "dns16_rk4_forced" else None
    tracers = g.Tracers(dom, positions=t["X0"].T.copy()) if t is not None else None
    sch = g.IFLowStorageRK4 if scheme == "rk4" else g.IFLowStorageRK3
    with g.Simulation(dom, conditions=fc, state=g.ResolvedFields(unpack("u0"), unpack("p0")),
                      mean_gradient=cond.mean_gradient, field_integration_scheme=sch, tracers=tracers) as sim:
        if t is not None:
            v, s = sim.tracer_samples()
            assert rel(v, t["v0"].T) < tol and rel(s, t["s0"]) < tol
        for step in range(1, 11):
            sim.advance(dt)
            if t is not None and step == 3:
                assert np.max(np.abs(sim.tracer_positions() - t["X3"].T)) < 10 * tol
                v, s = sim.tracer_samples()
                assert rel(v, t["v3"].T) < 10 * tol and rel(s, t["s3"]) < 10 * tol
            if step in (1, 10):
                st = sim.state()
                eu, ep = rel(st.velocity_spectral, unpack(f"u{step}")), rel(st.scalar_spectral, unpack(f"p{step}"))
                print(f"{name} {np.dtype(dtype).name} step {step}: u {eu:.2e} phi {ep:.2e}", flush=True)
                assert eu < tol and ep < tol
                r = sim.reductions()
                got = [r["kinetic_energy"], r["dissipation"], r["scalar_variance"], r["scalar_dissipation"]]
                np.testing.assert_allclose(got, z[f"reductions{step}"], rtol=10 * tol)
                assert sim.max_speed() == pytest.approx(float(z[f"max_speed{step}"]), rel=10 * tol)


@pytest.mark.parametrize("highest", [2, math.inf])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-11), (np.float32, 1e-5)])
def test_linear_forcing_matches_the_oracle(g, highest, dtype, tol):
    """Linear forcing (rate folded into the band's integrating factor), 10 IF-RK4 steps with the
    scalar, against the oracle; checked after 1 and 10 steps."""
    n, A, dt = 32, 0.4, 2e-3
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=dtype)
    cond = g.FlowConditions(viscosity=0.02, schmidt=1.0,
                            velocity_forcing=g.Linear(dom, rate=A, lowest_shell=1, highest_shell=highest))
    u, p = g.random_phase_velocity(n, seed=21), g.random_phase_scalar(n, seed=22)
    oc = dns.Conditions(viscosity=0.02, schmidt=1.0, mean_gradient=1.0, linear=(float(dtype(A)), 1, highest))
    d = dns.Domain(n)
    uo, po, dep = u, p, None
    with g.Simulation(dom, conditions=cond, state=g.ResolvedFields(u, p), mean_gradient=1.0) as sim:
        for s in range(1, 11):
            sim.advance(dt)
            uo, po, dep = dns.step(d, oc, "rk4", uo, po, dt, dep)
            if s in (1, 10):
                st = sim.state()
                assert rel(st.velocity_spectral, uo) < tol
                assert rel(st.scalar_spectral, po) < tol


def test_linear_forcing_grows_a_shear_wave_exactly_on_the_device(g):
    """u_x = 0.8 sin y (shell 1, forced at A) + 0.5 sin 3z (shell 3): no nonlinear term, so the two
    waves evolve as exp((A - nu) t) and exp(-9 nu t) exactly."""
    n, nu, A, dt, steps = 32, 0.05, 0.3, 1e-2, 25
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=np.float64)
    x = np.arange(n) * L / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    u = np.zeros((3, n, n, n))
    u[0] = 0.8 * np.sin(Y) + 0.5 * np.sin(3 * Z)
    uh = g.truncate_to_band(dns.rfft3(u), dom)
    cond = g.FlowConditions(viscosity=nu, schmidt=1.0, velocity_forcing=g.Linear(dom, rate=A))
    d = dns.Domain(n)
    band = d.shell_band(1, 2)
    expected = uh * np.where(band, np.exp((A - nu * d.k2) * steps * dt), np.exp(-nu * d.k2 * steps * dt))
    for scheme in ("rk4", "ssprk3", "heun"):
        with g.Simulation(dom, conditions=cond, state=g.ResolvedFields(uh, None),
                          field_integration_scheme=_scheme(g, scheme)) as sim:
            for _ in range(steps):
                sim.advance(dt)
            got = sim.state().velocity_spectral
        assert np.max(np.abs(got - expected)) <= 1e-12 * np.max(np.abs(expected)), scheme


@pytest.mark.parametrize("scheme", ["ssprk3", "heun"])
@pytest.mark.parametrize("forcing", ["constant_power", "linear"])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-11), (np.float32, 1e-5)])
def test_shu_osher_schemes_ten_steps_match_the_oracle(g, scheme, forcing, dtype, tol):
    """IFSSPRK3 (src/Propagator.jl:599-691) and IFHeun (:456-544) with the scalar and the mean
    gradient, 10 steps against the oracle's own arrangement of each; checked after 1 and 10 steps."""
    n, dt, nu = 32, 2e-3, 0.02
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=g.TwoThirds(), element_type=dtype)
    if forcing == "linear":
        vf = g.Linear(dom, rate=0.4, lowest_shell=1, highest_shell=2)
        okw = {"linear": (float(dtype(0.4)), 1, 2)}
    else:
        vf = g.ConstantPower(dom, injection_rate=0.1, lowest_shell=1, highest_shell=2)
        okw = {"forcing": (float(dtype(0.1)), 1, 2), "band_eps": float(np.finfo(dtype).eps)}
    cond = g.FlowConditions(viscosity=nu, schmidt=1.0, velocity_forcing=vf)
    u, p = g.random_phase_velocity(n, seed=31), g.random_phase_scalar(n, seed=32)
    oc = dns.Conditions(viscosity=nu, schmidt=1.0, mean_gradient=1.0, **okw)
    d = dns.Domain(n)
    uo, po, dep = u, p, None
    with g.Simulation(dom, conditions=cond, state=g.ResolvedFields(u, p), mean_gradient=1.0,
                      field_integration_scheme=_scheme(g, scheme)) as sim:
        for s in range(1, 11):
            sim.advance(dt)
            uo, po, dep = dns.step(d, oc, scheme, uo, po, dt, dep)
            if s in (1, 10):
                st = sim.state()
                eu, ep = rel(st.velocity_spectral, uo), rel(st.scalar_spectral, po)
                print(f"{scheme} {forcing} {np.dtype(dtype).name} step {s}: u {eu:.2e} phi {ep:.2e}", flush=True)
                assert eu < tol and ep < tol
                assert sim.max_speed() == pytest.approx(dep.max_speed, rel=10 * tol)


@pytest.mark.parametrize("scalar", [True, False])
def test_ssprk3_fp32_bulk_at_512_agrees_with_the_fp64_step(g, scalar):
    """The 512^3 kernels (the LDS-staged z forward variant, the split scalar kernel) on the
    IFSSPRK3 stages: the FP32-bulk state within 3e-6 of the FP64 state after 3 steps, by the
    device distance."""
    n, dt = 512, 1e-3
    sims = []
    for dtype in (np.float32, np.float64):
        sim, *_ = make_sim(g, n, dtype, scheme=g.IFSSPRK3, forcing=0.1, nu=2e-3, scalar=scalar)
        sims.append(sim)
    with sims[0], sims[1]:
        for _ in range(3):
            sims[0].advance(dt)
            sims[1].advance(dt)
        r = sims[0].distance(sims[1])
        print(f"ssprk3 512 scalar={scalar}:", {k: f"{v:.2e}" for k, v in r.items() if k.endswith("relative")},
              flush=True)
        assert r["velocity_relative"] < 3e-6
        if scalar:
            assert r["scalar_relative"] < 3e-6


def test_the_adjoint_refuses_ssprk3(g):
    sim, *_ = make_sim(g, 16, np.float64, scheme=g.IFSSPRK3)
    with sim:
        with pytest.raises(Exception, match="low-storage"):
            sim.adjoint_gradient(1e-3, 2)


def test_an_autotuned_ensemble_equals_standalone_runs_on_its_schedules(g):
    """Each batch of equal-size boxes tunes its own launch schedules, every box of the batch in one
    launch: the tuning leaves every state bitwise unchanged, the batches keep what they chose, and
    each box then equals a standalone run on its batch's schedule bit for bit (tracers included)."""
    sizes = [32, 48, 32, 64]
    nus = [0.005, 0.012, 0.02, 0.008]
    m, dt, steps = 64, 2e-3, 3
    members = [_member(g, n, np.float32, nu, 500 + 7 * b, m) for b, (n, nu) in enumerate(zip(sizes, nus))]
    with g.Ensemble(members, field_integration_scheme=g.IFLowStorageRK3, mean_gradient=1.0) as ens:
        before = [ens.box(b).state() for b in range(len(sizes))]
        report = ens.autotune(reps=1)
        assert sorted(report) == [32, 48, 64]
        chosen = ens.schedules()
        for n in chosen:
            assert chosen[n] == tuple(report[n][f]["chosen"] for f in g.SCHEDULE_FAMILIES), n
        for b in range(len(sizes)):
            after = ens.box(b).state()
            assert np.array_equal(before[b].velocity_spectral, after.velocity_spectral), b
            assert np.array_equal(before[b].scalar_spectral, after.scalar_spectral), b
        for s in range(steps):
            ens.step(dt, graph=s != 1)
            ens.advance_tracers(dt)
        print("schedules", chosen, flush=True)
        for b, mb in enumerate(members):
            with g.Simulation(mb.domain, conditions=mb.conditions, state=mb.state, tracers=mb.tracers,
                              field_integration_scheme=g.IFLowStorageRK3, mean_gradient=1.0) as alone:
                alone.schedule(chosen[sizes[b]])
                alone._call("hdns_deposit")
                for _ in range(steps):
                    alone.advance(dt, check=False)
                box = ens.box(b)
                a, e = alone.state(), box.state()
                assert np.array_equal(a.velocity_spectral, e.velocity_spectral), b
                assert np.array_equal(a.scalar_spectral, e.scalar_spectral), b
                assert alone.reductions() == box.reductions(), b
                assert np.array_equal(alone.tracer_positions(), box.tracer_positions()), b
        # a schedule set by hand holds; a size the ensemble lacks is refused
        want = {32: (0, 1, 1, 0, 0)}
        assert ens.schedules(want)[32] == want[32]
        with pytest.raises(ValueError, match="no box of size 128"):
            ens.schedules({128: (0, 0, 0, 0, 0)})
