"""Checkpoint and restart on the device: a run written after 3 steps and read back continues bit
for bit like the run that never stopped (the reference's test/hdf5_checkpoint.jl:202-230 asks the
same of an IFSSPRK3 run), through the state, the clock, the forcing, the tracer positions and the
velocity trail with its times, and the unwrapped positions (test/hdf5_checkpoint.jl:114-135).
The file is HDF5; a HouLi or PhaseShift run comes back with its dealiasing."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
L = 2 * math.pi


def _run(g, dtype, scheme, forcing, particles, unwrapped=True, dealiasing=None):
    n = 32
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=L, dealiasing=dealiasing or g.TwoThirds(),
                        element_type=dtype)
    vf = {"constant_power": lambda: g.ConstantPower(dom, injection_rate=0.1),
          "stochastic_ou": lambda: g.StochasticOU(dom, forcing_variance=0.3, correlation_time=0.05, seed=17),
          "controller": lambda: g.VarianceController(dom, target=0.6, turnover_time=0.7),
          "rescaling": lambda: g.VarianceRescaling(dom, target=0.55, over="total")}[forcing]()
    cond = g.FlowConditions(viscosity=0.02, schmidt=0.9, velocity_forcing=vf)
    tr = g.Tracers(dom, number_of_particles=300, seed=5, unwrapped=unwrapped)
    return g.Simulation(dom, conditions=cond, state=g.ResolvedFields(g.random_phase_velocity(n, seed=71),
                                                                     g.random_phase_scalar(n, seed=72)),
                        field_integration_scheme=scheme, tracers=tr, particle_integration_scheme=particles,
                        mean_gradient=0.8)


def _snapshot(sim):
    st = sim.state()
    times = sim.trail_times()
    return {"u": st.velocity_spectral, "phi": st.scalar_spectral, "x": sim.tracer_positions(),
            "unwrapped": sim.unwrapped_positions(), "times": times,
            "trail": [sim.trail_sample(-j) for j in range(len(times))], "clock": (sim.time, sim.step)}


@pytest.mark.parametrize("dtype,scheme,forcing,particles", [
    (np.float32, "IFSSPRK3", "constant_power", ("AB", 3)),
    (np.float64, "IFLowStorageRK4", "controller", ("ABM", 2)),
    (np.float32, "IFLowStorageRK3", "rescaling", ("ABM", 1)),
    (np.float32, "IFLowStorageRK4", "stochastic_ou", ("AB", 2)),
])
def test_a_restored_run_continues_bit_for_bit(g, tmp_path, dtype, scheme, forcing, particles):
    from gbg_amd import checkpoint as ck
    sch = getattr(g, scheme)
    ps = (lambda: g.AB(particles[1])) if particles[0] == "AB" else (lambda: g.ABM(particles[1]))
    dts = [1.5e-3, 2e-3, 1.7e-3, 2.2e-3, 1.9e-3, 2.1e-3]
    with _run(g, dtype, sch, forcing, ps()) as a:
        for dt in dts:
            a.advance(dt)
        want = _snapshot(a)
    path = tmp_path / "run.h5"
    with _run(g, dtype, sch, forcing, ps()) as b:
        for dt in dts[:3]:
            b.advance(dt)
        ck.write_simulation(path, b)
    with ck.read_simulation(path) as c:
        assert c.step == 3 and c.time == (dts[0] + dts[1]) + dts[2]
        for dt in dts[3:]:
            c.advance(dt)
        got = _snapshot(c)
    print(f"{scheme} {forcing} {particles} {np.dtype(dtype).name}: trail {got['times']}", flush=True)
    assert got["clock"] == want["clock"]
    assert np.array_equal(got["u"], want["u"])
    assert np.array_equal(got["phi"], want["phi"])
    assert np.array_equal(got["x"], want["x"])
    assert np.array_equal(got["unwrapped"], want["unwrapped"])
    assert got["times"] == want["times"]
    for a_, b_ in zip(got["trail"], want["trail"]):
        assert np.array_equal(a_, b_)


def test_the_checkpoint_keeps_the_settings(g, tmp_path):
    from gbg_amd import checkpoint as ck
    with _run(g, np.float64, g.IFHeun, "controller", g.AB(4)) as a:
        a.advance(1e-3)
        a.schedule(np.array([0, 1, 1, 0, 1], dtype=np.int32))
        ck.write_simulation(tmp_path / "s.h5", a)
    with ck.read_simulation(tmp_path / "s.h5") as b:
        assert b.field_integration_scheme is g.IFHeun
        assert repr(b.particle_integration_scheme) == "AB(4)"
        f = b.conditions.velocity_forcing
        assert isinstance(f, g.VarianceController) and (f.target, f.turnover_time) == (0.6, 0.7)
        assert b.mean_gradient == 0.8 and b.conditions.schmidt == 0.9
        assert list(b.schedule()) == [0, 1, 1, 0, 1]
        assert b.trail_times() == [1e-3, 0.0]


def test_an_untracked_set_stays_untracked(g, tmp_path):
    """test/hdf5_checkpoint.jl:74-78: a set without the unwrapped record reads back without it."""
    from gbg_amd import checkpoint as ck
    with _run(g, np.float64, g.IFLowStorageRK4, "constant_power", g.Heun(), unwrapped=False) as a:
        a.advance(1e-3)
        with pytest.raises(ValueError, match="no unwrapped positions"):
            a.unwrapped_positions()
        ck.write_simulation(tmp_path / "p.h5", a)
    with ck.read_simulation(tmp_path / "p.h5") as b:
        assert not b.tracers.unwrapped
        with pytest.raises(ValueError, match="no unwrapped positions"):
            g.unwrapped_positions(b)


@pytest.mark.parametrize("dealiasing", ["houli", "phase_shift"])
def test_a_restored_run_keeps_its_dealiasing(g, tmp_path, dealiasing):
    """write_domain / read_domain (ext/GBGHomogeneousTurbulenceHDF5Ext.jl:119-154): the strategy and
    HouLi's order and strength come back, and the run continues bit for bit."""
    from gbg_amd import checkpoint as ck
    from gbg_amd import hdf5
    make = {"houli": lambda: g.HouLi(order=36, strength=63.463), "phase_shift": g.PhaseShift}[dealiasing]
    dts = [2e-3, 1.8e-3, 2.1e-3, 1.9e-3]
    with _run(g, np.float64, g.IFSSPRK3, "constant_power", g.AB(4), dealiasing=make()) as a:
        for dt in dts:
            a.advance(dt)
        want = _snapshot(a)
    with _run(g, np.float64, g.IFSSPRK3, "constant_power", g.AB(4), dealiasing=make()) as b:
        for dt in dts[:2]:
            b.advance(dt)
        ck.write_simulation(tmp_path / "d.h5", b)
    assert hdf5.is_hdf5(tmp_path / "d.h5")
    with ck.read_simulation(tmp_path / "d.h5") as c:
        assert repr(c.domain.dealiasing) == repr(make())
        for dt in dts[2:]:
            c.advance(dt)
        got = _snapshot(c)
    assert got["clock"] == want["clock"]
    assert np.array_equal(got["u"], want["u"]) and np.array_equal(got["phi"], want["phi"])
    assert np.array_equal(got["x"], want["x"])
