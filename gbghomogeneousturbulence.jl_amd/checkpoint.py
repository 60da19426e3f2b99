"""Checkpoint and restart: write_simulation / read_simulation with the reference's schema
(ext/GBGHomogeneousTurbulenceHDF5Ext.jl:119-545, schema version 3).

The file is HDF5, as the reference's is: the same tree of groups, each tagged by its ``type`` and
``version`` attributes, with the parameters as attributes and the arrays as datasets, written by
the package's own HDF5 writer (``hdf5.py``: superblock 0, symbol-table groups, contiguous
datasets, the layout libhdf5 writes by default), so h5py, HDF5.jl and h5dump open it.  The
reference writes its leaves (spectral fields, point sets, the interpolation scheme) through its
field library's persistence, whose on-disk layout is not part of the reference; here a spectral
field is one complex dataset in the device's reference layout, a point set one (M, 3) array and
a trail sample one (3, M) array.  Loading runs nothing from the file.  A ``.npz`` file written
by an earlier version of this module (the same tree, attributes as JSON) is still read.

What a restored run continues from: the spectral state in the reference layout (an FP32-bulk run's
values survive the FP64 round trip exactly), the clock, the domain with its dealiasing, the
forcing and scheme settings, the tracer positions and the earlier samples of the velocity trail
with their times. The newest sample is measured again at the stored positions, as the reference
does.  Two extensions ride along as attributes: the mean scalar gradient and the launch schedule.
The schedule fixes the rounding of the device kernels, so with it the restored run continues bit
for bit.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from . import api, hdf5

SCHEMA_VERSION = 3
_TAG = "GBGHomogeneousTurbulence."


class Group:
    """One group of a Store: ``attrs`` (a dict of JSON values), datasets by name, subgroups."""

    def __init__(self, store: "Store", path: str):
        self._store = store
        self.path = path
        self.attrs = store.attrs.setdefault(path, {})

    def create_group(self, name: str) -> "Group":
        path = f"{self.path}/{name}" if self.path else name
        if path in self._store.attrs:
            raise ValueError(f"a group named {path!r} exists already")
        return Group(self._store, path)

    def __contains__(self, name: str) -> bool:
        path = f"{self.path}/{name}" if self.path else name
        return path in self._store.attrs or path in self._store.data

    def group(self, name: str, who: str) -> "Group":
        path = f"{self.path}/{name}" if self.path else name
        if path not in self._store.attrs:
            if path in self._store.data:
                raise ValueError(f"{who}: {name!r} is a dataset, not a group")
            raise ValueError(f"{who}: no object named {name!r} under the given parent")
        return Group(self._store, path)

    def __setitem__(self, name: str, value) -> None:
        self._store.data[f"{self.path}/{name}"] = np.ascontiguousarray(value)

    def __getitem__(self, name: str) -> np.ndarray:
        key = f"{self.path}/{name}"
        if key not in self._store.data:
            raise ValueError(f"no dataset {name!r} in group {self.path!r}")
        return self._store.data[key]


class Store:
    """A tree of groups, attributes and datasets, saved to and loaded from one HDF5 file."""

    def __init__(self):
        self.attrs: dict = {}
        self.data: dict = {}

    @property
    def root(self) -> Group:
        return Group(self, "")

    def save(self, path) -> None:
        """Write the tree as an HDF5 file (through a temporary file, then renamed)."""
        groups = {"": hdf5.Group()}

        def group(where: str) -> hdf5.Group:
            if where not in groups:
                parent, _, name = where.rpartition("/")
                groups[where] = hdf5.Group()
                group(parent).members[name] = groups[where]
            return groups[where]

        for where, attrs in self.attrs.items():
            group(where).attrs.update(attrs)
        for key, value in self.data.items():
            parent, _, name = key.rpartition("/")
            group(parent).members[name] = hdf5.Dataset(np.asarray(value))
        hdf5.write_file(path, groups[""])

    @classmethod
    def load(cls, path) -> "Store":
        s = cls()
        if hdf5.is_hdf5(path):
            def walk(g: hdf5.Group, where: str) -> None:
                s.attrs[where] = {k: (v.tolist() if isinstance(v, np.ndarray) else v) for k, v in g.attrs.items()}
                for name, m in g.members.items():
                    if isinstance(m, hdf5.Group):
                        walk(m, f"{where}/{name}" if where else name)
                    else:
                        s.data[f"{where}/{name}"] = m.data
            walk(hdf5.read_file(path), "")
            return s
        with np.load(path, allow_pickle=False) as z:
            for key in z.files:
                if key == "__attributes__":
                    s.attrs.update(json.loads(bytes(z[key]).decode("utf-8")))
                else:
                    s.data[key] = z[key]
        s.attrs.setdefault("", {})
        return s


def _require_tag(g: Group, who: str, tag: str) -> dict:
    a = g.attrs
    if "type" not in a:
        raise ValueError(f"{who}: group {g.path!r} has no `type` attribute; it was not written by this schema")
    if a["type"] != _TAG + tag:
        raise ValueError(f"{who}: group {g.path!r} holds a {a['type']}, not a {_TAG + tag}")
    if int(a.get("version", 0)) > SCHEMA_VERSION:
        raise ValueError(f"{who}: group {g.path!r} was written by schema version {a['version']}, newer than "
                         f"this reader's {SCHEMA_VERSION}")
    return a


def _tagged(parent: Group, name: str, tag: str) -> Group:
    g = parent.create_group(name)
    g.attrs["type"] = _TAG + tag
    g.attrs["version"] = SCHEMA_VERSION
    return g


def _shell(h):
    return math.inf if h == math.inf else int(h)


# --- the parts ---------------------------------------------------------------------------------

def write_domain(parent: Group, name: str, domain) -> Group:
    """ext/GBGHomogeneousTurbulenceHDF5Ext.jl:119-137: the grid, the box, the element type and the
    dealiasing strategy with its parameters (HouLi's order, its strength when set, and its
    materialized flag only when true)."""
    g = _tagged(parent, name, "FixedDomain")
    d = domain.dealiasing
    g.attrs.update(grid_points_per_axis=domain.grid_points_per_axis, side_length=domain.side_length,
                   eltype="Float32" if domain.element_type is np.float32 else "Float64",
                   dealiasing=type(d).__name__)
    if isinstance(d, api.HouLi):
        g.attrs["dealiasing_order"] = d.order
        if d.strength is not None:
            g.attrs["dealiasing_strength"] = d.strength
        if d.materialized:
            g.attrs["dealiasing_materialized"] = True
    return g


def read_domain(parent: Group, name: str):
    """ext/GBGHomogeneousTurbulenceHDF5Ext.jl:139-154."""
    g = parent.group(name, "read_domain")
    a = _require_tag(g, "read_domain", "FixedDomain")
    tag = a["dealiasing"]
    if tag == "TwoThirds":
        dealiasing = api.TwoThirds()
    elif tag == "PhaseShift":
        dealiasing = api.PhaseShift()
    elif tag == "HouLi":
        dealiasing = api.HouLi(order=int(a["dealiasing_order"]), strength=a.get("dealiasing_strength"),
                               materialized=bool(a.get("dealiasing_materialized", False)))
    else:
        raise ValueError(f"read_domain: unrecognized dealiasing tag {tag!r} in group {name!r}")
    et = {"Float32": np.float32, "Float64": np.float64}.get(a["eltype"])
    if et is None:
        raise ValueError(f"read_domain: unrecognized element type {a['eltype']!r}")
    return api.FixedDomain(grid_points_per_axis=int(a["grid_points_per_axis"]), side_length=float(a["side_length"]),
                           dealiasing=dealiasing, element_type=et)


def write_forcing(parent: Group, name: str, f) -> Group:
    """ext/GBGHomogeneousTurbulenceHDF5Ext.jl:305-363."""
    if f is None or isinstance(f, api.NoForcing):
        return _tagged(parent, name, "NoForcing")
    if isinstance(f, api.Linear):
        g = _tagged(parent, name, "Linear")
        g.attrs["rate"] = f.rate
    elif isinstance(f, api.ConstantPower):
        g = _tagged(parent, name, "ConstantPower")
        g.attrs["injection_rate"] = f.injection_rate
    elif isinstance(f, api.VarianceRescaling):
        g = _tagged(parent, name, "VarianceRescaling")
        g.attrs.update(target=f.target, over=f.over)
    elif isinstance(f, api.VarianceController):
        g = _tagged(parent, name, "VarianceController")
        g.attrs.update(target=f.target, gain=f.gain, turnover_time=f.turnover_time,
                       turnover_time_derived=f.turnover_time_derived)
    elif isinstance(f, api.StochasticOU):
        g = _tagged(parent, name, "StochasticOU")
        g.attrs.update(forcing_variance=f.forcing_variance, correlation_time=f.correlation_time, seed=f.seed)
    else:
        raise ValueError(f"write_forcing: no schema for {f!r}")
    g.attrs.update(lowest_shell=f.lowest_shell, highest_shell=f.highest_shell)
    return g


def read_forcing(parent: Group, name: str, domain):
    g = parent.group(name, "read_forcing")
    a = g.attrs
    if "type" not in a:
        raise ValueError(f"read_forcing: group {name!r} has no `type` attribute; it was not written by write_forcing")
    tag = a["type"].removeprefix(_TAG)
    lo, hi = int(a.get("lowest_shell", 1)), _shell(a.get("highest_shell", 2))
    if tag == "NoForcing":
        return api.NoForcing()
    if tag == "Linear":
        return api.Linear(domain, rate=a["rate"], lowest_shell=lo, highest_shell=hi)
    if tag == "ConstantPower":
        return api.ConstantPower(domain, injection_rate=a["injection_rate"], lowest_shell=lo, highest_shell=hi)
    if tag == "VarianceRescaling":
        return api.VarianceRescaling(domain, target=a["target"], over=a["over"], lowest_shell=lo, highest_shell=hi)
    if tag == "StochasticOU":
        return api.StochasticOU(domain, forcing_variance=a["forcing_variance"], correlation_time=a["correlation_time"],
                                lowest_shell=lo, highest_shell=hi, seed=int(a.get("seed", 0)))
    if tag == "VarianceController":
        derived = bool(a.get("turnover_time_derived", False))
        return api.VarianceController(domain, target=a["target"], gain=a["gain"],
                                      turnover_time=None if derived else a["turnover_time"],
                                      lowest_shell=lo, highest_shell=hi)
    raise ValueError(f"read_forcing: unrecognized forcing tag {a['type']!r} in group {name!r}")


_FIELD_SCHEMES = {"IFLowStorageRK3": api.IFLowStorageRK3, "IFLowStorageRK4": api.IFLowStorageRK4,
                  "IFSSPRK3": api.IFSSPRK3, "IFHeun": api.IFHeun}


def write_simulation(path, sim, name: str = "simulation") -> None:
    """write_simulation (ext/GBGHomogeneousTurbulenceHDF5Ext.jl:426-456) into a new file at `path`."""
    store = Store()
    g = _tagged(store.root, name, "Simulation")
    c = sim.conditions
    g.attrs.update(time=sim.time, step=sim.step, viscosity=c.viscosity, schmidt=c.schmidt, cfl=sim.cfl,
                   max_dt=sim.max_dt, scalar=sim.carries_scalar,
                   field_integration_scheme=sim.field_integration_scheme.__name__, array_family="MI355X",
                   mean_gradient=sim.mean_gradient, schedule=[int(x) for x in sim.schedule()])
    write_domain(g, "domain", sim.domain)
    st = _tagged(g, "state", "ResolvedFields")
    st.attrs["scalar"] = sim.carries_scalar
    fields = sim.state()
    st["velocity_spectral"] = fields.velocity_spectral
    if sim.carries_scalar:
        st["scalar_spectral"] = fields.scalar_spectral
    # the forcing states (src/Forcing.jl StochasticOUState; ext/GBGHomogeneousTurbulenceHDF5Ext.jl:153-168):
    # the process and its draw count, with which the stored seed fixes every later draw
    for field, key in ((0, "modeled_velocity_forcing"), (1, "modeled_scalar_forcing")):
        proc = sim.ou.get(field)
        if proc is not None:
            og = _tagged(st, key, "StochasticOUState")
            og.attrs["draws"] = proc.draws
            og["process"] = proc.process
    write_forcing(g, "velocity_forcing", c.velocity_forcing)
    write_forcing(g, "scalar_forcing", c.scalar_forcing)
    if sim.tracers is not None:
        ps = sim.particle_integration_scheme
        g.attrs.update(particle_integration_scheme="AB" if isinstance(ps, api.AB) else "ABM",
                       particle_scheme_steps=ps.steps, shortened_first_steps=sim.shortened_first_steps)
        tr = _tagged(g, "tracers", "Tracers")
        tr["positions"] = sim.tracer_positions()
        if sim.tracers.unwrapped:   # the unwrapped record, when tracked (ext/GBGHomogeneousTurbulenceHDF5Ext.jl:229-230)
            tr["unwrapped_positions"] = sim.unwrapped_positions()
        tr.attrs["interpolation_scheme"] = "CatmullRom"
        # the newest sample is measured again at the stored positions on read; the earlier ones
        # are stored, oldest first, with their times (src/Particles.jl trail order)
        times = sim.trail_times()
        tg = _tagged(g, "velocity_at_particles_trail", "ParticleTrail")
        earlier = list(range(len(times) - 1, 0, -1))          # offsets back from the newest, oldest first
        tg.attrs.update(capacity=ps.trail_capacity, tensor_order=1, times=[times[j] for j in earlier])
        for i, j in enumerate(earlier, 1):
            tg[f"sample{i}"] = sim.trail_sample(-j)
    store.save(path)


def stored_trail_samples(trail, m: int) -> np.ndarray:
    """The earlier velocity samples of a stored trail group, newest first, [n, 3, m].  Each stored
    sample must hold one velocity per tracer; one that does not is refused by its number
    (ext/GBGHomogeneousTurbulenceHDF5Ext.jl:268-283)."""
    n = len(trail.attrs["times"])
    for i in range(1, n + 1):
        got = np.shape(trail[f"sample{i}"])
        if got != (3, m):
            raise ValueError(f"read_simulation: the stored sample {i} does not fit the trail it is read into: "
                             f"expected shape {(3, m)} (components, tracers), got {got}")
    if n == 0:
        return np.zeros((0, 3, m))
    return np.stack([np.asarray(trail[f"sample{i}"], dtype=np.float64) for i in range(n, 0, -1)])


def read_simulation(path, name: str = "simulation", **simulation_kwargs):
    """read_simulation (ext/GBGHomogeneousTurbulenceHDF5Ext.jl:458-545): a Simulation that continues
    the stored run; `simulation_kwargs` pass device options (keep_vorticity, tracer_capacity, ...)."""
    store = Store.load(path)
    g = store.root.group(name, "read_simulation")
    a = _require_tag(g, "read_simulation", "Simulation")
    domain = read_domain(g, "domain")
    vf = read_forcing(g, "velocity_forcing", domain)
    sf = read_forcing(g, "scalar_forcing", domain)
    scheme = _FIELD_SCHEMES.get(a["field_integration_scheme"])
    if scheme is None:
        raise ValueError(f"read_simulation: unrecognized field scheme tag {a['field_integration_scheme']!r}")
    sg = g.group("state", "read_simulation")
    _require_tag(sg, "read_simulation", "ResolvedFields")
    scalar = bool(a["scalar"])
    state = api.ResolvedFields(sg["velocity_spectral"], sg["scalar_spectral"] if scalar else None)
    tracers, ps, shortened, trail, unwrapped = None, None, False, None, None
    if "tracers" in g:
        if "velocity_at_particles_trail" not in g:
            raise ValueError(f"read_simulation: group {name!r} carries tracers without the times of its velocity "
                             "samples, so the run cannot be continued from it")
        tg = g.group("tracers", "read_simulation")
        _require_tag(tg, "read_simulation", "Tracers")
        tracers = api.Tracers(domain, positions=tg["positions"], unwrapped="unwrapped_positions" in tg)
        unwrapped = tg["unwrapped_positions"] if tracers.unwrapped else None
        k = int(a["particle_scheme_steps"])
        ps = {"AB": api.AB, "ABM": api.ABM}[a["particle_integration_scheme"]](k)
        if isinstance(ps, api.ABM) and k == 1:
            ps = api.Heun()
        shortened = bool(a["shortened_first_steps"])
        trail = g.group("velocity_at_particles_trail", "read_simulation")
        _require_tag(trail, "read_simulation", "ParticleTrail")
        trail_samples = stored_trail_samples(trail, tracers.count)
    mode = str(simulation_kwargs.get("traversal_mode", "automatic")).lstrip(":")
    if tracers is None and mode != "automatic":   # ext/GBGHomogeneousTurbulenceHDF5Ext.jl:507-511
        raise ValueError(f"read_simulation: `traversal_mode = :{mode}` was given, but group {name!r} stores no "
                         "tracers to order the gathers over")
    cond = api.FlowConditions(viscosity=a["viscosity"], schmidt=a["schmidt"], velocity_forcing=vf, scalar_forcing=sf)
    stored_ou = {}
    for field, key in ((0, "modeled_velocity_forcing"), (1, "modeled_scalar_forcing")):
        if key in sg:
            og = sg.group(key, "read_simulation")
            oa = _require_tag(og, "read_simulation", "StochasticOUState")
            stored_ou[field] = (og["process"], int(oa["draws"]))
    sim = api.Simulation(domain, conditions=cond, state=state, field_integration_scheme=scheme, tracers=tracers,
                         particle_integration_scheme=ps, mean_gradient=float(a.get("mean_gradient", 0.0)),
                         cfl=float(a["cfl"]), max_dt=float(a["max_dt"]), shortened_first_steps=shortened,
                         **simulation_kwargs)
    try:
        for field, (process, draws) in stored_ou.items():
            if field not in sim.ou or sim.ou[field].process.shape != process.shape:
                raise ValueError(f"read_simulation: group {name!r} stores a forcing state that does not fit its "
                                 "stored forcing")
            sim.ou[field] = api.OUProcess(sim.ou[field].forcing, sim.ou[field].order, process, draws)
        if unwrapped is not None:
            sim.set_unwrapped_positions(unwrapped)
        if "schedule" in a:
            sim.schedule(np.asarray(a["schedule"], dtype=np.int32))
            sim.redeposit()
        sim.time = float(a["time"])
        sim.step = int(a["step"])
        if trail is not None:
            times = list(trail.attrs["times"])                # oldest first
            n = len(times)
            sim.restore_trail(times[::-1], trail_samples, sim.time)
    except Exception:
        sim.close()
        raise
    return sim
