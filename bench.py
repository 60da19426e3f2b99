#!/usr/bin/env python3
"""Headline benchmark: Mpoint*steps/s of the 512^3 IF-RK4 DNS step, FP32-bulk beside plain FP64.

Workload (one GPU): 512^3 triple-periodic box of side 2 pi, 2/3 dealiasing, IF-RK4
(Carpenter-Kennedy, 5 stages), velocity + passive scalar (Sc = 1) with a uniform mean gradient
G = 1 along z, ConstantPower eps = 0.1 on shells 1-2, nu from k_max eta = 1.5 (8.419e-4), fixed
dt = 1e-3.  Synthetic input: random-phase solenoidal u with E(k) ~ k^4 exp(-2 (k/4)^2), k = 1/2,
seed 2026; phi = 0.  Tracers off.  A step is every stage plus the closing deposit, with the state
resident in HBM before the timed region starts.  The headline times the step as its captured HIP
graph (hdns_step_graph, bitwise the eager step) without per-kernel events; `timing` adds the eager
launches with and without those events, and the kernel table comes from the run with them.

Prints one JSON line (rank 0, one GPU).  Extra keys: fp64 (the plain-FP64 step on the same GPU),
kernels (per-kernel event times), roofline (the dominant kernel against 8 TB/s), cpu_baseline (the
FP64 oracle on the host cores, a bounded sample).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
KIND_NAMES = ["y_forward", "z_forward_stage_update", "reality_projection", "z_inverse_reductions",
              "reduction_finalize", "y_inverse", "x_pass_products", "other"]
KERNEL_SYMBOLS = ["k_yfwd", "k_zf", "k_reality", "k_zi", "k_finalize", "k_yinv", "k_xpass", ""]


def hip():
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            h = ctypes.CDLL(name)
            break
        except OSError:
            continue
    else:
        raise RuntimeError("libamdhip64.so not found")
    h.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    h.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    h.hipEventSynchronize.argtypes = [ctypes.c_void_p]
    h.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
    h.hipEventDestroy.argtypes = [ctypes.c_void_p]
    return h


def algorithmic_bytes(n: int, real_bytes: int, stages: int, scalar: bool = True, ssp: bool = False) -> dict:
    """Compulsory HBM bytes per launch of each kernel, counting retained modes only (kx <= M)."""
    m = n // 3
    k = 2 * m + 1
    kx = m + 1
    cb = 2 * real_bytes
    nf, ni, ns = (6, 7, 4) if scalar else (3, 6, 3)
    # accumulator passes per stage: the 2N schemes neither read it at the first stage nor write it at
    # the last; the Shu-Osher schemes write the initial state once and read it at every later stage
    acc = 1.0 if ssp else 2 * (stages - 1) / stages
    comp = k * k * kx
    return {
        "y_forward": nf * n * kx * (n + k) * cb,
        "z_forward_stage_update": (nf * n * k * kx + 2 * ns * comp + ns * comp * acc) * cb,
        "z_inverse_reductions": (ns * comp + ni * n * k * kx) * cb,
        "y_inverse": ni * n * kx * (k + n) * cb,
        "x_pass_products": (ni + nf) * n * n * kx * cb,
        "reality_projection": 2 * ns * k * k * cb,
        "reduction_finalize": 0,
    }


def time_gpu(g, n: int, dtype, steps: int, warmup: int, H, seed: int = 2026, scheme=None):
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=2 * math.pi, dealiasing=g.TwoThirds(), element_type=dtype)
    nu = g.predicted_viscosity(dom, 0.1, kmax_eta=1.5)
    u = g.random_phase_velocity(n, seed=seed)
    p = np.zeros((n, n, n // 2 + 1), dtype=np.complex128)
    cond = g.FlowConditions(viscosity=nu, schmidt=1.0,
                            velocity_forcing=g.ConstantPower(dom, injection_rate=0.1, lowest_shell=1, highest_shell=2))
    dt = 1e-3
    with g.Simulation(dom, conditions=cond, state=g.ResolvedFields(u, p), mean_gradient=1.0,
                      field_integration_scheme=scheme or g.IFLowStorageRK4, keep_physical=False) as sim:
        del u, p
        tuned = sim.autotune(reps=2)
        for _ in range(warmup):
            sim.step_fields(dt)
        sim.synchronize()
        stream = sim.stream()
        e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
        H.hipEventCreate(ctypes.byref(e0))
        H.hipEventCreate(ctypes.byref(e1))
        sim.profile(True)
        sim.profile_read(reset=True)
        H.hipDeviceSynchronize()
        t0 = time.perf_counter()
        H.hipEventRecord(e0, ctypes.c_void_p(stream))
        for _ in range(steps):
            sim.step_fields(dt)
        H.hipEventRecord(e1, ctypes.c_void_p(stream))
        H.hipEventSynchronize(e1)
        H.hipDeviceSynchronize()
        wall = time.perf_counter() - t0
        ms = ctypes.c_float()
        H.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
        kms, kcnt = sim.profile_read(reset=True)
        sim.profile(False)

        # the headline timing: the same steps without the per-kernel events (their records are
        # measurement overhead), launched kernel by kernel and as the captured graph of the step
        def timed(graph: bool) -> float:
            sim.step_fields(dt, graph=graph)          # (the graph is captured here, once)
            H.hipDeviceSynchronize()
            H.hipEventRecord(e0, ctypes.c_void_p(stream))
            for _ in range(steps):
                sim.step_fields(dt, graph=graph)
            H.hipEventRecord(e1, ctypes.c_void_p(stream))
            H.hipEventSynchronize(e1)
            out = ctypes.c_float()
            H.hipEventElapsedTime(ctypes.byref(out), e0, e1)
            return float(out.value)

        ms_plain = timed(False)
        ms_graph = timed(True)
        H.hipEventDestroy(e0)
        H.hipEventDestroy(e1)
        speed = sim.max_speed()
        red = sim.reductions()
        mem = sim.memory_bytes()
    if not (math.isfinite(speed) and math.isfinite(red["scalar_variance"])):
        raise RuntimeError("the benchmark run blew up")
    return {"ms_total": float(ms.value), "ms_plain": ms_plain, "ms_graph": ms_graph,
            "wall_s": wall, "kernel_ms": kms, "kernel_launches": kcnt,
            "max_speed": speed, "kinetic_energy": red["kinetic_energy"], "memory_bytes": mem, "nu": nu,
            "schedule": tuned}


def tracer_bytes(n: int, m: int, real_bytes: int) -> dict:
    """Compulsory HBM bytes of the tracer kernels of one Heun step, M tracers on an n^3 grid: each
    gather kernel reads its fields once (every grid point lies in some tracer's 4^3 stencil at
    these densities) plus each tracer's own records (positions, trail or samples), which the
    rebuild keeps in traversal order so the gathers visit them in slot order."""
    w = real_bytes
    return {"advect": 3 * n ** 3 * w + m * (3 * w + 3 * w + 3 * w),        # u; X, f, X written
            "sample": 4 * n ** 3 * w + m * (3 * w + 3 * w + w),            # u, phi; X, u and phi written
            # one pass: u and phi once; X, f read; X, the u and phi samples written
            "advect_sample": 4 * n ** 3 * w + m * (3 * w + 3 * w + 3 * w + 3 * w + w),
            # X and the slot map read, keys and indices written (sort aside); the records moved
            # into the new order (X, two trail slots, phi; read and written) with their slot maps
            "rebuild": m * (3 * w + 4 + 8) + m * (2 * 10 * w + 16)}


def time_tracers(g, n: int, m: int, dtype, steps: int, H) -> dict:
    """Tracer throughput: traversal rebuild + Heun push + closing sample per step, M tracers on an
    n^3 field (the field step itself is excluded from this figure), as the run loop makes them (the
    push and the sample in one launch, hdns_advect_deposit); each phase between its own HIP events
    on the context's stream, and the push-and-sample's rate against the HBM roofline.  A second run
    times the push and the sample as two launches (hdns_advect, hdns_deposit_tracers)."""
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=2 * math.pi, dealiasing=g.TwoThirds(), element_type=dtype)
    u = g.random_phase_velocity(n, seed=2026)
    p = np.zeros((n, n, n // 2 + 1), dtype=np.complex128)
    cond = g.FlowConditions(viscosity=g.predicted_viscosity(dom, 0.1, kmax_eta=1.5), schmidt=1.0)
    dt = 1e-3
    w = np.dtype(dtype).itemsize
    ab = tracer_bytes(n, m, w)

    def run(fused: bool, sched: int = -1) -> dict:
        phases = ("rebuild", "advect_sample") if fused else ("rebuild", "advect", "sample")
        total = dict.fromkeys(phases, 0.0)
        tr = g.Tracers(dom, number_of_particles=m, seed=2028)
        with g.Simulation(dom, conditions=cond, state=g.ResolvedFields(u, p), mean_gradient=1.0, tracers=tr) as sim:
            total["gather_launch"] = sim.tracer_schedule(sched)
            stream = ctypes.c_void_p(sim.stream())
            ev = [ctypes.c_void_p() for _ in range(len(phases) + 1)]
            for e in ev:
                H.hipEventCreate(ctypes.byref(e))
            t = 0.0
            for s in range(steps + 1):
                sim.step_fields(dt)
                t += dt
                H.hipEventRecord(ev[0], stream)
                sim._call("hdns_rebuild_traversal")
                H.hipEventRecord(ev[1], stream)
                if fused:
                    sim.advect_and_deposit(dt, t)
                else:
                    sim.advect(dt)
                    H.hipEventRecord(ev[2], stream)
                    sim.deposit_tracers(t)
                H.hipEventRecord(ev[-1], stream)
                H.hipEventSynchronize(ev[-1])
                if s > 0:
                    for i, ph in enumerate(phases):
                        ms = ctypes.c_float()
                        H.hipEventElapsedTime(ctypes.byref(ms), ev[i], ev[i + 1])
                        total[ph] += ms.value
            for e in ev:
                H.hipEventDestroy(e)
        out = {ph: total[ph] / steps for ph in phases}
        out["gather_launch"] = total["gather_launch"]
        return out

    per = run(True)
    launch = per.pop("gather_launch")
    other = run(True, 1 - launch)
    other.pop("gather_launch")
    sep = run(False)
    sep.pop("gather_launch")
    ms_step = sum(per.values())
    gms = per["advect_sample"]
    gb = ab["advect_sample"]
    traffic = None
    tfile = ROOT / "profiles" / "traffic.json"
    if tfile.exists():
        try:
            t = json.loads(tfile.read_text()).get(f"k_advect_sample<{'float' if w == 4 else 'double'}>@{n}x{m}")
            traffic = t["bytes_per_launch"] if t else None
        except (ValueError, KeyError):
            traffic = None
    return {"grid": n, "tracers": m, "dtype": np.dtype(dtype).name, "ms_per_step": round(ms_step, 4),
            "ms": {ph: round(v, 4) for ph, v in per.items()},
            "tracer_steps_per_s": round(m / (ms_step * 1e-3), 1),
            "two_launches_ms": {ph: round(v, 4) for ph, v in sep.items()},
            "gather_launch": {"in_force": "capped" if launch else "uncapped",
                              "other_launch_ms": round(other["advect_sample"], 4)},
            "roofline": {"bound": "hbm", "kernel": "k_advect_sample", "algorithmic_bytes": gb,
                         "achieved": round(gb / (gms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gb / (gms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "traffic": traffic}}


def time_ensemble(g, steps: int, H, boxes_per_size: int = 32, tracers: int = 4096) -> dict:
    """The batched ensemble: 32 boxes each of 32^3, 48^3 and 64^3, nu log-uniform in [0.005, 0.05],
    own seeded random-phase fields and scalars, IF-RK3 FP32-bulk, 4096 tracers per box; each
    batch's launch schedules autotuned, then one graph launch per step and the tracer half of the
    iteration."""
    sizes = [32] * boxes_per_size + [48] * boxes_per_size + [64] * boxes_per_size
    rng = np.random.default_rng(2029)
    nus = np.exp(rng.uniform(np.log(0.005), np.log(0.05), len(sizes)))
    members = []
    for b, (n, nu) in enumerate(zip(sizes, nus)):
        dom = g.FixedDomain(grid_points_per_axis=n, side_length=2 * math.pi, dealiasing=g.TwoThirds(),
                            element_type=np.float32)
        state = g.ResolvedFields(g.random_phase_velocity(n, seed=3000 + b), g.random_phase_scalar(n, seed=5000 + b))
        members.append(g.EnsembleMember(dom, g.FlowConditions(viscosity=float(nu), schmidt=1.0), state,
                                        g.Tracers(dom, number_of_particles=tracers, seed=7000 + b)))
    dt = 1e-3
    with g.Ensemble(members, field_integration_scheme=g.IFLowStorageRK3, mean_gradient=1.0) as ens:
        stream = ctypes.c_void_p()
        h = ctypes.c_void_p()
        g._lib.check(g._lib.lib().hdns_ensemble_member(ens._h, 0, ctypes.byref(h)))
        g._lib.check(g._lib.lib().hdns_stream(h, ctypes.byref(stream)))   # the ensemble's shared stream
        tuned = ens.autotune(reps=2)   # each batch of equal-size boxes, as the headline run is tuned
        for _ in range(3):
            ens.step(dt)
            ens.advance_tracers(dt)
        e0, e1, e2 = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        for e in (e0, e1, e2):
            H.hipEventCreate(ctypes.byref(e))
        field_ms = 0.0
        H.hipEventRecord(e0, stream)
        for _ in range(steps):
            H.hipEventRecord(e1, stream)
            ens.step(dt)
            H.hipEventRecord(e2, stream)
            ens.advance_tracers(dt)
            H.hipEventSynchronize(e2)
            ms = ctypes.c_float()
            H.hipEventElapsedTime(ctypes.byref(ms), e1, e2)
            field_ms += ms.value
        H.hipEventRecord(e1, stream)
        H.hipEventSynchronize(e1)
        total = ctypes.c_float()
        H.hipEventElapsedTime(ctypes.byref(total), e0, e1)
        for e in (e0, e1, e2):
            H.hipEventDestroy(e)
    points = sum(n ** 3 for n in sizes)
    per = total.value / steps
    fms = field_ms / steps
    # the field step's compulsory bytes: every box's stage kernels under the same model as the
    # headline's kernel table, three stages of IF-RK3
    step_bytes = 0.0
    for n in set(sizes):
        ab = algorithmic_bytes(n, 4, 3)
        per_stage = sum(v for k, v in ab.items() if k != "reality_projection")
        step_bytes += sizes.count(n) * (3 * per_stage + ab["reality_projection"])
    return {"boxes": len(sizes), "sizes": {"32": boxes_per_size, "48": boxes_per_size, "64": boxes_per_size},
            "scheme": "IFLowStorageRK3", "dtype": "f32", "graph": True, "steps": steps,
            "autotuned_schedule": {str(n): [t[f]["chosen"] for f in g.SCHEDULE_FAMILIES] for n, t in tuned.items()},
            "ms_per_iteration": round(per, 4), "field_ms_per_step": round(fms, 4),
            "gridpoint_steps_per_s": round(points / (fms * 1e-3), 1),
            "tracer_steps_per_s": round(len(sizes) * tracers / (per * 1e-3), 1),
            "tracers_per_box": tracers,
            "roofline": {"bound": "hbm", "what": "the field step (every box, every stage kernel)",
                         "algorithmic_bytes": int(step_bytes), "achieved": round(step_bytes / (fms * 1e-3) / 1e9, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(step_bytes / (fms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}


def _forced_flow(g, n: int, dtype, tracers: int = 0):
    """The configurations' flow: random-phase k^4 exp(-2(k/4)^2) velocity (k = 0.5, seed 2026),
    ConstantPower eps = 0.1 on shells 1-2, nu for k_max eta = 1.5, a scalar with mean gradient
    G = 1 along z and zero initial fluctuation, IF-RK4; `tracers` seeded uniform tracers."""
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=2 * math.pi, dealiasing=g.TwoThirds(), element_type=dtype)
    cond = g.FlowConditions(viscosity=g.predicted_viscosity(dom, 0.1, kmax_eta=1.5), schmidt=1.0,
                            velocity_forcing=g.ConstantPower(dom, injection_rate=0.1, lowest_shell=1, highest_shell=2))
    u = g.random_phase_velocity(n, seed=2026)
    p = np.zeros((n, n, n // 2 + 1), dtype=np.complex128)
    tr = g.Tracers(dom, number_of_particles=tracers, seed=2028) if tracers else None
    return g.Simulation(dom, conditions=cond, state=g.ResolvedFields(u, p), mean_gradient=1.0,
                        field_integration_scheme=g.IFLowStorageRK4, tracers=tr, keep_physical=tracers > 0)


def time_iterations(g, n: int, m: int, dtype, steps: int, dt: float) -> dict:
    """Whole iterations of the run loop (advance!: step + deposit, traversal rebuild, Heun push and
    sample of M tracers, the blow-up guards' readbacks), wall clock, after the autotuner (as for
    the headline figure) and 2 untimed iterations."""
    t0 = time.perf_counter()
    with _forced_flow(g, n, dtype, tracers=m) as sim:
        sim.autotune(reps=2)
        setup = time.perf_counter() - t0
        for _ in range(2):
            sim.advance(dt)
        sim.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            sim.advance(dt)
        sim.synchronize()
        wall = time.perf_counter() - t0
        mem = sim.memory_bytes()
    per = wall / steps
    return {"grid": n, "tracers": m, "dtype": np.dtype(dtype).name, "scheme": "IFLowStorageRK4", "dt": dt,
            "iterations": steps, "ms_per_iteration": round(per * 1e3, 3),
            "gridpoint_steps_per_s": round(n ** 3 / per, 1), "tracer_steps_per_s": round(m / per, 1),
            "memory_GiB": round(mem / 2 ** 30, 2), "setup_s": round(setup, 1)}


PUBLISHED_MS_PER_STEP = 14.4   # docs/src/validation.md:211-215 (AMD RX 7900 XTX, 265 s / 18,400 steps)


def time_published(g, steps: int, dealiasing: str = "hou_li") -> dict:
    """The reference's only published GPU step time, like for like: 128^3 FP64, IFSSPRK3,
    HouLi(order 36, strength 63.463), ConstantPower eps = 1 on shells 1-2, nu = 9.032e-3 (k_max eta
    = 1.5 on the resolved edge), fixed dt = 2e-3, no scalar, 10^4 AB(4) tracers with Catmull-Rom
    (docs/src/validation.md:29-39).  Whole iterations (advance: step + deposit, traversal rebuild,
    tracer push and sample, blow-up guard readback), wall clock after the autotuner and 20 untimed
    ones.  The initial
    field is the seeded random-phase one, not the reference's TurTLE hand-off state."""
    n, dt, m = 128, 2e-3, 10_000
    dl = g.HouLi(order=36, strength=63.463) if dealiasing == "hou_li" else g.PhaseShift()
    dom = g.FixedDomain(grid_points_per_axis=n, side_length=2 * math.pi, dealiasing=dl, element_type=np.float64)
    nu = g.predicted_viscosity(g.FixedDomain(grid_points_per_axis=n, side_length=2 * math.pi,
                                             dealiasing=g.HouLi(order=36, strength=63.463), element_type=np.float64),
                               1.0, kmax_eta=1.5)
    cond = g.FlowConditions(viscosity=nu, schmidt=1.0,
                            velocity_forcing=g.ConstantPower(dom, injection_rate=1.0, lowest_shell=1, highest_shell=2))
    with g.Simulation(dom, conditions=cond, state=g.ResolvedFields(g.random_phase_velocity(n, seed=2026), None),
                      field_integration_scheme=g.IFSSPRK3, tracers=g.Tracers(dom, number_of_particles=m, seed=2028),
                      particle_integration_scheme=g.AB(4)) as sim:
        sim.autotune(reps=2)
        for _ in range(20):
            sim.advance(dt)
        sim.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            sim.advance(dt)
        sim.synchronize()
        per = (time.perf_counter() - t0) / steps
    ms = per * 1e3
    out = {"grid": n, "dtype": "f64", "dealiasing": repr(dl), "scheme": "IFSSPRK3", "forcing": "ConstantPower eps 1, shells 1-2",
           "nu": round(nu, 6), "dt": dt, "tracers": m, "particle_scheme": "AB(4)", "iterations": steps,
           "ms_per_step": round(ms, 3), "gridpoint_steps_per_s": round(n ** 3 / per, 1)}
    if dealiasing == "hou_li":
        out.update({"published_ms_per_step": PUBLISHED_MS_PER_STEP,
                    "published_hardware": "AMD Radeon RX 7900 XTX (docs/src/validation.md:211-215)",
                    "speedup_vs_published": round(PUBLISHED_MS_PER_STEP / ms, 2)})
    return out


def accuracy_report(g, n: int, steps: int, dt: float) -> dict:
    """FP32-bulk against a plain FP64 run of the same flow on the same GPU, per step: relative RMS
    distance of u and phi, relative error of k, eps, <phi^2>, eps_phi, and the largest relative
    error over the live shells of E(k) and E_phi(k)."""
    keys = ("kinetic_energy", "dissipation", "scalar_variance", "scalar_dissipation")
    out = {"grid": n, "steps": steps, "dt": dt, "per_step": []}

    def spectrum_error(a, b, which):
        _, ea = a.energy_spectrum(which)
        _, eb = b.energy_spectrum(which)
        live = eb > 1e-12 * eb.max()
        return float(np.max(np.abs(ea[live] - eb[live]) / eb[live]))

    with _forced_flow(g, n, np.float32) as a, _forced_flow(g, n, np.float64) as b:
        for s in range(1, steps + 1):
            a.advance(dt)
            b.advance(dt)
            d = a.distance(b)
            ra, rb = a.reductions(), b.reductions()
            row = {"step": s, "u": float(f"{d['velocity_relative']:.3e}"), "phi": float(f"{d['scalar_relative']:.3e}")}
            row.update({k: float(f"{abs(ra[k] - rb[k]) / abs(rb[k]):.3e}") for k in keys})
            row["E_u"] = float(f"{spectrum_error(a, b, 0):.3e}")
            row["E_phi"] = float(f"{spectrum_error(a, b, 1):.3e}")
            out["per_step"].append(row)
    return out


def _parseval_weights(n: int) -> np.ndarray:
    """1 on the kx = 0 and kx = N/2 planes, 2 inside: <a, b>_w = sum w Re(conj(a) b)."""
    w = np.full(n // 2 + 1, 2.0)
    w[0] = w[-1] = 1.0
    return w


def time_adjoint(g, n: int = 128, steps: int = 200, every: int = 20, directions: int = 16) -> dict:
    """The adjoint workload: gradient of k(T) + <phi^2>(T) after `steps` forced IF-RK4 steps with
    respect to u_hat(0), checkpoints every `every` steps, FP64 and FP32-bulk.  The FP64 gradient is
    checked against fourth-order central differences of device runs along `directions` seeded
    random-phase solenoidal directions (unit weighted norm, step 200 against |u_hat|_w ~ N^3)."""
    dt = 2e-3
    w = _parseval_weights(n)

    def inner(a, b):
        return float(np.sum(w * np.real(np.conj(a) * b)))

    def run(dtype):
        dom = g.FixedDomain(grid_points_per_axis=n, side_length=2 * math.pi, dealiasing=g.TwoThirds(),
                            element_type=dtype)
        cond = g.FlowConditions(viscosity=g.predicted_viscosity(dom, 0.1, kmax_eta=1.5), schmidt=1.0,
                                velocity_forcing=g.ConstantPower(dom, injection_rate=0.1, lowest_shell=1,
                                                                 highest_shell=2))
        u = g.random_phase_velocity(n, seed=2031)
        p = np.zeros((n, n, n // 2 + 1), dtype=np.complex128)
        return g.Simulation(dom, conditions=cond, state=g.ResolvedFields(u, p), mean_gradient=1.0), u, p

    out = {"grid": n, "steps": steps, "checkpoint_every": every, "scheme": "IFLowStorageRK4", "dt": dt,
           "objective": "k(T) + <phi^2>(T)"}
    grads = {}
    for dtype, key in ((np.float64, "fp64"), (np.float32, "fp32_bulk")):
        sim, u, p = run(dtype)
        with sim:
            sim.adjoint_gradient(dt, 2, 1)   # first-call allocations and code loads
            t0 = time.perf_counter()
            J, gu, _ = sim.adjoint_gradient(dt, steps, every)
            tg = time.perf_counter() - t0
            slots_uniform = steps // every + (steps % every > 0) + every
            # binomial (revolve) checkpointing from the same state, with a third of those state
            # slots: the same bits (the gradient leaves the state where it was)
            slots = max(2, slots_uniform // 3)
            t0 = time.perf_counter()
            Jr, gur, _ = sim.adjoint_gradient(dt, steps, snapshots=slots)
            tr = time.perf_counter() - t0
            sim.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                sim.step_fields(dt, graph=True)
            sim.synchronize()
            tf = time.perf_counter() - t0
            r = {"J": J, "gradient_s": round(tg, 4), "forward_ms_per_step": round(tf * 1e3 / steps, 4),
                 "gradient_over_forward": round(tg / tf, 3), "state_slots": slots_uniform}
            rep = 1
            while math.comb(slots + rep, rep) < steps:
                rep += 1
            r["revolve"] = {"state_slots": slots, "gradient_s": round(tr, 4), "repetitions": rep,
                            "forward_steps": rep * steps - math.comb(slots + rep, rep - 1),
                            "bitwise_equal": bool(Jr == J and np.array_equal(gur, gu))}
            if key == "fp64" and directions:
                def J_of(u0):
                    sim.set_fields(g.ResolvedFields(u0, p))
                    for _ in range(steps):
                        sim.step_fields(dt, graph=True)
                    red = sim.reductions()
                    return red["kinetic_energy"] + red["scalar_variance"]

                e, worst = 200.0, 0.0
                for k in range(directions):
                    v = g.random_phase_velocity(n, seed=9100 + k)
                    v /= math.sqrt(inner(v, v))
                    jp2, jp1, jm1, jm2 = (J_of(u + s_ * e * v) for s_ in (2, 1, -1, -2))
                    fd = (-jp2 + 8 * jp1 - 8 * jm1 + jm2) / (12 * e)
                    worst = max(worst, abs(inner(gu, v) - fd) / abs(fd))
                r.update({"fd_directions": directions, "fd_max_rel": float(f"{worst:.3e}")})
            grads[key] = (J, gu)
            out[key] = r
    J64, g64 = grads["fp64"]
    J32, g32 = grads["fp32_bulk"]
    out["fp32_bulk"]["J_rel_vs_fp64"] = float(f"{abs(J32 - J64) / J64:.3e}")
    out["fp32_bulk"]["grad_rel_vs_fp64"] = float(f"{np.linalg.norm(g32 - g64) / np.linalg.norm(g64):.3e}")
    return out


def copy_calibration(H, nbytes: int = 1 << 30) -> None:
    """Kernels of known traffic for scaling the FETCH_SIZE / WRITE_SIZE counters: two
    device-to-device copies (1 GiB read + 1 GiB written each), then, where torch is importable,
    a 1 GiB fill (written only) and a sum over it (read only) as an independent cross-check."""
    H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    H.hipFree.argtypes = [ctypes.c_void_p]
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    H.hipMalloc(ctypes.byref(a), nbytes)
    H.hipMalloc(ctypes.byref(b), nbytes)
    for _ in range(2):
        H.hipMemcpy(b, a, nbytes, 3)   # hipMemcpyDeviceToDevice
    H.hipDeviceSynchronize()
    H.hipFree(a)
    H.hipFree(b)
    try:
        import torch
    except ImportError:
        return
    x = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda")
    x.fill_(1.0)
    total = float(x.sum())
    torch.cuda.synchronize()
    if abs(total - nbytes // 4) > 1e-3 * (nbytes // 4):
        raise RuntimeError("the calibration sum is off")
    del x


def kernel_table(res: dict, n: int, real_bytes: int, stages: int, ssp: bool = False) -> dict:
    ab = algorithmic_bytes(n, real_bytes, stages, ssp=ssp)
    out = {}
    for i, name in enumerate(KIND_NAMES[:7]):
        cnt = int(res["kernel_launches"][i])
        if cnt == 0:
            continue
        avg_ms = float(res["kernel_ms"][i]) / cnt
        b = ab.get(name, 0)
        sym = KERNEL_SYMBOLS[i]
        if sym == "k_xpass" and (real_bytes, n) in ((4, 256), (4, 512), (4, 1024), (8, 256), (8, 512)):
            sym = "k_xpass_wave"   # the register-resident x pass: FP32 at N = 256, 512, 1024; FP64 at 256, 512
        out[name] = {"symbol": sym, "launches": cnt, "avg_ms": round(avg_ms, 5),
                     "share": round(float(res["kernel_ms"][i]) / max(res["ms_total"], 1e-9), 4),
                     "algorithmic_bytes": b, "achieved_GBs": round(b / (avg_ms * 1e-3) / 1e9, 1) if b else None}
    return out


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(steps: int = 100) -> dict:
    """The FP64 oracle on the host cores, on the CPU-runnable configuration: 64^3, nu = 0.01,
    unforced, a mean-gradient scalar from zero, IF-RK4 at dt = 2e-3 for `steps` steps with 4096
    Heun tracers; and one IF-RK4 step at 128^3 (the grid of the reference's published CPU figures,
    0.202-0.248 s/step on a 16-thread i7-12700K, validation/RUNBOOK.md:257-260,282-284)."""
    from oracle import dns
    from oracle import tracers as trc

    import gbg_amd as g
    cores = min(os.cpu_count() or 1, 16)
    dns.set_workers(cores)
    L = 2 * math.pi
    n, dt = 64, 2e-3
    d = dns.Domain(n)
    oc = dns.Conditions(viscosity=0.01, schmidt=1.0, mean_gradient=1.0)
    u = g.random_phase_velocity(n, seed=2026)
    p = np.zeros((n, n, n // 2 + 1), dtype=np.complex128)
    X = g.uniform_positions(4096, L, 2028).T.copy()
    t0 = time.perf_counter()
    dep = dns.deposit(d, u, p)
    v = trc.interpolate(dep.u, X, L)
    for _ in range(steps):
        u, p, dep = dns.step(d, oc, "rk4", u, p, dt, dep)
        X, v = trc.heun_advect(X, v, dep.u, dt, L)
    sec = time.perf_counter() - t0
    n2 = 128
    d2 = dns.Domain(n2)
    u2 = g.random_phase_velocity(n2, seed=2026)
    p2 = np.zeros((n2, n2, n2 // 2 + 1), dtype=np.complex128)
    dep2 = dns.deposit(d2, u2, p2)
    t1 = time.perf_counter()
    dns.step(d2, dns.Conditions(viscosity=0.01, schmidt=1.0, forcing=(0.1, 1, 2), mean_gradient=1.0), "rk4", u2, p2,
             dt, dep2)
    sec128 = time.perf_counter() - t1
    return {"value": round(n ** 3 * steps / sec / 1e6, 3), "unit": "Mpoint*steps/s", "cores": cores, "kind": "port",
            "cpu_model": _cpu_model(),
            "sample": f"the 64^3 configuration: {steps} IF-RK4 steps of dt 2e-3 (nu 0.01, unforced, mean-gradient "
                      f"scalar) with 4096 Heun tracers in the FP64 numpy oracle, {sec:.1f} s ({sec / steps * 1e3:.0f} "
                      f"ms/step); transforms in scipy.fft (pocketfft) on {cores} threads, pointwise work and the "
                      f"tracer gathers in numpy on one thread",
            "s_per_step_128": round(sec128, 3),
            "sample_128": f"one IF-RK4 step (+ deposit) at 128^3 with forcing: {sec128:.2f} s/step "
                          f"({128 ** 3 / sec128 / 1e6:.2f} Mpoint*steps/s); the reference's own CPU path ran 0.202-0.248 "
                          f"s/step at 128^3 on 16 threads (validation/RUNBOOK.md:257-260,282-284)"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain-FP64 comparison run")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=100, help="oracle steps of the 64^3 CPU baseline")
    ap.add_argument("--no-published", action="store_true",
                    help="skip the reference's published configuration (128^3 FP64 HouLi IFSSPRK3 + 10^4 AB(4) tracers)")
    ap.add_argument("--no-tracers", action="store_true", help="skip the tracer throughput figure")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 96-box ensemble figure")
    ap.add_argument("--ensemble-steps", type=int, default=200)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the 128^3 adjoint figure")
    ap.add_argument("--no-iterations", action="store_true",
                    help="skip the whole-iteration figures (256^3 FP64 + 1M tracers, 1024^3 FP32-bulk + 16M tracers)")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the FP32-bulk vs FP64 error report")
    ap.add_argument("--no-schemes", action="store_true",
                    help="skip the other field schemes (IFLowStorageRK3, IFSSPRK3, IFHeun) on the same workload")
    ap.add_argument("--adjoint-directions", type=int, default=16,
                    help="finite-difference directions for the adjoint check (0: time only)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one 1 GiB device copy (the reference kernel for counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this DNS step runs on a single MI355X and --gpus {args.gpus} is not "
              "supported (the slab-decomposed multi-GPU path is a design note, not built)", file=sys.stderr)
        return 2
    import __graft_entry__
    g = __graft_entry__.load_package()
    H = hip()
    n, stages = args.n, 5
    print(f"bench: {n}^3 IF-RK4 FP32-bulk, warmup {args.warmup}, steps {args.steps}", file=sys.stderr, flush=True)
    r32 = time_gpu(g, n, np.float32, args.steps, args.warmup, H)
    # the headline: the step replayed as its captured graph, with no per-kernel events
    ms32 = r32["ms_graph"] / args.steps
    value = n ** 3 / (ms32 * 1e-3) / 1e6
    timing = {"graph_ms_per_step": round(ms32, 3), "eager_ms_per_step": round(r32["ms_plain"] / args.steps, 3),
              "eager_with_kernel_events_ms_per_step": round(r32["ms_total"] / args.steps, 3)}
    k32 = kernel_table(r32, n, 4, stages)
    fp64 = None
    if not args.no_fp64:
        print(f"bench: {n}^3 IF-RK4 plain FP64", file=sys.stderr, flush=True)
        r64 = time_gpu(g, n, np.float64, max(3, args.steps // 2), 1, H)
        ms64 = r64["ms_graph"] / max(3, args.steps // 2)
        fp64 = {"value": round(n ** 3 / (ms64 * 1e-3) / 1e6, 2), "ms_per_step": round(ms64, 3),
                "fp32_bulk_speedup": round(ms64 / ms32, 3), "kernels": kernel_table(r64, n, 8, stages)}
    dom_name = max(k32, key=lambda k: k32[k]["share"])
    dk = k32[dom_name]
    achieved = dk["achieved_GBs"]
    step_bytes = sum(dk_["algorithmic_bytes"] * dk_["launches"] for dk_ in k32.values()) / args.steps
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dk["symbol"],
                "kernel_role": dom_name, "algorithmic_bytes_per_launch": dk["algorithmic_bytes"],
                "step_algorithmic_GBs": round(step_bytes / (ms32 * 1e-3) / 1e9, 1),
                "survey_model_frac": round(value * 1e6 * 811.0 / (HBM_PEAK_GBS * 1e9), 4)}
    traffic_file = ROOT / "profiles" / "traffic.json"
    if traffic_file.exists():
        try:
            t = json.loads(traffic_file.read_text()).get(dk["symbol"])
            if t and t.get("n") == n:
                roofline["traffic"] = t["bytes_per_launch"]
                roofline["traffic_source"] = "rocprofv3 FETCH_SIZE/WRITE_SIZE, copy-calibrated: " + t["kernel"]
        except (ValueError, KeyError):
            pass
    schemes = None
    if not args.no_schemes:
        print(f"bench: {n}^3 FP32-bulk, IFLowStorageRK3, IFSSPRK3 and IFHeun", file=sys.stderr, flush=True)
        schemes = {}
        for name, stages in (("IFLowStorageRK3", 3), ("IFSSPRK3", 3), ("IFHeun", 2)):
            k = max(5, args.steps // 2)
            r = time_gpu(g, n, np.float32, k, 1, H, scheme=getattr(g, name))
            ms = r["ms_graph"] / k
            schemes[name] = {"value": round(n ** 3 / (ms * 1e-3) / 1e6, 2), "ms_per_step": round(ms, 3),
                             "steps": k, "kernels": kernel_table(r, n, 4, stages, ssp=name != "IFLowStorageRK3")}
    tracers = None
    if not args.no_tracers:
        print("bench: tracers, 256^3 field, 2^20 tracers", file=sys.stderr, flush=True)
        tracers = [time_tracers(g, 256, 1 << 20, np.float32, 5, H), time_tracers(g, 256, 1 << 20, np.float64, 5, H)]
    ensemble = None
    if not args.no_ensemble:
        print(f"bench: ensemble, 96 boxes (32^3, 48^3, 64^3), IF-RK3, {args.ensemble_steps} steps",
              file=sys.stderr, flush=True)
        ensemble = time_ensemble(g, args.ensemble_steps, H)
    iterations = None
    if not args.no_iterations:
        print("bench: whole iterations, 256^3 FP64 + 2^20 tracers and 1024^3 FP32-bulk + 2^24 tracers",
              file=sys.stderr, flush=True)
        iterations = [time_iterations(g, 256, 1 << 20, np.float64, 20, 2e-3),
                      time_iterations(g, 1024, 1 << 24, np.float32, 10, 5e-4)]
    accuracy = None
    if not args.no_accuracy:
        print(f"bench: FP32-bulk vs FP64 error report at {n}^3, 10 steps", file=sys.stderr, flush=True)
        accuracy = accuracy_report(g, n, 10, 1e-3)
    adjoint = None
    if not args.no_adjoint:
        print(f"bench: adjoint, 128^3, 200 IF-RK4 steps, checkpoints every 20, "
              f"{args.adjoint_directions} FD directions", file=sys.stderr, flush=True)
        adjoint = time_adjoint(g, directions=args.adjoint_directions)
    if args.copy_calibration:
        copy_calibration(H)
    published = None
    if not args.no_published:
        print("bench: the published configuration, 128^3 FP64 HouLi IFSSPRK3 + 10^4 AB(4) tracers (and PhaseShift)",
              file=sys.stderr, flush=True)
        published = [time_published(g, 200), time_published(g, 100, "phase_shift")]
    cpu = None
    if not args.no_cpu_baseline:
        print(f"bench: cpu baseline, 64^3 x {args.cpu_steps} steps + 4096 tracers, and one 128^3 step",
              file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.cpu_steps)
    line = {
        "metric": "Mpoint*steps/s at 512^3 IF-RK4, FP32-bulk vs FP64 (+% HBM roofline)",
        "value": round(value, 2), "unit": "Mpoint*steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms32, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (FP64 integrating factors, twiddles, reductions)",
        "data": "synthetic: random-phase k^4 exp(-2(k/4)^2) velocity, seed 2026, phi = 0",
        "config": {"workload": f"{n}^3 forced isotropic turbulence + passive scalar with mean gradient, IF-RK4, "
                               f"dt 1e-3, ConstantPower eps 0.1 shells 1-2, nu {r32['nu']:.4e}, tracers off",
                   "grid": n, "scheme": "IFLowStorageRK4", "parallelism": "single GPU"},
        "roofline": roofline, "cpu_baseline": cpu, "fp64": fp64, "kernels": k32, "schemes": schemes,
        "timing": timing,
        "wall_ms_per_step": round(r32["wall_s"] * 1e3 / args.steps, 3),
        "memory_GiB": round(r32["memory_bytes"] / 2 ** 30, 2),
        "autotuned_schedule": r32["schedule"],
        "tracers": tracers,
        "ensemble": ensemble,
        "adjoint": adjoint,
        "iterations": iterations,
        "accuracy": accuracy,
        "published_configuration": published,
    }
    # a compact summary last, so that a reader of the line's tail sees the figures that matter
    line["summary"] = {
        "fp32_bulk_ms_per_step": round(ms32, 3),
        "fp64_ms_per_step": fp64["ms_per_step"] if fp64 else None,
        "fp32_bulk_speedup_over_fp64": fp64["fp32_bulk_speedup"] if fp64 else None,
        "fp64_x_pass_ms": fp64["kernels"]["x_pass_products"]["avg_ms"] if fp64 else None,
        "roofline_frac": roofline["frac"], "survey_model_frac": roofline["survey_model_frac"],
        "published_hou_li_128_ms_per_step": published[0]["ms_per_step"] if published else None,
        "published_reference_ms_per_step": PUBLISHED_MS_PER_STEP,
        "cpu_baseline_Mpoint_steps_per_s": cpu["value"] if cpu else None,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
