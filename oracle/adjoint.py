"""FP64 adjoint (reverse-mode derivative) of the oracle step (test infrastructure only).

The reference has no adjoint: the north-star item asks for a hand-written rrule of one IF-RK step
with checkpointing, checked against central finite differences.  This module restates, in numpy,
the transpose of every linear operation the step applies, so that the device adjoint can be held
to it mode by mode, and so that both can be held to finite differences of the forward oracle.

Inner products.  A real field is stored as its half spectrum f_hat[.., kz, ky, kx]; the real
inner product of two fields is  <a, b>_w = sum_k w_k Re(conj(a_k) b_k)  with the Parseval weights
w (1 on kx = 0 and kx = N/2, 2 inside), which equals the full-spectrum sum.  Gradients are taken
with respect to this product: dJ = <grad J, delta>_w.  On the physical grid the product is the
plain sum over points.  With the oracle's transforms (forward unnormalised, inverse with 1/N^3):
    rfft3^T = N^3 irfft3,      irfft3^T = rfft3 / N^3
(irfft3 ignores the anti-Hermitian part of the kx = 0 plane, so these hold for any half
spectrum).  The mask, the Leray projection, the zeroed mean mode, the integrating factor and the
reality projection are real and symmetric per mode, so each is its own transpose; i k x (curl)
is its own transpose and (i k .)^T = -i k.

Linearisations (u, w, phi the stage's physical fields, m the physical adjoint of a velocity
product, s that of a flux):
    N_u = Z P M rfft3(u x w) + A 1_b u_hat,    A = eps / (2 e_b),  e_b = <1_b u, u>_w / (2 N^6)
        -> lam_u += rfft3(w x m)/N^3 + curl(rfft3(m x u))/N^3,  m = N^3 irfft3(M P Z mu_u)
           lam_u += A 1_b mu_u - (A / e_b) (<mu_u, 1_b u_hat>_w / N^6) 1_b u_hat
    N_phi = -M i k . rfft3(u phi) - G u_z_hat
        -> lam_u += rfft3(phi s)/N^3, lam_phi += rfft3(s . u)/N^3,  s = N^3 irfft3(i k M mu_phi)
           lam_u_z += -G mu_phi
The stage recurrences (h1 = a h + dt N; q' = E (q + b h1); h' = E h1) transpose to
    lam_h1 = b E lam_q' + E lam_h';  lam_q = E lam_q' + dN^T (dt lam_h1);  lam_h = a lam_h1.

The state lives in the retained modes (the 2/3 ball): a perturbation outside it is not a state.
So the adjoints are restricted to the ball (the transposed products are masked).  Components
outside the ball never flow back into it (every path into the products goes through the mask),
so the restriction changes nothing inside.
"""
from __future__ import annotations

import numpy as np

from . import dns


def inner(d: dns.Domain, a: np.ndarray, b: np.ndarray) -> float:
    """<a, b>_w summed over components (FP64, plain numpy sum: a derivative check, not a baseline)."""
    return float(np.sum(d.weights * np.real(np.conj(a) * b)))


def objective(d: dns.Domain, uh: np.ndarray, ph: np.ndarray | None) -> float:
    """J = k + <phi^2> = <u,u>_w/(2 N^6) + <phi,phi>_w/N^6 (the reductions a27)."""
    j = 0.5 * inner(d, uh, uh) / d.norm
    if ph is not None:
        j += inner(d, ph, ph) / d.norm
    return j


def objective_gradient(d: dns.Domain, uh: np.ndarray, ph: np.ndarray | None):
    """dJ/du_hat = u_hat/N^6, dJ/dphi_hat = 2 phi_hat/N^6 (with respect to <,>_w)."""
    return uh / d.norm, (2.0 * ph / d.norm if ph is not None else None)


def _stage_record(d, cond, scheme, uh, ph, dt, dep):
    """Forward step, keeping every stage's input state, physical fields and forcing amplitude."""
    a, b, dd = dns.scheme_coefficients(scheme)
    S = len(a)
    rec = []
    uh = uh.copy()
    ph = ph.copy() if ph is not None else None
    if dep is None:
        dep = dns.deposit(d, uh, ph)
    hu = hp = None
    for i in range(S):
        if i == 0:
            u, w, phi = dep.u, dep.w, dep.phi
        else:
            u = dns.irfft3(uh, d.n)
            w = dns.irfft3(dns.curl(d, uh), d.n)
            phi = dns.irfft3(ph, d.n) if ph is not None else None
        amp, eb = 0.0, 0.0
        if cond.forcing is not None:
            amp = dns.constant_power_amplitude(d, cond, uh)
            eb = 0.5 * dns.weighted_inner(d, uh, d.shell_band(cond.forcing[1], cond.forcing[2]))
        rec.append((uh.copy(), None if ph is None else ph.copy(), u, w, phi, amp, eb))
        nu_hat = dns.velocity_rhs(d, cond, uh, u, w)
        np_hat = dns.scalar_rhs(d, cond, uh, u, phi) if ph is not None else None
        hu = dt * nu_hat if i == 0 else a[i] * hu + dt * nu_hat
        sub = dd[i] * dt
        eu = np.exp(-cond.viscosity * sub * d.k2)
        uh = eu * (uh + b[i] * hu)
        if i < S - 1:
            hu = eu * hu
        if ph is not None:
            hp = dt * np_hat if i == 0 else a[i] * hp + dt * np_hat
            ep = np.exp(-cond.scalar_diffusivity * sub * d.k2)
            ph = ep * (ph + b[i] * hp)
            if i < S - 1:
                hp = ep * hp
    return rec, (a, b, dd)


def rhs_vjp(d, cond, uh, u, w, phi, amp, eb, mu_u, mu_p):
    """(dN/dq)^T (mu_u, mu_phi) at one stage: the transposed right-hand sides (module docstring)."""
    n3 = float(d.n) ** 3
    nu_ = d.mask * mu_u
    nu_[:, 0, 0, 0] = 0.0
    nu_ = dns.leray(d, nu_) * d.mask
    m = n3 * dns.irfft3(nu_, d.n)
    lam_u = dns.rfft3(np.cross(w, m, axis=0)) / n3 + dns.curl(d, dns.rfft3(np.cross(m, u, axis=0)) / n3)
    if cond.forcing is not None and amp != 0.0:
        band = d.shell_band(cond.forcing[1], cond.forcing[2])
        proj = inner(d, mu_u, band * uh) / d.norm
        lam_u = lam_u + amp * band * mu_u - (amp / eb) * proj * band * uh
    lam_p = None
    if mu_p is not None:
        kx, ky, kz = d.kvec
        mp = d.mask * mu_p
        xi = np.stack([1j * kx * mp, 1j * ky * mp, 1j * kz * mp])
        s = n3 * dns.irfft3(xi, d.n)
        lam_u = lam_u + dns.rfft3(phi[None] * s) / n3
        lam_p = dns.rfft3(np.sum(s * u, axis=0)) / n3
        if cond.mean_gradient != 0.0:
            lam_u[2] = lam_u[2] - cond.mean_gradient * mu_p
    return lam_u, lam_p


def step_vjp(d: dns.Domain, cond: dns.Conditions, scheme: str, uh: np.ndarray, ph: np.ndarray | None,
             dt: float, lam_u: np.ndarray, lam_p: np.ndarray | None, dep=None):
    """The transpose of one step's derivative at (uh, ph): (lam_u, lam_phi) of the step's output
    to those of its input.  The forward stages are recomputed and kept (the step is the
    checkpointing unit)."""
    rec, (a, b, dd) = _stage_record(d, cond, scheme, uh, ph, dt, dep)
    S = len(a)
    # the closing projections, transposed: u_out = R(P(u_S)), phi_out = R(phi_S)
    lq_u = dns.leray(d, dns.reality_projection(d, lam_u)) * d.mask
    lq_p = dns.reality_projection(d, lam_p) * d.mask if lam_p is not None else None
    lh_u = np.zeros_like(lq_u)
    lh_p = np.zeros_like(lq_p) if lq_p is not None else None
    for i in range(S - 1, -1, -1):
        sub = dd[i] * dt
        eu = np.exp(-cond.viscosity * sub * d.k2)
        lh1_u = b[i] * eu * lq_u + (eu * lh_u if i < S - 1 else 0.0)
        new_u = eu * lq_u
        lh1_p = None
        if lq_p is not None:
            ep = np.exp(-cond.scalar_diffusivity * sub * d.k2)
            lh1_p = b[i] * ep * lq_p + (ep * lh_p if i < S - 1 else 0.0)
            new_p = ep * lq_p
        uh_i, ph_i, u, w, phi, amp, eb = rec[i]
        gu, gp = rhs_vjp(d, cond, uh_i, u, w, phi, amp, eb, dt * lh1_u,
                         dt * lh1_p if lh1_p is not None else None)
        lq_u = (new_u + gu) * d.mask
        lh_u = a[i] * lh1_u * d.mask
        if lq_p is not None:
            lq_p = (new_p + gp) * d.mask
            lh_p = a[i] * lh1_p * d.mask
    return lq_u, lq_p


def run(d, cond, scheme, uh, ph, dt, steps):
    """The forward run: the final state."""
    dep = None
    for _ in range(steps):
        uh, ph, dep = dns.step(d, cond, scheme, uh, ph, dt, dep)
    return uh, ph


def gradient(d: dns.Domain, cond: dns.Conditions, scheme: str, uh0: np.ndarray, ph0: np.ndarray | None,
             dt: float, steps: int, every: int):
    """J(T) and its gradient with respect to u_hat(0) (and phi_hat(0)), by checkpointing: the
    forward run keeps the state every `every` steps; the reverse sweep recomputes each segment
    from its checkpoint, keeping the segment's states, and walks it backwards step by step.
    Returns (J, lam_u0, lam_phi0)."""
    checkpoints = {0: (uh0.copy(), None if ph0 is None else ph0.copy())}
    uh, ph = uh0, ph0
    for s in range(steps):
        uh, ph, _ = dns.step(d, cond, scheme, uh, ph, dt)
        if (s + 1) % every == 0 and s + 1 < steps:
            checkpoints[s + 1] = (uh.copy(), None if ph is None else ph.copy())
    J = objective(d, uh, ph)
    lam_u, lam_p = objective_gradient(d, uh, ph)
    starts = sorted(checkpoints)
    for seg in range(len(starts) - 1, -1, -1):
        s0 = starts[seg]
        s1 = starts[seg + 1] if seg + 1 < len(starts) else steps
        states = [checkpoints[s0]]
        for _ in range(s0, s1 - 1):
            u, p, _ = dns.step(d, cond, scheme, states[-1][0], states[-1][1], dt)
            states.append((u, p))
        for k in range(len(states) - 1, -1, -1):
            lam_u, lam_p = step_vjp(d, cond, scheme, states[k][0], states[k][1], dt, lam_u, lam_p)
    return J, lam_u, lam_p


def solenoidal_direction(d: dns.Domain, seed: int) -> np.ndarray:
    """A random real, dealiased, solenoidal, zero-mean velocity perturbation (half spectrum)."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((3, d.n, d.n, d.n))
    f = dns.rfft3(x) * d.mask
    f[:, 0, 0, 0] = 0.0
    f = dns.leray(d, f)
    return f / np.sqrt(inner(d, f, f))
