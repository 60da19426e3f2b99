"""FP64 restatement of the reference's spin-up observables (test infrastructure only).

integral_length follows src/Observables.jl:556-581 and derivative_skewness src/Observables.jl:406,
465-477.  The arrays follow oracle/dns.py: u_hat [3, kz, ky, kx] with the unnormalised forward
transform.
"""
from __future__ import annotations

import math

import numpy as np

from . import dns

# C of S = C * 3 <det S> / <S:S>^(3/2) (Betchov 1956): (4/7) sqrt(15/2) (src/Observables.jl:406)
SKEWNESS_ISOTROPY_CONSTANT = (4.0 / 7.0) * math.sqrt(15.0 / 2.0)

# the six independent strain components, in the order the library uses: xx, yy, zz, xy, xz, yz
STRAIN_PAIRS = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))


def strain_physical(d: dns.Domain, uh: np.ndarray) -> np.ndarray:
    """S_ij = (d_i u_j + d_j u_i) / 2 in physical space, [6, z, y, x] in STRAIN_PAIRS order
    (the reference's symmetric_gradient then to_physical!, src/Observables.jl:466-468)."""
    k = d.kvec
    out = []
    for i, j in STRAIN_PAIRS:
        s = 0.5j * (k[i] * uh[j] + k[j] * uh[i])
        out.append(dns.irfft3(s, d.n))
    return np.stack(out)


def det_symmetric(s6: np.ndarray) -> np.ndarray:
    """det of [[a, d, e], [d, b, f], [e, f, c]] per point, s6 = (a, b, c, d, e, f)."""
    a, b, c, dd, e, f = s6
    return a * (b * c - f * f) - dd * (dd * c - e * f) + e * (dd * f - b * e)


def strain_dirichlet(d: dns.Domain, uh: np.ndarray) -> float:
    """<S:S> = (1/2) <|w|^2> = (1/2) sum |k|^2 |u_hat|^2 / N^6 (src/Observables.jl:469-471)."""
    return 0.5 * dns.weighted_inner(d, uh, d.k2)


def derivative_skewness(d: dns.Domain, uh: np.ndarray) -> float:
    """S = C * 3 <det S> / <S:S>^(3/2), 0 for a strain-free field (src/Observables.jl:465-477)."""
    ss = strain_dirichlet(d, uh)
    if ss <= 0:
        return 0.0
    mean_det = float(np.mean(det_symmetric(strain_physical(d, uh))))
    return 3.0 * SKEWNESS_ISOTROPY_CONSTANT * mean_det / ss ** 1.5


def integral_length(kappa: np.ndarray, energy: np.ndarray) -> float:
    """L11 = (3 pi / 4) sum_{kappa>0} E / kappa / sum_{kappa>0} E (src/Observables.jl:567-581)."""
    kappa = np.asarray(kappa, dtype=np.float64)
    energy = np.asarray(energy, dtype=np.float64)
    live = kappa > 0
    total = float(np.sum(energy[live]))
    if not total > 0:
        raise ValueError("this velocity field carries no energy, so it has no integral scale")
    return 3.0 * math.pi / 4.0 * float(np.sum(energy[live] / kappa[live])) / total


def below_scale(d: dns.Domain, uh: np.ndarray, viscosity: float):
    """kinetic_energy_and_dissipation_below_scale (src/Observables.jl:195-264): for l = n L / N,
    k(l) = (sum w |u|^2 - sum w W_l^2 |u|^2) / (2 N^6) and eps(l) = nu (sum w |k|^2 |u|^2 -
    sum w |k|^2 W_l^2 |u|^2) / N^6, W_l(k) = prod_a sinc(k_a l / 2) (sin x / x)."""
    whole = dns.weighted_inner(d, uh)
    dirichlet = dns.weighted_inner(d, uh, d.k2)
    scales, ks, es = [], [], []
    for n in range(1, d.n + 1):
        l = n * d.side_length / d.n
        w2 = 1.0
        for k in d.kvec:
            x = k * l / 2
            w2 = w2 * np.where(x == 0, 1.0, np.sin(x) / np.where(x == 0, 1.0, x)) ** 2
        scales.append(l)
        ks.append((whole - dns.weighted_inner(d, uh, w2)) / 2)
        es.append(viscosity * (dirichlet - dns.weighted_inner(d, uh, d.k2 * w2)))
    return np.array(scales), np.array(ks), np.array(es)
