"""StochasticOU, the Eswaran-Pope forcing (src/Forcing.jl:1120-1413): an additive force on a band
of shells whose Fourier amplitudes are independent complex Ornstein-Uhlenbeck processes.

The process lives on the host, on a companion grid of 2 (highest_shell + 1) points per side over
the same box.  That grid holds the band strictly below its Nyquist frequency and costs a few
kilobytes. Once per step, before the step, the process takes its exact transition over dt,
    a <- alpha a + sqrt(sigma^2 (1 - alpha^2)) xi,  alpha = exp(-dt / T_L),
(src/Forcing.jl:1360-1385). Here xi is a fresh stationary draw: a real Gaussian field on the
companion grid, transformed, cut to the band, projected transverse to kappa for the velocity, and
scaled by the square root of the expected retained fraction. The band force is then uploaded to
the device, which adds it at every stage (``hdns_band_source_upload``).

The draw stream is keyed like the reference's, by (seed, draw count, field), so a realization is a
pure function of the three and resumes from the count.  The generator differs: the reference
draws with Julia's Xoshiro and its normal sampler, which this image cannot reproduce, so this
uses numpy's counter-based Philox.  A run's draws therefore match the reference's statistically,
not number for number.
"""
from __future__ import annotations

import math

import numpy as np

_MASK64 = (1 << 64) - 1


class StochasticOU:
    """StochasticOU(domain; forcing_variance, correlation_time, lowest_shell=1, highest_shell=2,
    seed) (src/Forcing.jl:1298-1348), with the reference's guards."""

    def __init__(self, domain, *, forcing_variance: float, correlation_time: float, lowest_shell: int = 1,
                 highest_shell=2, seed: int | None = None):
        cast = domain.element_type
        sigma2, tl = float(cast(forcing_variance)), float(cast(correlation_time))
        if not (math.isfinite(sigma2) and sigma2 >= 0):
            raise ValueError(f"`forcing_variance` must be finite and non-negative, got {sigma2}: it is the stationary "
                             "mean square of the term the forcing adds, and its square root sets the amplitude of "
                             "every draw")
        if not (math.isfinite(tl) and tl > 0):
            raise ValueError(f"`correlation_time` must be finite and positive, got {tl}: it is the "
                             "Ornstein-Uhlenbeck relaxation time T_L, and each step decays the amplitudes by "
                             "exp(-dt/T_L)")
        if lowest_shell < 1:
            raise ValueError(f"`lowest_shell` must be at least 1, got {lowest_shell}: shell 0 is the mean "
                             "(kappa = 0) mode, which a band forcing cannot legitimately drive")
        if highest_shell == math.inf:
            raise ValueError("StochasticOU needs a finite `highest_shell`, got Inf: it adds its process rather than "
                             "multiplying the field, so its band must stay below the dealias cutoff, and the "
                             "coarse grid carrying the process is sized from the band's top shell")
        if not float(highest_shell).is_integer() or highest_shell < lowest_shell:
            raise ValueError(f"`highest_shell` must be an integer shell at least `lowest_shell`, got {highest_shell}")
        hi = int(highest_shell)
        n = domain.grid_points_per_axis
        if 9 * (2 * hi + 1) ** 2 > 4 * n * n:   # (hi + 1/2) k0 past (2/3)(N/2) k0
            raise ValueError(f"the forcing band (shells {lowest_shell}-{hi}) reaches past the dealias cutoff; an "
                             "additive force would fill modes the dealiasing holds empty. Lower the shells, or "
                             "enlarge the grid.")
        self.forcing_variance = sigma2
        self.correlation_time = tl
        self.lowest_shell = int(lowest_shell)
        self.highest_shell = hi
        self.seed = int(seed) if seed is not None else int(np.random.default_rng().integers(-(2 ** 63), 2 ** 63))
        self.companion_points = 2 * (hi + 1)
        self.element_type = cast
        nc = self.companion_points
        m = np.fft.fftfreq(nc, 1.0 / nc).astype(np.int64)              # signed bins of the companion grid
        four = 4 * (m[:, None, None] ** 2 + m[None, :, None] ** 2 + m[None, None, :] ** 2)
        full = (four >= (2 * self.lowest_shell - 1) ** 2) & (four <= (2 * hi + 1) ** 2)
        self.modes = int(np.count_nonzero(full))   # both members of every pair (src/Forcing.jl:1288-1295)
        if self.modes == 0:
            raise ValueError(f"the band (shells {lowest_shell}-{hi}) holds no lattice wavevector, so the forcing "
                             "would drive nothing")
        self.band = full[:, :, : nc // 2 + 1]                           # the half spectrum [kz, ky, kx]

    def __repr__(self) -> str:
        return (f"StochasticOU(forcing_variance {self.forcing_variance}, correlation_time {self.correlation_time}, "
                f"shells {self.lowest_shell}-{self.highest_shell}, seed {self.seed})")


class OUProcess:
    """The forcing state of one field (StochasticOUState, src/Forcing.jl:1236-1264): the process on the
    companion grid's half spectrum, unnormalised ([3, Nc, Nc, Nc/2+1] for the velocity,
    [Nc, Nc, Nc/2+1] for the scalar) and the number of increments drawn."""

    def __init__(self, forcing: StochasticOU, order: int, process: np.ndarray | None = None, draws: int = 0):
        self.forcing = forcing
        self.order = int(order)                      # 1: velocity, 0: scalar
        nc = forcing.companion_points
        shape = ((3,) if order == 1 else ()) + (nc, nc, nc // 2 + 1)
        self.process = np.zeros(shape, dtype=np.complex128) if process is None else np.array(process, np.complex128)
        self.draws = int(draws)
        if process is None:                          # the stationary start (src/Forcing.jl:1388-1392)
            self._advance(0.0)

    def increment(self, draws: int) -> np.ndarray:
        """The stationary unit draw xi of position `draws` in the stream: a real Gaussian field on the
        companion grid, transformed (unnormalised), cut to the band, projected transverse to kappa
        for the velocity, and scaled to unit expected mean square."""
        f = self.forcing
        nc = f.companion_points
        key = np.array([f.seed & _MASK64, ((draws << 2) | self.order) & _MASK64], dtype=np.uint64)
        rng = np.random.Generator(np.random.Philox(key=key))
        shape = ((3,) if self.order == 1 else ()) + (nc, nc, nc)
        xi = np.fft.rfftn(rng.standard_normal(shape), axes=(-3, -2, -1)) * f.band
        if self.order == 1:
            m = np.fft.fftfreq(nc, 1.0 / nc)
            kz, ky, kx = m[:, None, None], m[None, :, None], np.arange(nc // 2 + 1)[None, None, :]
            k2 = kx * kx + ky * ky + kz * kz
            k2 = np.where(k2 == 0, 1.0, k2)
            s = (kx * xi[0] + ky * xi[1] + kz * xi[2]) / k2
            xi = np.stack([xi[0] - kx * s, xi[1] - ky * s, xi[2] - kz * s])
        retained = (2 if self.order == 1 else 1) * f.modes
        return xi / math.sqrt(retained / nc ** 3)

    def _advance(self, alpha: float) -> None:
        f = self.forcing
        a = float(alpha)
        self.process = a * self.process + math.sqrt(f.forcing_variance * (1.0 - a * a)) * self.increment(self.draws)
        self.draws += 1

    def advance(self, dt: float) -> None:
        """prepare_forcing! (src/Forcing.jl:1395-1405): one exact transition over dt, with alpha
        formed in the run's element type."""
        T = self.forcing.element_type
        self._advance(float(np.exp(-T(dt) / T(self.forcing.correlation_time))))

    def band_source(self, n: int) -> np.ndarray:
        """The force in the run's unnormalised spectral convention on the cube |m| <= r of the upload
        layout ([c][mz + r][my + r][mx], c = 3 or 1): the companion's modes times (N / Nc)^3, which
        keeps the physical field."""
        f = self.forcing
        r, nc = f.highest_shell, f.companion_points
        idx = np.arange(-r, r + 1) % nc
        p = self.process if self.order == 1 else self.process[None]
        cube = p[:, idx[:, None], idx[None, :], : r + 1] * (n / nc) ** 3
        return np.ascontiguousarray(cube)

    def full_source(self, n: int) -> np.ndarray:
        """The same force in the run's full half spectrum ([3,] N, N, N/2+1), for the oracle."""
        f = self.forcing
        r, nc = f.highest_shell, f.companion_points
        shape = ((3,) if self.order == 1 else ()) + (n, n, n // 2 + 1)
        out = np.zeros(shape, dtype=np.complex128)
        ms = np.arange(-r, r + 1)
        p = self.process if self.order == 1 else self.process[None]
        o = out if self.order == 1 else out[None]
        o[:, (ms % n)[:, None], (ms % n)[None, :], : r + 1] = p[:, (ms % nc)[:, None], (ms % nc)[None, :], : r + 1] * (n / nc) ** 3
        return out
