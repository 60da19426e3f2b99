"""The allocation-free spectrum, hdns_spectrum_noalloc, through the C ABI: the same output as
hdns_spectrum, bit for bit (both sum the shells in a fixed order with the same kernels), on calls
in a row that share the device's scratch array, the second on a smaller grid."""
import ctypes
import math

import numpy as np
import pytest

from test_gpu_parity import L, make_sim

pytestmark = pytest.mark.gpu


def _spectra(g, sim, which):
    lib = g._lib.lib()
    n = sim.domain.grid_points_per_axis
    nb = int(math.floor(math.sqrt(3.0) * n / 2)) + 2
    want = np.zeros(nb + 3)
    got = np.full(nb + 3, np.nan)
    assert lib.hdns_spectrum(sim._ctx, which, g._lib.dptr(want), nb + 3) == 0
    assert lib.hdns_spectrum_noalloc(sim._ctx, which, g._lib.dptr(got), nb + 3) == 0
    return got, want, nb


@pytest.mark.parametrize("first,second", [((64, np.float32), (32, np.float64)), ((96, np.float64), (48, np.float32))])
def test_the_allocation_free_spectrum_equals_the_spectrum(g, first, second):
    lib = g._lib.lib()
    sims = [make_sim(g, n, dtype, forcing=0.1, seed=11 + i)[0] for i, (n, dtype) in enumerate((first, second))]
    try:
        for sim in sims:                         # the larger grid first, then the smaller
            sim.advance(2e-3)
            for which in (0, 1):
                got, want, nb = _spectra(g, sim, which)
                assert np.array_equal(got, want)
                assert np.all(got[nb:] == 0) and np.all(np.isfinite(got))
                k, e = sim.energy_spectrum(which)
                assert np.sum(e[: nb]) > 0
        got, want, _ = _spectra(g, sims[0], 0)   # and the larger again after the smaller
        assert np.array_equal(got, want)
        n = sims[1].domain.grid_points_per_axis
        nb = int(math.floor(math.sqrt(3.0) * n / 2)) + 2
        short = np.zeros(nb - 1)
        assert lib.hdns_spectrum_noalloc(sims[1]._ctx, 0, g._lib.dptr(short), nb - 1) == 1
        assert b"nshells must be at least" in lib.hdns_last_error(sims[1]._ctx)
    finally:
        for sim in sims:
            sim.close()
