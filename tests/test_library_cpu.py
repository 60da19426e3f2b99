"""Host-side checks of the C ABI that need no GPU: the library loads, exports every symbol the
header declares, its coefficient and weight tables equal the oracle's, and its argument guards fire
before any device call."""
import ctypes
from pathlib import Path
import math

import numpy as np
import pytest

from oracle import dns
from oracle import tracers as trc


def test_library_exports_every_declared_symbol(g):
    L = g._lib.lib()
    names = g._lib.declared_symbols()
    assert len(names) >= 25
    for name in names:
        assert hasattr(L, name), name


@pytest.mark.parametrize("scheme,code", [("rk3", 3), ("rk4", 4)])
@pytest.mark.parametrize("dtype,prec", [(np.float64, 1), (np.float32, 0)])
def test_scheme_coefficients_equal_the_oracle_rounding(g, scheme, code, dtype, prec):
    got = g._lib.scheme_coefficients(code, prec)
    want = dns.scheme_coefficients(scheme, dtype)
    for gv, wv in zip(got, want):
        assert tuple(map(float, gv)) == tuple(map(float, wv))


def test_heun_weights_equal_the_adams_weights(g):
    for h in (1e-3, 2e-3, 0.05):
        b1, beta1, c2 = g._lib.heun_weights(h)
        (wb1,) = trc.adams_weights((0.0,), h)
        wbeta = trc.adams_weights((h, 0.0), h)
        assert (b1, beta1, c2) == (wb1, wbeta[0], wbeta[1] - wb1)


def _cfg(g, **kw):
    c = g._lib.HdnsConfig()
    c.grid_points, c.side_length, c.precision, c.scheme = 32, 2 * math.pi, 1, 4
    c.viscosity, c.schmidt = 0.01, 1.0
    for k, v in kw.items():
        setattr(c, k, v)
    return c


@pytest.mark.parametrize("kw,code", [
    ({"viscosity": 0.0}, 1), ({"schmidt": -1.0}, 1), ({"grid_points": 40}, 5), ({"grid_points": 2048}, 5),
    ({"scheme": 2}, 5), ({"velocity_forcing": 1, "lowest_shell": 0, "highest_shell": 2}, 1),
    ({"velocity_forcing": 1, "lowest_shell": 3, "highest_shell": 2}, 1), ({"side_length": -1.0}, 1),
])
def test_create_rejects_bad_configurations_before_touching_a_device(g, kw, code):
    L = g._lib.lib()
    h = ctypes.c_void_p()
    status = L.hdns_ctx_create(ctypes.byref(_cfg(g, **kw)), ctypes.byref(h))
    assert status == code
    assert h.value is None
    assert L.hdns_last_error(None)


def test_config_struct_layout_matches_the_header(g, tmp_path):
    """Every field's offset and the size of hdns_config as a C compiler lays out include/hdns.h."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    fields = [f for f, _ in g._lib.HdnsConfig._fields_]
    src = tmp_path / "layout.c"
    src.write_text("#include <stdio.h>\n#include <stddef.h>\n#include \"hdns.h\"\nint main(void) {\n"
                   + "".join(f'  printf("%zu\\n", offsetof(hdns_config, {f}));\n' for f in fields)
                   + '  printf("%zu\\n", sizeof(hdns_config));\n  return 0;\n}\n')
    exe = tmp_path / "layout"
    inc = Path(__file__).resolve().parent.parent / "include"
    subprocess.run([cc, "-I", str(inc), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    want = [getattr(g._lib.HdnsConfig, f).offset for f in fields] + [ctypes.sizeof(g._lib.HdnsConfig)]
    assert got == want
