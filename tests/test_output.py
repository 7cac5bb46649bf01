"""flang's list-directed editing of REAL(8) values, as the Tecplot writers
apply it (fvk_fmt_list_real, host only), against what a flang-built probe
printed for 14,575 values covering random bit patterns, short decimals,
powers of ten and the edges of the F/E choice (tests/golden/flang_list_real.npz,
made by tests/golden/make_list_real.py)."""
import os

import numpy as np

from tests.helpers import ROOT, require_lib


def test_list_directed_real_matches_flang():
    fvk = require_lib()
    z = np.load(os.path.join(ROOT, "tests", "golden", "flang_list_real.npz"))
    vals = z["bits"].view(np.float64)
    bad = [(float(v), t, fvk.fmt_list_real(v)) for v, t in zip(vals, z["text"]) if fvk.fmt_list_real(v) != t]
    assert not bad, bad[:10]
    assert len(vals) > 14000


def test_known_forms():
    fvk = require_lib()
    cases = [(0.0, "0."), (-0.0, "-0."), (1.0, "1."), (0.1, ".1"), (12.5, "12.5"), (100.0, "100."),
             (0.03125, "3.125E-02"), (1e16, "1.E+16"), (1e-300, "1.E-300"), (0.095, ".095"),
             (0.0949, "9.49E-02"), (1e14, "100000000000000."), (1e15, "1.E+15"),
             (1.0 / 3.0, ".3333333333333333"), (1e23, "9.999999999999999E+22"),
             (2.2719716328605813e-222, "2.2719716328605812E-222")]
    for v, t in cases:
        assert fvk.fmt_list_real(v) == t, (v, fvk.fmt_list_real(v), t)
