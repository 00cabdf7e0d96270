"""CPU checks of the C-ABI library: it loads, exports every entry point include/*.h declares, and its
host-only functions (descriptor defaults, the seeded generator, errors) behave without a GPU."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

from purify_amd import _lib, utilities

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(pf_[A-Za-z0-9_]+)\s*\(", txt):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the ctypes table covers every declaration
    assert set(names) <= set(_lib.SIGNATURES)


def test_desc_defaults_match_reference():
    d = _lib.pf_op_desc()
    _lib.lib().pf_op_desc_init(C.byref(d))
    # operators.h:373-380 defaults
    assert d.oversample_ratio == 2.0 and d.power_iters == 100 and d.power_tol == 1e-4
    assert d.Ju == 4 and d.Jv == 4 and d.Jw == 6 and d.w_term == 0 and d.kernel == 0
    assert d.cellx == 1.0 and d.celly == 1.0


def test_version_string():
    assert b"gfx950" in _lib.lib().pf_version()


def test_synth_uv_is_seeded_and_gaussian():
    u1, v1 = utilities.synth_uv(1001, 200_000, np.pi / 3)
    u2, v2 = utilities.synth_uv(1001, 200_000, np.pi / 3)
    assert np.array_equal(u1, u2) and np.array_equal(v1, v2)
    u3, _ = utilities.synth_uv(1002, 1000, np.pi / 3)
    assert not np.array_equal(u1[:1000], u3)
    assert abs(np.mean(u1)) < 0.01
    assert np.std(u1) == pytest.approx(np.pi / 3, rel=0.01)
    assert np.max(np.abs(u1)) <= np.pi


def test_synth_image_sources():
    img = utilities.synth_image(1000, 64, 48, 20)
    assert img.shape == (64 * 48,)
    assert np.all(img.imag == 0) and np.all(img.real >= 0)
    assert img.real.max() > 0.1
    assert np.array_equal(img, utilities.synth_image(1000, 64, 48, 20))


def test_synth_cnormal_variance():
    z = utilities.synth_cnormal(7, 400_000, 2.0)
    assert np.mean(np.abs(z) ** 2) == pytest.approx(4.0, rel=0.01)


def test_errors_reported_without_gpu():
    L = _lib.lib()
    assert L.pf_synth_uv(1, -1, 1.0, None, None) == _lib.PF_ERR_ARG
    assert b"bad arguments" in L.pf_last_error()
    h = C.c_void_p()
    assert L.pf_op_create(None, C.byref(h)) == _lib.PF_ERR_ARG


def test_l2_radius_helpers_match_oracle():
    import oracle
    for M in (10, 1000, 13107):
        assert utilities.calculate_l2_radius(M, 0.5) == pytest.approx(oracle.calculate_l2_radius(M, 0.5), rel=1e-15)
    y = utilities.synth_cnormal(3, 1000)
    assert utilities.SNR_to_standard_deviation(y, 30) == pytest.approx(oracle.snr_to_standard_deviation(y, 30), rel=1e-14)
