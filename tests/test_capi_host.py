"""C ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and its host-side functions (modulation tables, table layout,
Birkhoff weights, argument errors) match the oracle and the reference source."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O
from henon_map_cpp_amd import _capi as C

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "henon_hip.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(hm_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(C.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the ctypes binding covers all of them
    assert set(names) <= set(C.SIGNATURES), set(names) - set(C.SIGNATURES)


def test_version_and_errors():
    assert b"gfx950" in C.lib().hm_version()
    out = np.empty(8)
    rc = C.lib().hm_sps_modulation(0.1, 0.0, 10, 5, C.ptr(out))
    assert rc == C.HM_ERR_ARG
    assert "start" in C.lib().hm_last_error().decode()
    with pytest.raises(ValueError):
        C.check(rc)
    # hm_track refuses a finite kick before any device work (hm_track_kicked
    # takes it, with the generator's seed and draw index)
    with pytest.raises(ValueError, match="hm_track_kicked"):
        C.check(C.lib().hm_track(None, None, None, None, None, 4, None, 4, 10, 0.0, 1.0, 0.5, 0, 0, None, None))
    with pytest.raises(ValueError, match="finite"):
        C.check(C.lib().hm_track_kicked(None, None, None, None, None, 4, None, 4, 10, 0.0, 1.0, float("nan"), 0, 0,
                                        1, 0, None))


@pytest.mark.parametrize("tune,eps,start,end", [(0.168, 0.0, 0, 777), (0.31, 32.0, 0, 70000), (0.2, 16.0, 99, 1500)])
def test_sps_modulation_bit_exact(tune, eps, start, end):
    got = C.sps_modulation(tune, eps, start, end)
    assert np.array_equal(got, O.sps_modulation(tune, eps, start, end))
    R = O.ref_lib()
    if R is not None:
        r = np.empty(end - start)
        R.ref_sps_modulation(tune, eps, start, end, r.ctypes.data_as(ctypes.c_void_p))
        assert np.array_equal(got, r)


def test_basic_and_lines_modulation_bit_exact():
    assert np.array_equal(C.basic_modulation(0.3, 0.01, 5.0, 3, 900), O.basic_modulation(0.3, 0.01, 5.0, 3, 900))
    c, f = C.sps_lines()
    oc, of = O.sps_lines()
    assert np.array_equal(c, oc) and np.array_equal(f, of)
    rng = np.random.default_rng(0)
    a = rng.uniform(0.1, 1.0, 7)
    coef = a / a.sum() * 1e-4
    freq = np.array([1, 2, 3, 6, 7, 10, 12]) * (2 * np.pi / 868.12)
    assert np.array_equal(C.lines_modulation(0.168, 32.0, coef, freq, 0, 5000),
                          O.lines_modulation(0.168, 32.0, coef, freq, 0, 5000))
    assert np.array_equal(C.lines_modulation(0.168, 32.0, c, f, 0, 500), C.sps_modulation(0.168, 32.0, 0, 500))


def test_table_layout_and_values():
    wx = C.sps_modulation(0.168, 32.0, 0, 3000)
    wy = C.sps_modulation(0.201, 32.0, 0, 3000)
    tab = C.build_table(wx, wy)
    t = O.Tables(wx, wy)
    assert tab.shape == (3000, 4)
    assert np.array_equal(tab, t.interleaved())


def test_birkhoff_weights_bit_exact():
    for n in (1, 2, 10, 2047, 100000):
        assert np.array_equal(C.birkhoff_weights(n), O.birkhoff_weights(n), equal_nan=True)


def test_lines_modulation_derivatives():
    # d omega_k/d omega_0 and d omega_k/d eps of the line modulation against
    # central differences of the modulation itself (which is linear in eps and
    # in tune, so the differences are exact up to rounding)
    from henon_map_cpp_amd.adjoint import lines_derivatives
    c, f = C.sps_lines()
    tune, eps, start, end = 0.201, 12.0, 7, 2007
    d = lines_derivatives(tune, eps, c, f, start, end)
    h = 1.0  # omega is linear in eps and in tune: wide steps only cut rounding
    de = (C.lines_modulation(tune, eps + h, c, f, start, end) - C.lines_modulation(tune, eps - h, c, f, start, end))
    np.testing.assert_allclose(d[:, 1], de / (2 * h), rtol=0, atol=1e-12 * np.abs(d[:, 1]).max())
    h = 0.01
    dq = (C.lines_modulation(tune + h, eps, c, f, start, end) - C.lines_modulation(tune - h, eps, c, f, start, end))
    np.testing.assert_allclose(d[:, 0], dq / (2 * h) / (2 * np.pi), rtol=1e-12)
    # without modulation omega_k = omega_0
    d0 = lines_derivatives(tune, 0.0, c, f, 0, 10)
    assert np.all(d0[:, 0] == 1.0)
    with pytest.raises(ValueError):
        lines_derivatives(tune, eps, c, f, 5, 5)
