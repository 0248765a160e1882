"""GPU parity of the tangent-map kernels through the C ABI: the reference's
trace-form MEGNO, the fused indicator kernel (exact mode: state and tangent
matrix bit-exact, logarithm-based outputs within 1e-13), fast mode, exact-orbit
mode, and the per-call construct kernels."""
import os

import numpy as np
import pytest

import oracle as O
from henon_map_cpp_amd import _capi as C
from henon_map_cpp_amd import workloads as W

pytestmark = pytest.mark.gpu
NT = min(os.cpu_count() or 1, 16)


def tables(n, tx, ty, eps, kind="sps"):
    if kind == "custom":
        coef, freq = W.seeded_line_coefficients(), W.line_frequencies()
        wx = C.lines_modulation(tx, eps, coef, freq, 0, n)
        wy = C.lines_modulation(ty, eps, coef, freq, 0, n)
    else:
        wx, wy = C.sps_modulation(tx, eps, 0, n), C.sps_modulation(ty, eps, 0, n)
    tab = C.build_table(wx, wy)
    return O.Tables(wx, wy), tab, C.DeviceBuffer.from_array(tab)


@pytest.mark.parametrize("mu", [0.0, 0.5])
def test_megno_trace_matches_reference_cpu_megno(mu):
    ot, tab, tdev = tables(600, 0.01, 0.012, 4.0)
    x, px, y, py = W.ball(700, 0.3, seed=21)
    samples = [0, 5, 100, 599, 600]
    dev = C.DeviceParticles(x, px, y, py)
    got = C.megno_trace(dev, tdev, 600, 600, mu, 10.0, samples)
    p = O.Particles(x, px, y, py)
    want = O.Tracker.from_tables(ot).megno(p, 600, mu, barrier=10.0, turn_samples=samples, n_threads=NT,
                                           carry_off_diagonal=False)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14, equal_nan=True)
    assert np.all(got[:, -1] == 0.0)  # sample n_turns reads the never-written slot
    gx, gpx, gy, gpy, gs = dev.get()
    assert np.array_equal(gx, p.x, equal_nan=True) and np.array_equal(gs, p.steps)


@pytest.mark.parametrize("mu,renorm", [(0.5, 0), (0.5, 100), (0.0, 50)])
def test_indicators_exact_mode(mu, renorm):
    ot, tab, tdev = tables(3000, 0.168, 0.201, 32.0, "custom")
    x, px, y, py = W.ball(2000, 0.5, seed=22)
    samples = [1, 10, 1000, 3000]
    birk = [(0, 3000), (100, 2148)]
    dev = C.DeviceParticles(x, px, y, py)
    got = C.track_indicators(dev, tdev, 3000, 3000, mu, 10.0, C.MODE_EXACT, renorm, samples, (1, 1, 0, 0),
                             matrix=True, birkhoff=birk)
    p = O.Particles(x, px, y, py)
    want = O.indicators(p, ot, 3000, mu, 10.0, renorm, samples, (1, 1, 0, 0), birkhoff=birk, n_threads=NT)
    gx, gpx, gy, gpy, gs = dev.get()
    # the map and the tangent matrix carry no transcendental: bit for bit
    assert np.array_equal(gx, p.x, equal_nan=True) and np.array_equal(gs, p.steps)
    if renorm == 0:
        assert np.array_equal(got["matrix"], want["matrix"], equal_nan=True)
    else:  # the scale factor goes through log10 only in log_scale; M itself is exact
        assert np.array_equal(got["matrix"], want["matrix"], equal_nan=True)
        np.testing.assert_allclose(got["log_scale"], want["log_scale"], rtol=1e-13, equal_nan=True)
    for key in ("megno", "fli"):
        np.testing.assert_allclose(got[key], want[key], rtol=1e-12, atol=1e-13, equal_nan=True)
    np.testing.assert_allclose(got["birkhoff"], want["birkhoff"], rtol=0, atol=1e-13, equal_nan=True)


def test_indicators_fast_mode_regular_orbits():
    ot, tab, tdev = tables(2000, 0.168, 0.201, 32.0, "custom")
    x, px, y, py = W.ball(1024, 0.15, seed=23)
    samples = [100, 2000]
    a = C.track_indicators(C.DeviceParticles(x, px, y, py), tdev, 2000, 2000, 0.5, 10.0, C.MODE_EXACT, 100,
                           samples, birkhoff=[(0, 2000)])
    b = C.track_indicators(C.DeviceParticles(x, px, y, py), tdev, 2000, 2000, 0.5, 10.0, C.MODE_FAST, 100,
                           samples, birkhoff=[(0, 2000)])
    np.testing.assert_allclose(b["megno"], a["megno"], rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(b["fli"], a["fli"], rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(b["birkhoff"], a["birkhoff"], atol=1e-11)


@pytest.mark.parametrize("mu", [0.0, 0.5])
def test_indicators_exact_orbit_mode(mu):
    """Exact orbit: the state, losses, Birkhoff tunes and trajectory window of
    exact mode bit for bit (chaotic particles and losses included); MEGNO and
    FLI from fused tangent products, against the oracle's unfused ones."""
    ot, tab, tdev = tables(3000, 0.168, 0.201, 32.0, "custom")
    x, px, y, py = W.ball(2000, 0.5, seed=26)
    samples = [1, 10, 1000, 3000]
    birk = [(0, 3000), (100, 2148)]
    de, do = C.DeviceParticles(x, px, y, py), C.DeviceParticles(x, px, y, py)
    a = C.track_indicators(de, tdev, 3000, 3000, mu, 10.0, C.MODE_EXACT, 100, samples, (1, 1, 0, 0),
                           matrix=True, birkhoff=birk)
    b = C.track_indicators(do, tdev, 3000, 3000, mu, 10.0, C.MODE_EXACT_ORBIT, 100, samples, (1, 1, 0, 0),
                           matrix=True, birkhoff=birk)
    for u, v in zip(de.get(), do.get()):
        assert np.array_equal(u, v, equal_nan=True)
    assert np.array_equal(a["birkhoff"], b["birkhoff"], equal_nan=True)
    lost = np.isnan(de.get()[0])
    assert lost.any() and (~lost).any()
    assert np.array_equal(np.isnan(a["megno"]), np.isnan(b["megno"]))
    assert not np.array_equal(a["matrix"], b["matrix"], equal_nan=True)  # the products are fused
    # regular orbits: fused and unfused products agree closely; chaotic ones
    # amplify the rounding difference like any other perturbation of M
    want = O.indicators(O.Particles(x, px, y, py), ot, 3000, mu, 10.0, 100, samples, (1, 1, 0, 0),
                        birkhoff=birk, n_threads=NT)
    regular = np.all(np.isfinite(want["megno"]), axis=0) & (np.abs(want["megno"][-1]) < 2.5)
    assert regular.sum() > 100
    for key in ("megno", "fli"):
        np.testing.assert_allclose(b[key][:, regular], want[key][:, regular], rtol=1e-9, atol=1e-10)


@pytest.mark.parametrize("renorm", [0, 100])
def test_birkhoff_window_variant_equals_turn_by_turn(renorm):
    """Birkhoff windows alone run the segmented kernel with the windows summed
    inside its straight runs of plain turns (losses inside a run, windows that
    end inside a run, a NaN start); adding a construct window selects the
    turn-by-turn kernel. Every output agrees bit for bit."""
    ot, tab, tdev = tables(3000, 0.168, 0.201, 32.0, "custom")
    x, px, y, py = W.ball(3000, 0.55, seed=27)
    x[5] = np.nan
    samples = [10, 1000, 3000]
    birk = [(0, 3000), (100, 2148), (1500, 1537), (7, 400)]
    a = C.track_indicators(C.DeviceParticles(x, px, y, py), tdev, 3000, 3000, 0.5, 10.0, C.MODE_EXACT, renorm,
                           samples, (1, 1, 0, 0), matrix=True, birkhoff=birk)
    dv = C.DeviceParticles(x, px, y, py)
    b = C.track_indicators(dv, tdev, 3000, 3000, 0.5, 10.0, C.MODE_EXACT, renorm, samples, (1, 1, 0, 0),
                           matrix=True, birkhoff=birk, tune_birkhoff=[50])
    for key in ("megno", "fli", "matrix", "birkhoff"):
        assert np.array_equal(a[key], b[key], equal_nan=True), key
    lost = np.isnan(dv.get()[0])
    steps = dv.get()[4]
    # particles lost inside each window's range, and windows finished before the loss
    assert np.any(lost & (steps >= 1500) & (steps < 1537))
    assert np.isnan(a["birkhoff"][:, :, 5]).all()
    ok = ~lost
    assert np.isfinite(a["birkhoff"][:, :, ok]).all()
    early = lost & (steps > 500)
    assert np.isfinite(a["birkhoff"][3, :, early]).all() and np.isnan(a["birkhoff"][0, :, early]).all()


def test_exact_orbit_mode_is_indicators_only():
    tab = C.DeviceBuffer.from_array(C.build_table(C.sps_modulation(0.168, 0.0, 0, 10),
                                                  C.sps_modulation(0.201, 0.0, 0, 10)))
    p = C.DeviceParticles(*W.ball(8, 0.1, seed=1))
    with pytest.raises(ValueError, match="unknown mode"):
        C.track(p, tab, 10, 10, 0.0, 10.0, mode=C.MODE_EXACT_ORBIT)


def test_history_window():
    ot, tab, tdev = tables(500, 0.31, 0.32, 16.0)
    x, px, y, py = W.ball(333, 0.4, seed=24)
    got = C.track_indicators(C.DeviceParticles(x, px, y, py), tdev, 500, 500, 0.5, 10.0, hist=(0, 501))
    p = O.Particles(x, px, y, py)
    t = O.Tracker.from_tables(ot)
    h = [np.stack([p.x, p.px, p.y, p.py])]
    for _ in range(500):
        t.track(p, 1, 0.5, barrier=10.0)
        h.append(np.stack([p.x, p.px, p.y, p.py]))
    assert np.array_equal(got["hist"], np.array(h), equal_nan=True)


def test_construct_kernels_match_reference_formulas():
    ot, tab, tdev = tables(50, 0.168, 0.201, 0.0)
    n = 300
    x, px, y, py = W.ball(n, 0.3, seed=25)
    dev = C.DeviceParticles(x, px, y, py)
    mats = C.DeviceBuffer.from_array(np.tile(np.eye(4), (n, 1, 1)))
    C.check(C.lib().hm_tangent_apply(mats.ptr, *dev.ptrs()[:4], n, tdev.ptr, 50, 3, 0.2, 1, None))
    C.check(C.lib().hm_tangent_apply(mats.ptr, *dev.ptrs()[:4], n, tdev.ptr, 50, 4, 0.2, 1, None))
    got = mats.download(np.float64, n * 16).reshape(n, 16)
    want = O.structured_multiply(np.tile(np.eye(4), (n, 1, 1)).reshape(n, 16), x, px, y, py, ot, 3, 0.2)
    want = O.structured_multiply(want, x, px, y, py, ot, 4, 0.2)
    assert np.array_equal(got, want)
    # matvec (in place, reference order) and normalize
    v = np.random.default_rng(1).normal(size=(n, 4))
    vd = C.DeviceBuffer.from_array(v)
    C.check(C.lib().hm_matvec(mats.ptr, vd.ptr, vd.ptr, n, 1, None))
    assert np.array_equal(vd.download(np.float64, 4 * n).reshape(n, 4), O.matvec(want, v))
    C.check(C.lib().hm_normalize(vd.ptr, n, None))
    np.testing.assert_allclose(vd.download(np.float64, 4 * n).reshape(n, 4), O.normalize(O.matvec(want, v)),
                               rtol=2e-16, atol=1e-300)
    # megno construct: add(M2, M1) with idx 1
    l = [C.DeviceBuffer.from_array(np.zeros(n)) for _ in range(3)]
    m1 = C.DeviceBuffer.from_array(np.tile(np.eye(4), (n, 1, 1)))
    C.check(C.lib().hm_megno_add(l[0].ptr, l[1].ptr, l[2].ptr, mats.ptr, m1.ptr, n, 1, None))
    r = [np.zeros(n) for _ in range(3)]
    O.megno_construct_add(r[0], r[1], r[2], want, np.tile(np.eye(4), (n, 1, 1)).reshape(n, 16), 1)
    np.testing.assert_allclose(l[2].download(np.float64, n), r[2], rtol=1e-14)


def test_full_indicator_workload_sampled():
    """The full chaos-indicator workload (256x256 grid + 65536 ball points, 1e5
    turns, mu=0.5, seeded 7-line modulation at eps=32, renormalisation every 100
    turns, samples 1e3, 1e4, 1e5) in exact mode; every 257th particle (511,
    grid and ball) against the oracle: state, survival turns and the tangent
    matrix bit for bit, the log10 scale to 1e-13, MEGNO and FLI to 1e-11 (sums
    of 1e5 logarithms, whose last bits differ between the device and glibc)."""
    c = W.CHAOS_INDICATORS
    T = c["n_turns"]
    x, px, y, py = W.indicator_points()
    ot, tab, tdev = tables(T, c["tune_x"], c["tune_y"], c["epsilon"], "custom")
    dev = C.DeviceParticles(x, px, y, py)
    got = C.track_indicators(dev, tdev, T, T, c["mu"], c["barrier"], C.MODE_EXACT, c["renorm_every"], c["samples"],
                             matrix=True)
    gx, gpx, gy, gpy, gs = dev.get()
    pick = np.arange(0, x.size, 257)
    p = O.Particles(x[pick], px[pick], y[pick], py[pick])
    want = O.indicators(p, ot, T, c["mu"], c["barrier"], c["renorm_every"], c["samples"], n_threads=NT)
    assert np.array_equal(gx[pick], p.x, equal_nan=True) and np.array_equal(gpy[pick], p.py, equal_nan=True)
    assert np.array_equal(gs[pick], p.steps)
    assert np.array_equal(got["matrix"][pick], want["matrix"], equal_nan=True)
    np.testing.assert_allclose(got["log_scale"][pick], want["log_scale"], rtol=1e-13, equal_nan=True)
    for key in ("megno", "fli"):
        np.testing.assert_allclose(got[key][:, pick], want[key], rtol=1e-11, atol=1e-12, equal_nan=True)
    lost = np.isnan(p.x)
    assert lost.any() and (~lost).any()
