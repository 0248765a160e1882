"""GPU tunes through the C ABI: stored trajectories (bit-exact), Birkhoff and
FFT tunes of windows against the oracle, and the reference notebook's printed
all_tunes values."""
import os

import numpy as np
import pytest

import oracle as O
from henon_map_cpp_amd import _capi as C
from henon_map_cpp_amd import workloads as W

pytestmark = pytest.mark.gpu
NT = min(os.cpu_count() or 1, 16)


def setup(n_turns, tx, ty, eps, x, px, y, py, mu, barrier=10.0):
    wx, wy = C.sps_modulation(tx, eps, 0, n_turns), C.sps_modulation(ty, eps, 0, n_turns)
    tab = C.build_table(wx, wy)
    tdev = C.DeviceBuffer.from_array(tab)
    dev = C.DeviceParticles(x, px, y, py)
    hist = C.track_history(dev, tdev, n_turns, n_turns, mu, barrier)
    return O.Tables(wx, wy), dev, hist


def oracle_history(tables, x, px, y, py, n_turns, mu, barrier=10.0):
    p = O.Particles(x, px, y, py)
    t = O.Tracker.from_tables(tables)
    h = [np.stack([p.x, p.px, p.y, p.py])]
    for _ in range(n_turns):
        t.track(p, 1, mu, barrier=barrier, n_threads=NT)
        h.append(np.stack([p.x, p.px, p.y, p.py]))
    return np.array(h), p


def test_history_bit_exact():
    x, px, y, py = W.ball(500, 0.45, seed=31)
    tables, dev, hist = setup(700, 0.31, 0.32, 16.0, x, px, y, py, 0.5)
    h, p = oracle_history(tables, x, px, y, py, 700, 0.5)
    got = hist.download(np.float64, 701 * 4 * 500).reshape(701, 4, 500)
    assert np.array_equal(got, h, equal_nan=True)
    assert np.array_equal(dev.get()[4], p.steps)


@pytest.mark.parametrize("mean", [False, True])
def test_window_tunes_match_oracle(mean):
    n_turns = 2100
    x, px, y, py = W.ball(300, 0.3, seed=32)
    tables, dev, hist = setup(n_turns, 0.168, 0.201, 8.0, x, px, y, py, 0.2)
    h, _ = oracle_history(tables, x, px, y, py, n_turns, 0.2)
    if mean:
        h = h - h.mean(axis=0, keepdims=True)  # replaced below by the oracle's sequential mean
        hh, _ = oracle_history(tables, x, px, y, py, n_turns, 0.2)
        s = np.zeros((4, 300))
        for m in range(n_turns + 1):
            s = s + hh[m]
        h = hh - (s / (n_turns + 1))[None]
    windows = [(0, 2047), (52, 2099), (0, 500), (1000, 2100), (7, 1030)]  # 2048, 2048, 501, 1101, 1024
    fft = C.history_tunes(hist, 300, n_turns + 1, windows, "fft", subtract_mean=mean)
    birk = C.history_tunes(hist, 300, n_turns + 1, windows, "birkhoff", subtract_mean=mean)
    for i in range(300):
        want_b = O.birkhoff_tune_vec(h[:, 0, i], h[:, 1, i], h[:, 2, i], h[:, 3, i], [w[0] for w in windows],
                                     [w[1] for w in windows])
        np.testing.assert_allclose(birk[:, :, i].ravel(), want_b[:-2], rtol=0, atol=2e-15, equal_nan=True)
        for q, (a, b) in enumerate(windows):
            if np.any(np.isnan(h[a:b + 1, :, i])):
                assert np.all(np.isnan(fft[q, :, i]))
                continue
            want = [O.fft_tune(h[a:b + 1, 0, i], h[a:b + 1, 1, i]), O.fft_tune(h[a:b + 1, 2, i], h[a:b + 1, 3, i])]
            np.testing.assert_allclose(fft[q, :, i], want, rtol=0, atol=1e-12)


def test_all_tunes_golden_on_gpu(golden):
    """tests/new_tune.ipynb: (0.1, 0, 0.1, 0), 1e4 turns, sps tunes .168/.201, mu = 0:
    the full-signal FFT tune (samples 1..n, N = 10000: direct DFT kernel), the
    FFT tune of the 10001-sample window and the Birkhoff tunes, after the
    per-coordinate mean removal of henon.cu:3041-3051."""
    e = golden["new_tune_all_tunes"]["exact"]
    x = np.array([0.1, 0.1])
    z = np.zeros(2)
    tables, dev, hist = setup(10000, 0.168, 0.201, 0.0, x, z, x, z, 0.0)
    windows = [(0, 500), (500, 1000), (0, 10000), (1, 10000)]
    fft = C.history_tunes(hist, 2, 10001, windows, "fft", subtract_mean=True)
    birk = C.history_tunes(hist, 2, 10001, windows, "birkhoff", subtract_mean=True)
    assert fft[3, 0, 0] == pytest.approx(e["fft_full_x"], abs=4e-15)
    assert fft[3, 1, 0] == pytest.approx(e["fft_full_y"], abs=4e-15)
    assert fft[2, 0, 0] == pytest.approx(e["fft_window_0_10000_x"], abs=4e-15)
    assert fft[2, 1, 0] == pytest.approx(e["fft_window_0_10000_y"], abs=4e-15)
    assert birk[0, 0, 0] == pytest.approx(e["birkhoff_window_0_500_x"], abs=1e-14)
    assert birk[1, 0, 0] == pytest.approx(e["birkhoff_window_500_1000_x"], abs=1e-14)
    assert birk[2, 0, 0] == pytest.approx(e["birkhoff_full_x"], abs=4e-15)


def test_single_tone_pow2_and_edges():
    # analytic tone z_m = exp(2 pi i f m): peak at f N, interpolation returns |1 - f|
    for N, f in [(1024, 0.2371), (2048, 0.0913), (777, 0.3333), (4096, 0.4107), (8192, 0.0517)]:
        m = np.arange(N)
        z = np.exp(2j * np.pi * f * m) * 0.01
        h = np.zeros((N, 4, 1))
        h[:, 0, 0], h[:, 1, 0] = z.real, z.imag
        h[:, 2, 0], h[:, 3, 0] = z.real, -z.imag
        d = C.DeviceBuffer.from_array(np.ascontiguousarray(h))
        got = C.history_tunes(d, 1, N, [(0, N - 1)], "fft")
        want = [O.fft_tune(h[:, 0, 0], h[:, 1, 0]), O.fft_tune(h[:, 2, 0], h[:, 3, 0])]
        np.testing.assert_allclose(got[0, :, 0], want, atol=1e-12)
        assert got[0, 0, 0] == pytest.approx(1 - f, abs=1e-6)


def test_fft_tunes_in_particle_batches(monkeypatch):
    """The power-of-two path corner-turns signals for batches of particles; a
    small staging budget forces several batches (the first and last particle of
    each batch must land in the right output slot)."""
    N, n = 256, 1000
    rng = np.random.default_rng(9)
    f = rng.uniform(0.02, 0.48, n)
    m = np.arange(N)[:, None]
    zx = 0.01 * np.exp(2j * np.pi * f[None, :] * m)
    zy = 0.02 * np.exp(-2j * np.pi * f[None, :] * m)
    h = np.zeros((N, 4, n))
    h[:, 0], h[:, 1], h[:, 2], h[:, 3] = zx.real, zx.imag, zy.real, zy.imag
    h[5, 2, 17] = np.nan  # one signal with a NaN sample
    d = C.DeviceBuffer.from_array(np.ascontiguousarray(h))
    monkeypatch.setenv("HM_FFT_ROW_BYTES", str(64 * 4 * N * 8 * 3))  # 192 particles per batch
    got = C.history_tunes(d, n, N, [(0, N - 1)], "fft")
    monkeypatch.delenv("HM_FFT_ROW_BYTES")
    ref = C.history_tunes(d, n, N, [(0, N - 1)], "fft")
    assert np.array_equal(got, ref, equal_nan=True)
    for i in [0, 1, 191, 192, 193, 383, 384, 999, 17]:
        want = [O.fft_tune(h[:, 0, i], h[:, 1, i]), O.fft_tune(h[:, 2, i], h[:, 3, i])]
        if i == 17:
            assert np.isnan(got[0, 0, i]) and np.isnan(got[0, 1, i])
        else:
            np.testing.assert_allclose(got[0, :, i], want, atol=1e-12)


def test_rows_history_and_tunes_equal_sample_major():
    """The particle-row trajectory store (hm_track_rows) holds the sample-major
    history transposed, bit for bit (padding = the final state), and every tune
    kernel gives the same result from either layout: the in-place row FFT
    (2048), the one-signal FFT (8192), the DFT (501), Birkhoff windows, with and
    without the mean removal."""
    n_turns, n = 8300, 257
    x, px, y, py = W.ball(n, 0.5, seed=34)  # 3 particles lost before sample 2047, 6 after 2147
    wx, wy = C.sps_modulation(0.168, 8.0, 0, n_turns), C.sps_modulation(0.201, 8.0, 0, n_turns)
    tdev = C.DeviceBuffer.from_array(C.build_table(wx, wy))
    d1, d2 = C.DeviceParticles(x, px, y, py), C.DeviceParticles(x, px, y, py)
    hist = C.track_history(d1, tdev, n_turns, n_turns, 0.3, 10.0)
    rows, fused_means = C.track_rows(d2, tdev, n_turns, n_turns, 0.3, 10.0, means=True)
    L = C.row_length(n_turns)
    # the means summed while the rows are written equal a separate pass bit for bit
    sep = C.DeviceBuffer(8 * 4 * n)
    C.check(C.lib().hm_rows_means(rows.ptr, n, L, n_turns + 1, sep.ptr, None))
    fm = fused_means.download(np.float64, 4 * n)
    assert np.array_equal(fm, sep.download(np.float64, 4 * n), equal_nan=True)
    assert np.isnan(fm).any() and np.isfinite(fm).any()
    assert L % 16 == 0 and L >= n_turns + 1
    h = hist.download(np.float64, (n_turns + 1) * 4 * n).reshape(n_turns + 1, 4, n)
    r = rows.download(np.float64, n * 4 * L).reshape(n, 4, L)
    assert np.array_equal(r[:, :, :n_turns + 1], np.transpose(h, (2, 1, 0)), equal_nan=True)
    assert np.array_equal(r[:, :, n_turns + 1:], np.repeat(r[:, :, n_turns:n_turns + 1], L - n_turns - 1, axis=2),
                          equal_nan=True)
    assert np.array_equal(d1.get()[4], d2.get()[4])
    windows = [(0, 2047), (100, 2147), (0, 8191), (3, 503), (0, n_turns)]
    for mean in (False, True):
        a = C.history_tunes(hist, n, n_turns + 1, windows, "fft", subtract_mean=mean)
        b = C.history_tunes(rows, n, n_turns + 1, windows, "fft", subtract_mean=mean, row_len=L)
        assert np.array_equal(a, b, equal_nan=True), mean
        if mean:
            c = C.history_tunes(rows, n, n_turns + 1, windows, "fft", row_len=L, means=fused_means)
            assert np.array_equal(a, c, equal_nan=True)
        # NaN exactly where a sample of the window (any coordinate) is NaN, and
        # with the mean removal also where the mean over all samples is NaN:
        # particles lost inside a window and particles lost only after it occur
        for w, (f, t) in enumerate(windows[:2]):
            bad = np.isnan(r[:, :, f:t + 1]).any(axis=(1, 2))
            later = ~bad & np.isnan(r[:, :, :n_turns + 1]).any(axis=(1, 2))
            assert bad.any() and later.any()
            for plane in (0, 1):
                assert np.array_equal(np.isnan(b[w, plane]), bad | later if mean else bad), (mean, w, plane)
        # Birkhoff from rows sums the same terms in a wavefront-parallel order:
        # sums of up to 8300 terms of size ~1e-3 agree to rounding (1e-14)
        a = C.history_tunes(hist, n, n_turns + 1, windows, "birkhoff", subtract_mean=mean)
        b = C.history_tunes(rows, n, n_turns + 1, windows, "birkhoff", subtract_mean=mean, row_len=L)
        assert np.array_equal(np.isnan(a), np.isnan(b))
        np.testing.assert_allclose(b, a, rtol=0, atol=2e-14, equal_nan=True)


@pytest.mark.parametrize("layout", ["rows", "sample_major"])
@pytest.mark.parametrize("kind", ["fft", "birkhoff"])
def test_tunes_nan_rule_on_given_samples(kind, layout):
    """A NaN in any of the four coordinates inside a window makes both planes'
    tunes NaN, also when only the other plane holds it (samples written
    directly, not by the tracker, which sets all four together); with the mean
    removal, a NaN anywhere in a coordinate makes its mean, and so every
    mean-removed sample and both tunes, NaN (dynamic_indicator.cpp:228-241
    checks the mean-removed vectors). Finite samples give finite tunes either
    way, for the FFT and the Birkhoff tunes and both history layouts."""
    L, n_samples = 2064, 2049
    t = np.arange(L)
    base = np.stack([np.cos(2 * np.pi * 0.21 * t), -np.sin(2 * np.pi * 0.21 * t),
                     np.cos(2 * np.pi * 0.17 * t), -np.sin(2 * np.pi * 0.17 * t)])
    rows = np.repeat(base[None], 4, axis=0)       # [particle][coord][L]
    rows[1, 3, 10] = np.nan                        # py inside the window
    rows[2, 2, 2048:] = np.nan                     # y after the window only (sample 2048 counts in the mean)
    rows[3, 0, 2047] = np.nan                      # x at the window's last sample
    if layout == "rows":
        dev, row_len = C.DeviceBuffer.from_array(np.ascontiguousarray(rows)), L
    else:                                          # [sample][coord][particle]
        hist = np.ascontiguousarray(rows[:, :, :n_samples].transpose(2, 1, 0))
        dev, row_len = C.DeviceBuffer.from_array(hist), None
    for mean in (False, True):
        out = C.history_tunes(dev, 4, n_samples, [(0, 2047)], kind, subtract_mean=mean, row_len=row_len)[0]
        want_nan = np.array([False, True, mean, True])
        for plane in (0, 1):
            assert np.array_equal(np.isnan(out[plane]), want_nan), (mean, plane)
        assert out[0, 0] == pytest.approx(0.21, abs=1e-5) and out[1, 0] == pytest.approx(0.17, abs=1e-5)
