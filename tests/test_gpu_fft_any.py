"""FFT tunes of windows of any length (the whole-signal kernel for windows up
to about 10^4 points, the four-step mixed-radix transform, chirp-z over smooth
and power-of-two lengths) through the C ABI and the drop-in API: against the oracle's long-double DFT where that finishes in
seconds (N up to 10001), against numpy's FFT magnitudes with the oracle's peak
interpolation above that, against single tones, with NaN samples, in particle
batches, and all_tunes over 1e5 turns (windows of 100000 and 100001 samples)."""
import numpy as np
import pytest

import oracle as O
from henon_map_cpp_amd import _capi as C
from henon_map_cpp_amd import workloads as W

pytestmark = pytest.mark.gpu


def quasi_periodic(N, n, seed):
    """n particles of a two-plane quasi-periodic signal with harmonics and a
    little noise: sample-major history [N][4][n]."""
    rng = np.random.default_rng(seed)
    m = np.arange(N)[:, None]
    h = np.zeros((N, 4, n))
    for plane in (0, 1):
        f = rng.uniform(0.05, 0.45, n)[None, :]
        a2 = rng.uniform(0.05, 0.3, n)[None, :]
        z = np.exp(-2j * np.pi * f * m) + a2 * np.exp(-2j * np.pi * 2 * f * m) + 0.3 * np.exp(2j * np.pi * 0.5 * f * m)
        z = 0.1 * z + 1e-4 * (rng.normal(size=(N, n)) + 1j * rng.normal(size=(N, n)))
        h[:, 2 * plane], h[:, 2 * plane + 1] = z.real, z.imag
    return h


def numpy_fft_tune(x, px):
    """dynamic_indicator.cpp:178-196 with numpy's FFT for the magnitudes and the
    oracle's interpolation (dynamic_indicator.cpp:139-176)."""
    N = len(x)
    h = 2 * np.sin(np.pi * np.arange(N) / N) ** 2
    mag = np.abs(np.fft.fft((x + 1j * px) * h))
    return O.interpolation(mag)


@pytest.mark.parametrize("N", [2, 3, 4, 5, 31, 501, 777, 1101, 2049, 4095, 4096, 4097, 8192, 10000, 10001])
def test_chirpz_matches_oracle_dft(N):
    n = 3 if N > 4000 else 8
    h = quasi_periodic(N, n, seed=N)
    d = C.DeviceBuffer.from_array(np.ascontiguousarray(h))
    got = C.history_tunes(d, n, N, [(0, N - 1)], "fft")
    for i in range(n):
        want = [O.fft_tune(h[:, 0, i], h[:, 1, i]), O.fft_tune(h[:, 2, i], h[:, 3, i])]
        np.testing.assert_allclose(got[0, :, i], want, rtol=0, atol=1e-12, err_msg=f"N={N} particle {i}")


@pytest.mark.parametrize("N", [6, 12, 45, 49, 210, 343, 1000, 1500, 2187, 3125, 5040, 9604])
def test_mixed_radix_lengths_match_oracle_dft(N):
    """Lengths with only the factors 2, 3, 5, 7 take the mixed-radix
    transform (radices 8, 4, 2, 3, 5, 7 in every combination here)."""
    n = 3 if N > 4000 else 8
    h = quasi_periodic(N, n, seed=N + 1)
    d = C.DeviceBuffer.from_array(np.ascontiguousarray(h))
    got = C.history_tunes(d, n, N, [(0, N - 1)], "fft")
    for i in range(n):
        want = [O.fft_tune(h[:, 0, i], h[:, 1, i]), O.fft_tune(h[:, 2, i], h[:, 3, i])]
        np.testing.assert_allclose(got[0, :, i], want, rtol=0, atol=1e-12, err_msg=f"N={N} particle {i}")


@pytest.mark.parametrize("N", [254, 10023, 10176])
def test_whole_signal_prime_passes_match_oracle_dft(N):
    """One-kernel transforms whose plans hold prime passes: 2 x 127 (256-thread
    kernel), 3 x 13 x 257 (the largest prime pass, no chirp-z length fits),
    2^6 x 3 x 53 (about the whole 160 KiB of LDS); a NaN sample makes its
    particle's tunes NaN."""
    n = 3
    h = quasi_periodic(N, n, seed=N + 7)
    h[N // 2, 1, 1] = np.nan
    d = C.DeviceBuffer.from_array(np.ascontiguousarray(h))
    got = C.history_tunes(d, n, N, [(0, N - 1)], "fft")
    assert np.all(np.isnan(got[0, :, 1]))
    for i in (0, 2):
        want = [O.fft_tune(h[:, 0, i], h[:, 1, i]), O.fft_tune(h[:, 2, i], h[:, 3, i])]
        np.testing.assert_allclose(got[0, :, i], want, rtol=0, atol=1e-12, err_msg=f"N={N} particle {i}")


@pytest.mark.parametrize("N", [16384, 16411, 17640, 65537, 100000, 100001, 131072, 2_100_001])
def test_chirpz_long_windows_match_numpy_fft(N):
    """Windows too long for one workgroup: {2,3,5,7}-smooth lengths (17640 =
    2^3 3^2 5 7^2: every radix) as four-step transforms of the length, others up
    to 2^21 samples by chirp-z over a smooth length; 2100001 samples (no smooth
    length of two 2048-point factors is long enough) take the power-of-two
    chirp-z path."""
    n = 4 if N < 1_000_000 else 2
    h = quasi_periodic(N, n, seed=N)
    d = C.DeviceBuffer.from_array(np.ascontiguousarray(h))
    got = C.history_tunes(d, n, N, [(0, N - 1)], "fft")
    for i in range(n):
        want = [numpy_fft_tune(h[:, 0, i], h[:, 1, i]), numpy_fft_tune(h[:, 2, i], h[:, 3, i])]
        np.testing.assert_allclose(got[0, :, i], want, rtol=0, atol=1e-12, err_msg=f"N={N} particle {i}")


@pytest.mark.parametrize("N,f", [(10000, 0.1681), (10001, 0.3127), (100001, 0.2011), (12345, 0.0412)])
def test_chirpz_single_tone(N, f):
    m = np.arange(N)
    z = np.exp(2j * np.pi * f * m) * 0.01
    h = np.zeros((N, 4, 1))
    h[:, 0, 0], h[:, 1, 0] = z.real, z.imag
    h[:, 2, 0], h[:, 3, 0] = z.real, -z.imag
    d = C.DeviceBuffer.from_array(np.ascontiguousarray(h))
    got = C.history_tunes(d, 1, N, [(0, N - 1)], "fft")
    assert got[0, 0, 0] == pytest.approx(1 - f, abs=1e-6)
    assert got[0, 1, 0] == pytest.approx(f, abs=1e-6)


@pytest.mark.parametrize("layout", ["rows", "sample_major"])
def test_chirpz_windows_nan_and_batches(layout, monkeypatch):
    """Windows of 1500 and 3001 samples at offsets, NaN samples inside and
    outside them, both history layouts, the mean removal, and a scratch budget
    small enough to force several particle batches: the same results as one
    batch, and the oracle's per window."""
    N_all, n = 3200, 11
    h = quasi_periodic(N_all, n, seed=5)
    h[100, 3, 2] = np.nan    # py of particle 2 inside both windows
    h[3150, 0, 5] = np.nan   # x of particle 5 after the windows only
    windows = [(37, 1536), (99, 3099)]
    if layout == "rows":
        L = (N_all + 15) // 16 * 16
        rows = np.zeros((n, 4, L))
        rows[:, :, :N_all] = h.transpose(2, 1, 0)
        d, row_len = C.DeviceBuffer.from_array(np.ascontiguousarray(rows)), L
    else:
        d, row_len = C.DeviceBuffer.from_array(np.ascontiguousarray(h)), None
    for mean in (False, True):
        one = C.history_tunes(d, n, N_all, windows, "fft", subtract_mean=mean, row_len=row_len)
        monkeypatch.setenv("HM_FFT_ROW_BYTES", str(3 * (2 * 16 * 8192 + 2 * 8 * 3001 + 8)))  # 3 particles
        many = C.history_tunes(d, n, N_all, windows, "fft", subtract_mean=mean, row_len=row_len)
        monkeypatch.delenv("HM_FFT_ROW_BYTES")
        assert np.array_equal(one, many, equal_nan=True)
        hh = h - (np.sum(h, axis=0) / N_all)[None] if mean else h
        for w, (a, b) in enumerate(windows):
            for i in range(n):
                seg = hh[a:b + 1, :, i]
                if np.isnan(seg).any():
                    assert np.all(np.isnan(one[w, :, i])), (mean, w, i)
                    continue
                want = [O.fft_tune(seg[:, 0], seg[:, 1]), O.fft_tune(seg[:, 2], seg[:, 3])]
                np.testing.assert_allclose(one[w, :, i], want, rtol=0, atol=1e-12, err_msg=f"{mean} {w} {i}")
        assert np.all(np.isnan(one[:, :, 2]))
        assert np.all(np.isnan(one[:, :, 5])) == mean


def test_all_tunes_over_1e5_turns():
    """all_tunes appends the (0, n_turns) window and the full signal: 100001
    and 100000 samples at 1e5 turns. FFT tunes against numpy's FFT of the
    oracle's mean-removed trajectory, Birkhoff tunes against the oracle."""
    import henon_map_cpp_amd as hm

    n_turns = 100_000
    x = np.array([0.1, 0.2, 0.31])
    z = np.zeros(3)
    t = hm.henon_tracker(n_turns, 0.168, 0.201, "sps", epsilon=8.0)
    p = hm.particles(x, z, x, z)
    df = t.tune_all(p, n_turns, 0.3, from_idx=np.array([0, 20000]), to_idx=np.array([2047, 70000]))
    assert list(df["to"]) == [2047, 70000, n_turns]
    rb, rf = (np.asarray(v) for v in t.tracker.all_tunes(hm.particles(x, z, x, z).particles, n_turns, 0.3, 10.0,
                                                         float("nan"), False, [], []))
    tables = O.Tables(O.sps_modulation(0.168, 8.0, 0, n_turns), O.sps_modulation(0.201, 8.0, 0, n_turns))
    op = O.Particles(x, z, x, z)
    ot = O.Tracker.from_tables(tables)
    hist = [np.stack([op.x, op.px, op.y, op.py])]
    for _ in range(n_turns):
        ot.track(op, 1, 0.3, barrier=10.0)
        hist.append(np.stack([op.x, op.px, op.y, op.py]))
    hist = np.array(hist)
    assert np.array_equal(p.get_x(), op.x, equal_nan=True)
    s = np.zeros((4, 3))
    for m in range(n_turns + 1):
        s = s + hist[m]
    hh = hist - (s / (n_turns + 1))[None]
    for i in range(3):
        if np.isnan(hh[:, :, i]).any():
            assert np.isnan(df["tune_x_fft"][2][i])
            continue
        full = [numpy_fft_tune(hh[1:, 0, i], hh[1:, 1, i]), numpy_fft_tune(hh[1:, 2, i], hh[1:, 3, i])]
        win = [numpy_fft_tune(hh[:, 0, i], hh[:, 1, i]), numpy_fft_tune(hh[:, 2, i], hh[:, 3, i])]
        assert df["tune_x_fft"][2][i] == pytest.approx(full[0], abs=1e-12)
        assert df["tune_y_fft"][2][i] == pytest.approx(full[1], abs=1e-12)
        np.testing.assert_allclose(rf[i, 0:2], win, rtol=0, atol=1e-12)
        w2 = [numpy_fft_tune(hh[20000:70001, 0, i], hh[20000:70001, 1, i])]
        assert df["tune_x_fft"][1][i] == pytest.approx(w2[0], abs=1e-12)
        want_b = O.birkhoff_tune_vec(hh[:, 0, i], hh[:, 1, i], hh[:, 2, i], hh[:, 3, i], [0], [2047])
        assert df["tune_x_birkhoff"][0][i] == pytest.approx(want_b[0], abs=1e-13)
        assert df["tune_x_birkhoff"][2][i] == pytest.approx(want_b[2], abs=1e-13)
