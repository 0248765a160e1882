"""Birkhoff and FFT tunes of the same windows from one read of the particle
rows (hm_tunes_rows): the FFT tunes equal hm_fft_tunes_rows bit for bit, the
Birkhoff tunes equal hm_birkhoff_tunes_rows to rounding (another summation
order) and the oracle's birkhoff_tune (dynamic_indicator.cpp:81-131) of the
window; windows of power-of-two length up to 2048 take the fused kernel, others
the two separate stages. NaN samples, NaN means and both planes are covered."""
import numpy as np
import pytest

import oracle as O
from henon_map_cpp_amd import _capi as C

pytestmark = pytest.mark.gpu


def rows_of(n, n_samples, seed):
    """Particle rows [n][4][L] of a two-plane quasi-periodic signal with noise."""
    rng = np.random.default_rng(seed)
    m = np.arange(n_samples)[None, :]
    L = C.row_length(n_samples - 1)
    rows = np.zeros((n, 4, L))
    for plane in (0, 1):
        f = rng.uniform(0.05, 0.45, n)[:, None]
        amp = rng.uniform(0.02, 0.3, n)[:, None]
        z = amp * np.exp(-2j * np.pi * f * m) + 0.2 * amp * np.exp(-4j * np.pi * f * m)
        z = z + 1e-4 * (rng.normal(size=z.shape) + 1j * rng.normal(size=z.shape))
        rows[:, 2 * plane, :n_samples], rows[:, 2 * plane + 1, :n_samples] = z.real, z.imag
    return rows, L


WINDOWS = [(0, 2047), (5, 1028), (3, 10), (0, 999), (100, 1123), (1, 2048), (17, 32)]


@pytest.mark.parametrize("with_means", [False, True])
def test_fused_equals_separate_stages_and_oracle(with_means):
    n, S = 45, 2049
    rows, L = rows_of(n, S, seed=7 + with_means)
    rows[3, 1, 500] = np.nan    # inside every window but (3, 10), (17, 32)
    rows[8, 2, 2048] = np.nan   # only in (1, 2048)
    rows[11, 3, 7] = np.nan     # inside (0, 2047), (3, 10), (0, 999), (5, 1028)
    d = C.DeviceBuffer.from_array(np.ascontiguousarray(rows))
    means = None
    if with_means:
        # means over every sample (NaN for particles 3, 8, 11: every window of
        # theirs is NaN then, without a read)
        means = C.DeviceBuffer(8 * 4 * n)
        C.check(C.lib().hm_rows_means(d.ptr, n, L, S, means.ptr, None))
    birk, fft = C.tunes_rows(d, n, S, WINDOWS, L, means=means)
    want_f = C.history_tunes(d, n, S, WINDOWS, "fft", row_len=L, means=means)
    want_b = C.history_tunes(d, n, S, WINDOWS, "birkhoff", row_len=L, means=means)
    assert np.array_equal(fft, want_f, equal_nan=True)
    assert np.array_equal(np.isnan(birk), np.isnan(want_b))
    np.testing.assert_allclose(birk, want_b, rtol=0, atol=4e-15, equal_nan=True)
    mv = means.download(np.float64, 4 * n).reshape(4, n) if with_means else np.zeros((4, n))
    for w, (a, b) in enumerate(WINDOWS):
        for i in range(n):
            seg = rows[i, :, a:b + 1] - mv[:, i:i + 1]
            if np.isnan(seg).any():
                assert np.isnan(birk[w, :, i]).all() and np.isnan(fft[w, :, i]).all(), (w, i)
                continue
            ob = [O.birkhoff_tune(seg[0], seg[1]), O.birkhoff_tune(seg[2], seg[3])]
            np.testing.assert_allclose(birk[w, :, i], ob, rtol=0, atol=1e-14, err_msg=f"window {w} particle {i}")
            of = [O.fft_tune(seg[0], seg[1]), O.fft_tune(seg[2], seg[3])]
            np.testing.assert_allclose(fft[w, :, i], of, rtol=0, atol=1e-12, err_msg=f"window {w} particle {i}")
    assert np.isnan(birk[0, :, 3]).all() and np.isnan(birk[5, :, 8]).all()
    assert np.isfinite(birk[2, :, 3]).all() != with_means and np.isfinite(birk[0, :, 8]).all() != with_means


def test_nan_mean_skips_the_particle():
    n, S = 9, 2048
    rows, L = rows_of(n, S, seed=3)
    means = np.zeros((4, n))
    means[2, 4] = np.nan
    d = C.DeviceBuffer.from_array(np.ascontiguousarray(rows))
    m = C.DeviceBuffer.from_array(np.ascontiguousarray(means))
    birk, fft = C.tunes_rows(d, n, S, [(0, 2047)], L, means=m)
    assert np.isnan(birk[0, :, 4]).all() and np.isnan(fft[0, :, 4]).all()
    others = np.delete(np.arange(n), 4)
    assert np.isfinite(birk[0][:, others]).all() and np.isfinite(fft[0][:, others]).all()


def test_all_tunes_uses_the_fused_windows():
    """all_tunes over 2047 turns: the window (0, 2047) is fused; its Birkhoff
    tunes equal birkhoff_tunes of the mean-removed trajectory to rounding and
    its FFT tunes the oracle's."""
    import henon_map_cpp_amd as hm

    T = 2047
    rng = np.random.default_rng(11)
    x = rng.uniform(0.05, 0.3, 64)
    z = np.zeros(64)
    t = hm.henon_tracker(T, 0.168, 0.201, "sps", epsilon=16.0)
    rb, rf = (np.asarray(v) for v in t.tracker.all_tunes(hm.particles(x, z, x, z).particles, T, 0.4, 10.0,
                                                         float("nan"), False, [0, 512], [1023, 1535]))
    tables = O.Tables(O.sps_modulation(0.168, 16.0, 0, T), O.sps_modulation(0.201, 16.0, 0, T))
    op, ot = O.Particles(x, z, x, z), O.Tracker.from_tables(tables)
    hist = [np.stack([op.x, op.px, op.y, op.py])]
    for _ in range(T):
        ot.track(op, 1, 0.4, barrier=10.0)
        hist.append(np.stack([op.x, op.px, op.y, op.py]))
    hist = np.array(hist)
    s = np.zeros((4, 64))
    for k in range(T + 1):
        s = s + hist[k]
    hh = hist - (s / (T + 1))[None]
    checked = 0
    for i in range(64):
        if np.isnan(hh[:, :, i]).any():
            assert np.isnan(rb[i]).all() and np.isnan(rf[i]).all()
            continue
        want_b = O.birkhoff_tune_vec(hh[:, 0, i], hh[:, 1, i], hh[:, 2, i], hh[:, 3, i], [0, 512], [1023, 1535])
        np.testing.assert_allclose(rb[i, :6], want_b[:6], rtol=0, atol=1e-14)
        for w, (a, b) in enumerate([(0, 1023), (512, 1535), (0, T)]):
            seg = hh[a:b + 1, :, i]
            want = [O.fft_tune(seg[:, 0], seg[:, 1]), O.fft_tune(seg[:, 2], seg[:, 3])]
            np.testing.assert_allclose(rf[i, 2 * w:2 * w + 2], want, rtol=0, atol=1e-12)
        checked += 1
    assert checked > 40


@pytest.mark.parametrize("N,n", [(8, (1 << 22) + 3), (6, (1 << 21) + 5)])
def test_many_particles_take_several_launches(N, n):
    """More particles than one launch takes (4194304 per row-kernel launch;
    at 6 samples 2097152 per mixed-radix batch and per Birkhoff launch of two
    windows): the particles on both sides of each launch boundary and a random
    sample against the oracle, and the first thousand against a call of their
    own."""
    rng = np.random.default_rng(N)
    L = 16
    rows = np.zeros((n, 4, L))
    m = np.arange(N)
    f = rng.uniform(0.05, 0.45, (n, 2))
    for plane in (0, 1):
        z = 0.1 * np.exp(-2j * np.pi * f[:, plane:plane + 1] * m[None, :])
        rows[:, 2 * plane, :N], rows[:, 2 * plane + 1, :N] = z.real, z.imag
    rows[:, :, :N] += 1e-3 * rng.normal(size=(n, 4, N))
    d = C.DeviceBuffer.from_array(rows)
    windows = [(0, N - 1)] if N == 8 else [(0, N - 1), (1, N - 1)]
    birk, fft = C.tunes_rows(d, n, N, windows, L)
    boundary = (1 << 22) if N == 8 else (1 << 21)
    check = list(range(boundary - 3, min(n, boundary + 3))) + list(rng.integers(0, n, 64)) + [0, n - 1]
    for i in check:
        for w, (a, b) in enumerate(windows):
            seg = rows[i, :, a:b + 1]
            want_f = [O.fft_tune(seg[0], seg[1]), O.fft_tune(seg[2], seg[3])]
            want_b = [O.birkhoff_tune(seg[0], seg[1]), O.birkhoff_tune(seg[2], seg[3])]
            np.testing.assert_allclose(fft[w, :, i], want_f, rtol=0, atol=1e-12, err_msg=f"particle {i}")
            np.testing.assert_allclose(birk[w, :, i], want_b, rtol=0, atol=1e-14, err_msg=f"particle {i}")
    small = C.DeviceBuffer.from_array(np.ascontiguousarray(rows[:1000]))
    b1, f1 = C.tunes_rows(small, 1000, N, windows, L)
    assert np.array_equal(f1, fft[:, :, :1000]) and np.array_equal(b1, birk[:, :, :1000])
    assert np.isfinite(fft).all() and np.isfinite(birk).all()
