"""The chaos-indicator workload's other outputs at full size (256x256 grid +
65536 ball points r < 0.5, mu = 0.5, seeded 7-line modulation at eps = 32):
the reversibility error at 1e3, 1e4 and 1e5 turns and the 2048-turn Birkhoff
and FFT tunes, computed for all 131072 particles on the GPU and checked on a
sample of them (every 257th, grid and ball) against the oracle.

Tolerances: the forward and inverse maps are exact-mode (no transcendental), so
trajectories, survival turns and reversibility errors are bit for bit; the
Birkhoff tunes sum the same terms as the oracle in a wavefront-parallel order
(1e-13); the FFT tunes come from a different DFT than the oracle's (1e-12)."""
import os

import numpy as np
import pytest

import oracle as O
from henon_map_cpp_amd import _capi as C
from henon_map_cpp_amd import workloads as W

pytestmark = pytest.mark.gpu
NT = min(os.cpu_count() or 1, 16)
C3 = W.CHAOS_INDICATORS


def tables(T):
    coef, freq = W.seeded_line_coefficients(), W.line_frequencies()
    wx = C.lines_modulation(C3["tune_x"], C3["epsilon"], coef, freq, 0, T)
    wy = C.lines_modulation(C3["tune_y"], C3["epsilon"], coef, freq, 0, T)
    return O.Tables(wx, wy), C.DeviceBuffer.from_array(C.build_table(wx, wy))


def test_reversibility_error_full_set_sampled():
    T = C3["n_turns"]
    ot, tdev = tables(T)
    x0 = W.indicator_points()
    pick = np.arange(0, x0[0].size, 257)
    parts = C.DeviceParticles(*x0)
    done = 0
    for t in (1_000, 10_000, 100_000):
        C.track(parts, tdev, T, t - done, C3["mu"], C3["barrier"], False, C.MODE_EXACT)
        done = t
        g = parts.get()
        snap = C.DeviceParticles(*g[:4], steps=g[4])
        C.track(snap, tdev, T, t, C3["mu"], C3["barrier"], True, C.MODE_EXACT)
        back = snap.get()
        err = np.sqrt(sum((b - a) ** 2 for a, b in zip(x0, back[:4])))
        # the oracle: t turns forward, t turns back, on the sample
        p = O.Particles(*(a[pick] for a in x0))
        tr = O.Tracker.from_tables(ot)
        tr.track(p, t, C3["mu"], barrier=C3["barrier"], n_threads=NT)
        assert np.array_equal(g[0][pick], p.x, equal_nan=True) and np.array_equal(g[4][pick], p.steps), t
        tr.track(p, t, C3["mu"], barrier=C3["barrier"], inverse=True, n_threads=NT)
        assert np.array_equal(back[0][pick], p.x, equal_nan=True) and np.array_equal(back[3][pick], p.py,
                                                                                       equal_nan=True), t
        assert np.array_equal(back[4][pick], p.steps), t
        want = np.sqrt(sum((b - a[pick]) ** 2 for a, b in zip(x0, (p.x, p.px, p.y, p.py))))
        assert np.array_equal(err[pick], want, equal_nan=True), t
        ok = np.isfinite(err)
        assert ok.any() and (~ok).any()


def test_2048_turn_tunes_full_set_sampled():
    """all_tunes' stages on the full set: particle-row trajectory of samples
    0..2047 with the per-coordinate means, Birkhoff and FFT tunes of the
    mean-removed window."""
    T = 2048
    ot, tdev = tables(T)
    x0 = W.indicator_points()
    n = x0[0].size
    parts = C.DeviceParticles(*x0)
    rows, means = C.track_rows(parts, tdev, T, T - 1, C3["mu"], C3["barrier"], means=True)
    L = C.row_length(T - 1)
    fft = C.history_tunes(rows, n, T, [(0, T - 1)], "fft", row_len=L, means=means)[0]
    birk = C.history_tunes(rows, n, T, [(0, T - 1)], "birkhoff", row_len=L, means=means)[0]
    pick = np.arange(0, n, 257)
    p = O.Particles(*(a[pick] for a in x0))
    tr = O.Tracker.from_tables(ot)
    h = [np.stack([p.x, p.px, p.y, p.py])]
    for _ in range(T - 1):
        tr.track(p, 1, C3["mu"], barrier=C3["barrier"], n_threads=1)
        h.append(np.stack([p.x, p.px, p.y, p.py]))
    h = np.array(h)  # (2048, 4, n_pick)
    s = np.zeros((4, pick.size))
    for m in range(T):
        s = s + h[m]
    hh = h - (s / T)[None]
    n_fin = 0
    for q, i in enumerate(pick):
        if np.isnan(hh[:, :, q]).any():
            assert np.all(np.isnan(fft[:, i])) and np.all(np.isnan(birk[:, i]))
            continue
        n_fin += 1
        wb = O.birkhoff_tune_vec(hh[:, 0, q], hh[:, 1, q], hh[:, 2, q], hh[:, 3, q], [0], [T - 1])
        np.testing.assert_allclose(birk[:, i], wb[0:2], rtol=0, atol=1e-13, err_msg=str(i))
        if q % 5 == 0:  # the oracle's long-double DFT on every fifth sampled particle
            wf = [O.fft_tune(hh[:, 0, q], hh[:, 1, q]), O.fft_tune(hh[:, 2, q], hh[:, 3, q])]
            np.testing.assert_allclose(fft[:, i], wf, rtol=0, atol=1e-12, err_msg=str(i))
    assert n_fin > 300
