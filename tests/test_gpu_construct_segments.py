"""The fused indicator kernel with a trajectory window and construct-form turn
samples (megno_birkhoff_construct, tune_birkhoff_construct; src/henon.cu:
2302-2655) over several loss-compaction segments and straight runs of plain
turns: particles lost in every segment, particles NaN from the start (all four
coordinates or one of them), turn samples that end before, at and after the
call's last turn. Every output equals the oracle's tracker and constructs fed
turn by turn, and the MEGNO/FLI/Birkhoff outputs equal those of a call without
the extra windows bit for bit."""
import numpy as np
import pytest

import oracle as O
from henon_map_cpp_amd import _capi as C

pytestmark = pytest.mark.gpu


def start(n, seed, hi):
    rng = np.random.default_rng(seed)
    x = [rng.uniform(0.02, hi, n), rng.uniform(-0.03, 0.03, n), rng.uniform(0.02, hi, n), np.zeros(n)]
    x[0][0] = x[1][0] = x[2][0] = x[3][0] = np.nan  # NaN from the start
    x[0][1] = np.nan  # x only: the y plane stays finite
    x[3][2] = np.nan  # py only
    return x


def expected(x0, tables, T, mu, mb, tb, hist):
    """The oracle's tracker and constructs, one turn at a time. The fused kernel
    sets a particle's tangent matrix to NaN in the turn it is lost (the MEGNO
    sums of that turn are NaN); the same is done here before the add."""
    n = x0[0].size
    ot, op = O.Tracker.from_tables(tables), O.Particles(*x0)
    om = np.tile(np.eye(4), (n, 1, 1)).reshape(n, 16)
    omb, otb = O.MegnoBirkhoffConstruct(n, mb), O.TuneBirkhoffConstruct(n, tb)
    otb.first_add(op.x, op.px, op.y, op.py)
    h = [np.stack([op.x, op.px, op.y, op.py])]
    for k in range(T):
        was = ~(np.isnan(op.x) | np.isnan(op.px) | np.isnan(op.y) | np.isnan(op.py))
        om_new = O.structured_multiply(om, op.x, op.px, op.y, op.py, tables, k, mu)
        ot.track(op, 1, mu, barrier=10.0)
        om_new[was & np.isnan(op.x)] = np.nan
        omb.add(om_new, om)
        om = om_new
        otb.add(op.x, op.px, op.y, op.py)
        h.append(np.stack([op.x, op.px, op.y, op.py]))
    h = np.array(h)[hist[0]:hist[0] + hist[1]]
    return op, omb.get_values(), np.stack([otb.t1x, otb.t2x, otb.t1y, otb.t2y], axis=1), h


@pytest.mark.parametrize("n,T,hi,seed", [(160, 700, 0.6, 61), (3000, 420, 0.9, 62)])
def test_windows_and_constructs_over_segments(n, T, hi, seed):
    mu = 0.5
    wx, wy = C.sps_modulation(0.168, 32.0, 0, T + 1), C.sps_modulation(0.201, 32.0, 0, T + 1)
    tables = O.Tables(wx, wy)
    tdev = C.DeviceBuffer.from_array(C.build_table(wx, wy))
    x0 = start(n, seed, hi)
    mb = [1, 125, T // 2 + 1, T]
    tb = [1, 63, T // 2, T + 1]  # T + 1: the second half takes no sample
    hist = (100, T + 1 - 100)
    samples, birk = [125, T], [(0, 125), (200, T)]
    got = C.track_indicators(C.DeviceParticles(*x0), tdev, T + 1, T, mu, 10.0, samples=samples, birkhoff=birk,
                             hist=hist, megno_birkhoff=mb, tune_birkhoff=tb, matrix=True)
    op, want_mb, want_tb, want_h = expected(x0, tables, T, mu, mb, tb, hist)
    lost = np.isnan(op.x[3:])
    assert lost.any() and (~lost).any()
    assert np.array_equal(got["hist"], want_h, equal_nan=True)
    np.testing.assert_allclose(got["megno_birkhoff"], want_mb, rtol=1e-12, atol=1e-14, equal_nan=True)
    np.testing.assert_allclose(got["tune_birkhoff"], want_tb, rtol=0, atol=1e-12, equal_nan=True)
    assert np.isfinite(got["tune_birkhoff"][1:3, 2, 1]).all()  # x NaN from the start: finite y sums
    assert (got["tune_birkhoff"][3, 1] == 0).all()  # T + 1 > n_turns: the second half is empty
    # the same call without the extra windows (another kernel variant)
    plain = C.track_indicators(C.DeviceParticles(*x0), tdev, T + 1, T, mu, 10.0, samples=samples, birkhoff=birk,
                               matrix=True)
    for key in ("megno", "fli", "birkhoff", "matrix", "log_scale"):
        assert np.array_equal(got[key], plain[key], equal_nan=True), key
    want = O.indicators(O.Particles(*x0), tables, T, mu, 10.0, 0, samples, birkhoff=birk)
    np.testing.assert_allclose(got["megno"], want["megno"], rtol=1e-11, atol=1e-12, equal_nan=True)
    np.testing.assert_allclose(got["birkhoff"], want["birkhoff"], rtol=0, atol=1e-13, equal_nan=True)


@pytest.mark.parametrize("renorm", [0, 100])
def test_held_particles_in_batch_and_single_calls(renorm):
    """Particles with some but not all coordinates NaN are not tracked, but
    their tangent matrix and MEGNO go on at the held point, over every segment
    of a batched MEGNO job and of a single call, as the oracle's loop does."""
    n, T, mu = 700, 900, 0.5
    x0 = start(n, 63, 0.7)
    x0[1][4] = np.nan  # px only
    x0[0][5] = x0[2][5] = np.nan  # x and y
    wx, wy = C.sps_modulation(0.31, 8.0, 0, T), C.sps_modulation(0.32, 8.0, 0, T)
    tdev = C.DeviceBuffer.from_array(C.build_table(wx, wy))
    single = C.track_indicators(C.DeviceParticles(*x0), tdev, T, T, mu, 10.0, renorm_every=renorm,
                                samples=[300, T])
    parts, meg = C.DeviceParticles(*x0), C.DeviceBuffer(8 * n)
    C.track_batch([(parts, tdev, T, T, mu, 10.0, meg, renorm)], C.MODE_EXACT, C.BatchCfg(0, 300, 0, 0))
    assert C.take_status() == 0
    assert np.array_equal(meg.download(np.float64, n), single["megno"][1], equal_nan=True)
    got = parts.get()
    nan0 = [0, 1, 2, 4, 5]  # all NaN; x; py; px; x and y
    for q in range(4):
        assert np.array_equal(got[q][nan0], x0[q][nan0], equal_nan=True)  # not moved
    want = O.indicators(O.Particles(*x0), O.Tables(wx, wy), T, mu, 10.0, renorm, [300, T])
    np.testing.assert_allclose(single["megno"], want["megno"], rtol=1e-11, atol=1e-12, equal_nan=True)
    assert np.isfinite(single["megno"][:, [2, 4]]).all()  # py or px NaN: the matrix stays finite
    assert np.isnan(single["megno"][:, [0, 1, 5]]).all()  # all NaN, or x or y NaN
