"""GPU parity of the tracking kernel, through the C ABI (libhenon_hip.so).

Exact mode is compared bit for bit (coordinates, NaN pattern and survival
turns) with the oracle on the same inputs and tables; fast mode (fused
multiply-adds) within the stated tolerances."""
import ctypes
import os

import numpy as np
import pytest

import oracle as O
from henon_map_cpp_amd import _capi as C
from henon_map_cpp_amd import workloads as W

pytestmark = pytest.mark.gpu
NT = min(os.cpu_count() or 1, 16)


def make_tables(n_turns, tune_x, tune_y, kind, epsilon, offset=0):
    if kind == "custom":
        coef, freq = W.seeded_line_coefficients(), W.line_frequencies()
        wx = C.lines_modulation(tune_x, epsilon, coef, freq, offset, offset + n_turns)
        wy = C.lines_modulation(tune_y, epsilon, coef, freq, offset, offset + n_turns)
    else:
        wx = C.sps_modulation(tune_x, epsilon, offset, offset + n_turns)
        wy = C.sps_modulation(tune_y, epsilon, offset, offset + n_turns)
    return O.Tables(wx, wy), C.build_table(wx, wy)


def run_both(x, px, y, py, tables, tab, n_turns, mu, barrier, inverse=False, mode=C.MODE_EXACT, steps=None,
             cfg=None, pre_turns=0):
    p = O.Particles(x, px, y, py)
    if steps is not None:
        p.steps[:] = steps
    dev = C.DeviceParticles(x, px, y, py, steps)
    tdev = C.DeviceBuffer.from_array(tab)
    if pre_turns:
        O.lib().or_track(*p.state(), p.n, *tables.ptrs(), pre_turns, mu, barrier, np.nan, 0, NT)
        C.track(dev, tdev, len(tab), pre_turns, mu, barrier, False, mode, cfg)
    O.lib().or_track(*p.state(), p.n, *tables.ptrs(), n_turns, mu, barrier, np.nan, int(inverse), NT)
    C.track(dev, tdev, len(tab), n_turns, mu, barrier, inverse, mode, cfg)
    got = dev.get()
    assert C.take_status() == 0
    return (p.x, p.px, p.y, p.py, p.steps), got


def assert_bit_exact(ref, got):
    for a, b, name in zip(ref, got, ["x", "px", "y", "py", "steps"]):
        assert np.array_equal(a, b, equal_nan=True), f"{name}: {np.sum(~((a == b) | (np.isnan(a) & np.isnan(b))))} differ"


def test_unmodulated_da_scan_bit_exact():
    """Unmodulated DA scan, 100x100 grid, 1e4 turns, mu=0: the full workload."""
    c = W.UNMODULATED_DA
    x, px, y, py = W.da_grid(c["n_side"])
    tables, tab = make_tables(c["n_turns"], c["tune_x"], c["tune_y"], "sps", 0.0)
    ref, got = run_both(x, px, y, py, tables, tab, c["n_turns"], c["mu"], c["barrier"])
    assert_bit_exact(ref, got)
    lost = np.isnan(got[0])
    assert 0.05 < lost.mean() < 0.95  # the scan straddles the dynamic aperture


@pytest.mark.parametrize("kind", ["sps", "custom"])
def test_modulated_da_slice_bit_exact(kind):
    """Modulated map (mu=0.5, eps=32) on a 100x100 grid of the DA range, 1e4 turns."""
    c = W.MODULATED_DA
    x, px, y, py = W.da_grid(100)
    tables, tab = make_tables(10_000, c["tune_x"], c["tune_y"], kind, c["epsilon"])
    ref, got = run_both(x, px, y, py, tables, tab, 10_000, c["mu"], c["barrier"])
    assert_bit_exact(ref, got)


def test_full_modulated_scan_sampled_bit_exact():
    """The full modulated DA scan (1000x1000 grid, 1e5 turns, mu=0.5, seeded
    7-line modulation at eps=32: the benchmark's workload) with the default
    launch shape and loss compaction; every 499th particle (2005, across the
    whole grid) equals the oracle's 1e5-turn tracking bit for bit."""
    c = W.MODULATED_DA
    x, px, y, py = W.da_grid(c["n_side"])
    tables, tab = make_tables(c["n_turns"], c["tune_x"], c["tune_y"], "custom", c["epsilon"])
    dev = C.DeviceParticles(x, px, y, py)
    C.track(dev, C.DeviceBuffer.from_array(tab), len(tab), c["n_turns"], c["mu"], c["barrier"])
    got = dev.get()
    assert C.take_status() == 0
    pick = np.arange(0, x.size, 499)
    p = O.Particles(x[pick], px[pick], y[pick], py[pick])
    O.lib().or_track(*p.state(), p.n, *tables.ptrs(), c["n_turns"], c["mu"], c["barrier"], np.nan, 0, NT)
    assert_bit_exact((p.x, p.px, p.y, p.py, p.steps), tuple(g[pick] for g in got))
    survived = got[4] == c["n_turns"]
    assert 0.2 < survived.mean() < 0.8 and (got[4] < 1000).any()


@pytest.mark.parametrize("mu", [0.0, 0.5])
def test_inverse_tracking_bit_exact(mu):
    x, px, y, py = W.ball(4096, 0.45, seed=11)
    tables, tab = make_tables(3000, 0.31, 0.32, "sps", 16.0)
    ref, got = run_both(x, px, y, py, tables, tab, 1500, mu, 10.0, inverse=True, pre_turns=2000)
    assert_bit_exact(ref, got)
    back = got[4] == 500
    assert back.sum() > 100


def test_odd_turn_counts_and_resume():
    x, px, y, py = W.ball(1000, 0.5, seed=2)
    tables, tab = make_tables(200, 0.28, 0.31, "sps", 8.0)
    for n in (1, 2, 3, 7):
        ref, got = run_both(x, px, y, py, tables, tab, n, 0.3, 1.0)
        assert_bit_exact(ref, got)
    # track(60); track(61) == track(121)
    dev = C.DeviceParticles(x, px, y, py)
    tdev = C.DeviceBuffer.from_array(tab)
    C.track(dev, tdev, len(tab), 60, 0.3, 1.0)
    C.track(dev, tdev, len(tab), 61, 0.3, 1.0)
    ref, got = run_both(x, px, y, py, tables, tab, 121, 0.3, 1.0)
    assert_bit_exact(ref, dev.get())


def test_nan_inputs_and_barrier():
    x = np.array([np.nan, 0.1, 5.0, 0.0, 0.2] * 20)
    z = np.zeros_like(x)
    y = np.array([0.1, 0.1, 0.0, 0.0, 0.3] * 20)
    tables, tab = make_tables(100, 0.168, 0.201, "sps", 0.0)
    ref, got = run_both(x, z, y, z, tables, tab, 100, 0.0, 1.0)
    assert_bit_exact(ref, got)
    assert got[4][0] == 0 and np.isnan(got[0][0])
    assert got[4][2] == 0 and np.isnan(got[0][2])


def test_mixed_turn_indices_use_per_lane_tables():
    """Particles of one wavefront at different turn indices (a partial reset)."""
    x, px, y, py = W.ball(640, 0.4, seed=4)
    steps = np.where(np.arange(640) % 3 == 0, 0, 37).astype(np.uint32)
    tables, tab = make_tables(500, 0.31, 0.32, "sps", 16.0)
    ref, got = run_both(x, px, y, py, tables, tab, 300, 0.5, 10.0, steps=steps)
    assert_bit_exact(ref, got)


def test_out_of_table_index_is_flagged_not_read():
    x, px, y, py = W.ball(128, 0.2, seed=5)
    tables, tab = make_tables(100, 0.31, 0.32, "sps", 0.0)
    dev = C.DeviceParticles(x, px, y, py, np.full(128, 95, dtype=np.uint32))
    tdev = C.DeviceBuffer.from_array(tab)
    C.track(dev, tdev, len(tab), 10, 0.0, 10.0)  # indices 95..104 leave the table
    assert C.take_status() & 1
    got = dev.get()
    assert np.all(got[4] == 100)  # stopped at the table's end, not stepped beyond
    assert C.take_status() == 0


def test_inverse_past_table_end_is_flagged_not_read():
    """Inverse turns read records steps-1 downwards: particles whose turn index
    lies past the end of a shorter table are stopped unchanged and flagged."""
    x, px, y, py = W.ball(128, 0.2, seed=5)
    tables, tab = make_tables(100, 0.31, 0.32, "sps", 0.0)
    dev = C.DeviceParticles(x, px, y, py, np.full(128, 105, dtype=np.uint32))
    tdev = C.DeviceBuffer.from_array(tab)
    C.track(dev, tdev, len(tab), 10, 0.0, 10.0, inverse=True)  # indices 104..95: the first is past the end
    assert C.take_status() & 1
    got = dev.get()
    assert np.all(got[4] == 105)
    for a, b in zip(got[:4], (x, px, y, py)):
        assert np.array_equal(a, b)
    assert C.take_status() == 0


def test_inverse_after_longer_table_raises():
    """A particle set advanced past a tracker's table cannot be tracked back by it."""
    import henon_map_cpp_amd as H

    x, px, y, py = W.ball(64, 0.1, seed=6)
    p = H.particles(x, px, y, py)
    long_t = H.henon_tracker(200, 0.31, 0.32, "sps")
    short_t = H.henon_tracker(100, 0.31, 0.32, "sps")
    long_t.track(p, 150, 0.0, 10.0)
    with pytest.raises(RuntimeError, match="too large"):
        short_t.track(p, 10, 0.0, 10.0, inverse=True)


@pytest.mark.parametrize("every", [1, 7, 250])
def test_loss_compaction_is_bit_identical(every):
    c = W.MODULATED_DA
    x, px, y, py = W.da_grid(64)
    tables, tab = make_tables(2000, c["tune_x"], c["tune_y"], "custom", c["epsilon"])
    cfg = C.LaunchCfg(0, every, 0, 0)
    ref, got = run_both(x, px, y, py, tables, tab, 2000, c["mu"], c["barrier"], cfg=cfg)
    assert_bit_exact(ref, got)


@pytest.mark.parametrize("mode", [C.MODE_EXACT, C.MODE_FAST])
def test_one_and_two_particles_per_lane_agree(mode):
    """775^2 = 600625 particles exceed one round of resident wavefronts (256 CUs
    x 32 waves x 64 lanes), so particles_per_lane=2 runs the mixed launch; both
    launch shapes must give identical results, and exact mode the oracle's."""
    c = W.MODULATED_DA
    x, px, y, py = W.da_grid(775)
    tables, tab = make_tables(400, c["tune_x"], c["tune_y"], "custom", c["epsilon"])
    tdev = C.DeviceBuffer.from_array(tab)
    outs = []
    for pair in (1, 2):
        dev = C.DeviceParticles(x, px, y, py)
        C.track(dev, tdev, len(tab), 400, c["mu"], c["barrier"], False, mode, C.LaunchCfg(0, 100, pair, 0))
        outs.append(dev.get())
    assert C.take_status() == 0
    assert_bit_exact(outs[0], outs[1])
    if mode == C.MODE_EXACT:
        p = O.Particles(x, px, y, py)
        O.lib().or_track(*p.state(), p.n, *tables.ptrs(), 400, c["mu"], c["barrier"], np.nan, 0, NT)
        assert_bit_exact((p.x, p.px, p.y, p.py, p.steps), outs[0])


@pytest.mark.parametrize("inverse", [False, True])
@pytest.mark.parametrize("mode", [C.MODE_EXACT, C.MODE_FAST])
def test_table_staging_ring_and_scalar_agree(inverse, mode):
    """The two ways a wavefront reads its table records (scalar loads, or 8-turn
    chunks through an LDS ring) give identical results: losses mid-chunk, a
    turn count that is not a multiple of 8, compaction segments, the mixed
    two-per-lane launch, and exact mode equal to the oracle."""
    c = W.MODULATED_DA
    x, px, y, py = W.da_grid(48)
    n_turns = 1203
    tables, tab = make_tables(2 * n_turns, c["tune_x"], c["tune_y"], "custom", c["epsilon"])
    tdev = C.DeviceBuffer.from_array(tab)
    outs = []
    for staging in (1, 2):
        for pair, every in ((1, 0), (2, 250), (1, -1)):
            dev = C.DeviceParticles(x, px, y, py)
            cfg = C.LaunchCfg(0, every, pair, staging)
            C.track(dev, tdev, len(tab), n_turns, c["mu"], c["barrier"], False, mode, cfg)
            if inverse:
                C.track(dev, tdev, len(tab), 611, c["mu"], c["barrier"], True, mode, cfg)
            outs.append(dev.get())
    assert C.take_status() == 0
    for o in outs[1:]:
        assert_bit_exact(outs[0], o)
    lost = np.isnan(outs[0][0])
    assert 0.05 < lost.mean() < 0.95
    if mode == C.MODE_EXACT:
        p = O.Particles(x, px, y, py)
        O.lib().or_track(*p.state(), p.n, *tables.ptrs(), n_turns, c["mu"], c["barrier"], np.nan, 0, NT)
        if inverse:
            O.lib().or_track(*p.state(), p.n, *tables.ptrs(), 611, c["mu"], c["barrier"], np.nan, 1, NT)
        assert_bit_exact((p.x, p.px, p.y, p.py, p.steps), outs[0])
    with pytest.raises(ValueError, match="table_staging"):
        C.track(C.DeviceParticles(x, px, y, py), tdev, len(tab), 10, c["mu"], c["barrier"], False, mode,
                C.LaunchCfg(0, 0, 0, 3))


@pytest.mark.parametrize("block", [64, 128, 256])
def test_block_sizes(block):
    x, px, y, py = W.da_grid(40)
    tables, tab = make_tables(500, 0.168, 0.201, "sps", 32.0)
    ref, got = run_both(x, px, y, py, tables, tab, 500, 0.5, 3.0, cfg=C.LaunchCfg(block, 0, 0, 0))
    assert_bit_exact(ref, got)


def test_fast_mode_tolerances():
    """Fused multiply-adds round differently from the reference's separate
    multiplies and adds. Regular orbits agree to 1e-12 relative. On a DA grid
    the loss turn of a particle at the aperture border depends on every
    rounding, so fast-mode survival turns differ for a sizeable fraction of the
    grid (measured on MI355X: 77% identical for the unmodulated scan, 84% for the
    modulated one, at 1e4 turns); the fractions are printed, not asserted.
    Exact mode is the mode that reproduces the reference."""
    x, px, y, py = W.ball(4096, 0.15, seed=8)
    tables, tab = make_tables(1000, 0.168, 0.201, "custom", 32.0)
    ref, got = run_both(x, px, y, py, tables, tab, 1000, 0.5, 10.0, mode=C.MODE_FAST)
    r = np.sqrt(ref[0] ** 2 + ref[1] ** 2 + ref[2] ** 2 + ref[3] ** 2)
    for a, b in zip(ref[:4], got[:4]):
        assert np.nanmax(np.abs(a - b) / r) < 1e-12
    for name, c, kind in [("unmodulated", W.UNMODULATED_DA, "sps"), ("modulated", W.MODULATED_DA, "custom")]:
        x, px, y, py = W.da_grid(100)
        tables, tab = make_tables(10_000, c["tune_x"], c["tune_y"], kind, c["epsilon"])
        ref, got = run_both(x, px, y, py, tables, tab, 10_000, c["mu"], c["barrier"], mode=C.MODE_FAST)
        same = np.mean(ref[4] == got[4])
        stable = ref[4] == 10_000
        print(f"fast mode, {name} DA 100x100 x 1e4: survival turns identical for {same:.4f}; "
              f"stable core kept {np.mean(got[4][stable] == 10_000):.4f}")
        assert np.mean(got[4][stable] == 10_000) >= 0.9


def test_observables_match_reference_formulas():
    x, px, y, py = W.ball(1000, 0.5, seed=9)
    x[::7] = np.nan
    dev = C.DeviceParticles(x, px, y, py)
    out = C.DeviceBuffer(8 * 1000)
    want = {
        C.OBS_RADIUS: np.sqrt(x * x + px * px + y * y + py * py),
        C.OBS_ACTION: x * x + px * px + y * y + py * py,
        C.OBS_ACTION_X: x * x + px * px,
        C.OBS_ACTION_Y: y * y + py * py,
        C.OBS_ANGLE_X: np.arctan2(px, x),
        C.OBS_ANGLE_Y: np.arctan2(py, y),
    }
    for kind, w in want.items():
        C.check(C.lib().hm_observable(kind, *dev.ptrs()[:4], 1000, out.ptr, None))
        got = out.download(np.float64, 1000)
        if kind in (C.OBS_ANGLE_X, C.OBS_ANGLE_Y):
            np.testing.assert_allclose(got, w, rtol=1e-15, atol=1e-300, equal_nan=True)
        else:
            assert np.array_equal(got, w, equal_nan=True)


@pytest.mark.parametrize("reverse", [0, 1])
def test_tangent_matrix_kernel(reverse):
    x, px, y, py = W.ball(777, 0.4, seed=10)
    tables, tab = make_tables(50, 0.31, 0.32, "sps", 16.0)
    dev = C.DeviceParticles(x, px, y, py)
    valid = np.ones(777, dtype=np.uint8)
    valid[5] = 0
    vdev = C.DeviceBuffer.from_array(valid)
    tdev = C.DeviceBuffer.from_array(tab)
    out = C.DeviceBuffer(777 * 16 * 8)
    C.check(C.lib().hm_tangent_matrix(*dev.ptrs()[:4], vdev.ptr, 777, tdev.ptr, 50, 3, 0.5, reverse, out.ptr, None))
    got = out.download(np.float64, 777 * 16).reshape(777, 4, 4)
    p = O.Particles(x, px, y, py)
    p.valid[:] = valid
    p.global_steps = 3
    ref = O.Tracker.from_tables(tables).get_tangent_matrix(p, 0.5, reverse=bool(reverse))
    if reverse:
        np.testing.assert_allclose(got, ref, rtol=1e-15, atol=1e-16, equal_nan=True)
    else:
        assert np.array_equal(got, ref, equal_nan=True)


def test_autotune_leaves_inputs_and_results_unchanged():
    """hm_autotune_track works on copies; the configuration it remembers (used by
    hm_track with cfg=None) gives the oracle's results bit for bit."""
    c = W.MODULATED_DA
    x, px, y, py = W.da_grid(120)
    tables, tab = make_tables(3000, c["tune_x"], c["tune_y"], "custom", c["epsilon"])
    tdev = C.DeviceBuffer.from_array(tab)
    dev = C.DeviceParticles(x, px, y, py)
    try:
        best, ms = C.autotune_track(dev, tdev, len(tab), c["mu"], c["barrier"], warm_turns=500, trial_turns=1000)
        assert best.block in (64, 256) and best.compact_every in (500, 1000, 2000)
        assert best.particles_per_lane in (1, 2) and ms > 0
        before = dev.get()
        assert all(np.array_equal(a, b) for a, b in zip(before[:4], (x, px, y, py)))
        assert np.all(before[4] == 0)
        ref, got = run_both(x, px, y, py, tables, tab, 3000, c["mu"], c["barrier"])
        assert_bit_exact(ref, got)
    finally:
        C.check(C.lib().hm_autotune_clear())


def test_zero_particles_are_a_no_op():
    """Every particle-array entry point accepts n = 0 (the reference's calls over
    empty vectors return without work) and launches nothing: NULL particle
    arrays are not touched, the status word stays clear."""
    L = C.lib()
    nul = None
    tab = C.DeviceBuffer.from_array(C.build_table(np.full(16, 0.3), np.full(16, 0.4)))
    cfg = C.LaunchCfg(0, 0, 0, 0)
    spec = C.IndicatorSpec()
    f = np.zeros(1, np.uint32)
    ln = np.full(1, 8, np.uint32)
    calls = {
        "hm_track": lambda: L.hm_track(nul, nul, nul, nul, nul, 0, tab.ptr, 16, 8, 0.5, 10.0, float("nan"), 0,
                                       C.MODE_EXACT, ctypes.byref(cfg), nul),
        "hm_track_kicked": lambda: L.hm_track_kicked(nul, nul, nul, nul, nul, 0, tab.ptr, 16, 8, 0.5, 10.0, 1e-8, 0,
                                                     C.MODE_EXACT, 1, 0, nul),
        "hm_track_indicators": lambda: L.hm_track_indicators(nul, nul, nul, nul, nul, 0, tab.ptr, 16, 8, 0.5, 10.0,
                                                             C.MODE_EXACT, ctypes.byref(spec), nul),
        "hm_track_history": lambda: L.hm_track_history(nul, nul, nul, nul, nul, 0, tab.ptr, 16, 8, 0.5, 10.0, 0, nul,
                                                       nul),
        "hm_track_rows": lambda: L.hm_track_rows(nul, nul, nul, nul, nul, 0, tab.ptr, 16, 8, 0.5, 10.0, 0, nul, 16,
                                                 nul, nul),
        "hm_fft_tunes_rows": lambda: L.hm_fft_tunes_rows(nul, 0, 16, 8, nul, C.ptr(f), C.ptr(ln), 1, nul, nul),
        "hm_birkhoff_tunes_rows": lambda: L.hm_birkhoff_tunes_rows(nul, 0, 16, 8, nul, C.ptr(f), C.ptr(ln + f - 1),
                                                                   1, nul, nul),
        "hm_observable": lambda: L.hm_observable(C.OBS_RADIUS, nul, nul, nul, nul, 0, nul, nul),
    }
    st = C.U32(0)
    C.check(L.hm_take_status(ctypes.byref(st), None))  # clear what earlier calls left
    for name, call in calls.items():
        assert call() == 0, (name, C.lib().hm_last_error())
    C.check(L.hm_stream_synchronize(None))
    C.check(L.hm_take_status(ctypes.byref(st), None))
    assert st.value == 0
