"""hm_track_noalloc: tracking that allocates nothing (live lists in a static
device array). Its results must equal hm_track's bit for bit; two calls in a
row, the second smaller, check that the array's leftover contents from one call
do not leak into the next."""
import numpy as np
import pytest

from henon_map_cpp_amd import _capi as C
from henon_map_cpp_amd import workloads as W


def test_noalloc_rejects_oversize_before_device_work():
    """Needs no GPU: the size limit is checked before anything touches a device."""
    rc = C.lib().hm_track_noalloc(None, None, None, None, None, (1 << 20) + 1, None, 10, 10, 0.0, 1.0,
                                  float("nan"), 0, 0, None, None)
    assert rc == C.HM_ERR_ARG
    assert "1048576" in C.lib().hm_last_error().decode()


def tables(n_turns):
    c = W.MODULATED_DA
    coef, freq = W.seeded_line_coefficients(), W.line_frequencies()
    wx = C.lines_modulation(c["tune_x"], c["epsilon"], coef, freq, 0, n_turns)
    wy = C.lines_modulation(c["tune_y"], c["epsilon"], coef, freq, 0, n_turns)
    return C.DeviceBuffer.from_array(C.build_table(wx, wy))


def run(entry, grid, tdev, n_turns, mode, cfg):
    x, px, y, py = W.da_grid(grid)
    dev = C.DeviceParticles(x, px, y, py)
    c = ctypes_cfg(cfg)
    C.check(entry(*dev.ptrs(), dev.n, tdev.ptr, n_turns, n_turns, W.MODULATED_DA["mu"],
                  W.MODULATED_DA["barrier"], float("nan"), 0, mode, c, None))
    C.check(C.lib().hm_stream_synchronize(None))
    return dev.get()


def ctypes_cfg(cfg):
    import ctypes

    return None if cfg is None else ctypes.byref(cfg)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [C.MODE_EXACT, C.MODE_FAST])
def test_noalloc_equals_track_two_calls(mode):
    n_turns = 3000
    tdev = tables(n_turns)
    lib = C.lib()
    # first call: 775^2 particles, more than one round of resident wavefronts,
    # two per lane in the leading ones, default segments (125, 250, 500, 1000,
    # 1000, 125 turns); second call: fewer particles, 7-turn segments (429 of
    # them), 64-thread workgroups
    for grid, cfg in ((775, C.LaunchCfg(0, 0, 2, 0)), (130, C.LaunchCfg(64, 7, 1, 0))):
        want = run(lib.hm_track, grid, tdev, n_turns, mode, cfg)
        got = run(lib.hm_track_noalloc, grid, tdev, n_turns, mode, cfg)
        for a, b in zip(got, want):
            assert np.array_equal(a, b, equal_nan=True)
        assert np.sum(got[4] < n_turns) > 0  # losses: the compaction path ran
    assert C.take_status() == 0


@pytest.mark.gpu
def test_noalloc_rejects_too_many_segments():
    tdev = tables(70000)
    x, px, y, py = W.ball(64, 0.1, seed=1)
    dev = C.DeviceParticles(x, px, y, py)
    cfg = C.LaunchCfg(0, 1, 0, 0)
    import ctypes

    rc = C.lib().hm_track_noalloc(*dev.ptrs(), dev.n, tdev.ptr, 70000, 70000, 0.0, 10.0, float("nan"), 0, 0,
                                  ctypes.byref(cfg), None)
    assert rc == C.HM_ERR_ARG
    assert "65536" in C.lib().hm_last_error().decode()
    assert np.array_equal(dev.get()[0], x)  # nothing ran
