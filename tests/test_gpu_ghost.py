"""Shadow (ghost) particles through the drop-in API on the GPU: construction,
layout accessors, displacement module/direction, renormalisation and tracking,
bit for bit against the oracle (henon.cu:737-952, 1247-1442), and the reference
notebook's printed values (tests/test_new_standard.ipynb cells 44-54)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hm():
    import henon_map_cpp_amd as hm

    assert hm.gpu_available()
    return hm


def test_ghost_notebook_sequence(hm, golden):
    g = golden["ghost_3x5"]
    p = hm.particles(np.ones(3), np.ones(3), np.ones(3), np.ones(3))
    for d in ("x", "px", "y", "py", "random"):
        p.add_ghost(1e-4, d)
    np.testing.assert_allclose(p.get_x(), g["x"], rtol=0, atol=0.6 * 10.0 ** -g["x_digits"])
    direction = p.get_displacement_direction()
    assert direction.shape == (4, 3, 5)
    np.testing.assert_allclose(direction[:, :, 4], g["random_direction"], rtol=0,
                               atol=0.6 * 10.0 ** -g["direction_digits"])
    np.testing.assert_allclose(p.get_displacement_module(), np.full((3, 5), g["module"]), rtol=1e-11)
    p.renormalize(1)
    np.testing.assert_allclose(p.get_displacement_module(), g["module_after_renormalize_1"], rtol=1e-14)
    # the oracle, step by step: bit for bit
    o = O.Particles(np.ones(3), np.ones(3), np.ones(3), np.ones(3))
    for d in ("x", "px", "y", "py", "random"):
        o.add_ghost(1e-4, d)
    o.renormalize(1.0)
    assert np.array_equal(p.get_x(), o.x) and np.array_equal(p.get_py(), o.py)
    assert np.array_equal(p.get_displacement_direction(), o.displacement_direction())


def test_ghost_layout_accessors(hm):
    x = np.linspace(0.1, 0.4, 5)
    p = hm.particles(x, np.zeros(5), x, np.zeros(5))
    p.add_ghost(1e-3, "px")
    p.add_ghost(1e-2, "y")
    pp = p.particles
    assert int(pp.get_n_particles()) == 5 and int(pp.get_n_ghosts_per_particle()) == 2
    assert list(pp.get_ghost()) == [0] * 5 + [1] * 10
    assert list(pp.get_idx()) == list(range(5)) + [0] * 5 + [1] * 5
    assert list(pp.get_idx_base()) == list(range(5)) * 3
    assert p.get_x().size == 15 and p.get_steps().size == 15
    np.testing.assert_array_equal(p.get_px()[5:10], np.full(5, 1e-3))
    np.testing.assert_array_equal(p.get_y()[10:], x + 1e-2)
    assert p.get_displacement_module().shape == (5, 2)
    with pytest.raises(RuntimeError):
        p.add_ghost(1.0, "z")


def test_ghost_tracking_lyapunov_loop(hm):
    """test_new_standard.ipynb cell 22: track a few turns, accumulate the log of
    the ghost displacement, renormalise; the ghosts are tracked like the base
    particles (exact mode: bit for bit with the oracle)."""
    samples = 6
    v = np.linspace(0, 0.5, samples)
    xx, yy = (a.flatten() for a in np.meshgrid(v, v))
    z = np.zeros_like(xx)
    t = hm.henon_tracker(2000, 0.168, 0.201, "sps", epsilon=16.0)
    p = hm.particles(xx, z, yy, z)
    ot = O.Tracker(2000, 0.168, 0.201, "sps", epsilon=16.0)
    op = O.Particles(xx, z, yy, z)
    for d in ("x", "random"):
        p.add_ghost(1e-10, d)
        op.add_ghost(1e-10, d)
    lyap, olyap = np.zeros((xx.size, 2)), np.zeros((xx.size, 2))
    for _ in range(30):
        t.track(p, 10, 0.5, barrier=10.0)
        ot.track(op, 10, 0.5, barrier=10.0)
        lyap += np.log(p.get_displacement_module() / 1e-10)
        olyap += np.log(op.displacement_module() / 1e-10)
        p.renormalize(1e-10)
        op.renormalize(1e-10)
    assert np.array_equal(p.get_x(), op.x, equal_nan=True)
    assert np.array_equal(p.get_steps(), op.steps)
    assert np.array_equal(lyap, olyap, equal_nan=True)
    # a lost particle's ghosts report NaN displacements
    lost = ~np.isfinite(op.x[:xx.size])
    assert np.all(np.isnan(p.get_displacement_module()[lost]))
    assert np.all(np.isnan(p.get_displacement_direction()[:, lost]))
    # reset returns the ghosts to where they were created
    p.reset()
    op.reset()
    assert np.array_equal(p.get_x(), op.x0) and np.array_equal(p.get_py(), op.py0)


def test_ghost_added_after_tracking_uses_the_current_state(hm):
    x = np.linspace(0.05, 0.3, 7)
    z = np.zeros(7)
    t = hm.henon_tracker(500, 0.28, 0.31, "sps")
    p = hm.particles(x, z, x, z)
    t.track(p, 100, 0.0, barrier=10.0)
    p.add_ghost(1e-6, "py")
    ot = O.Tracker(500, 0.28, 0.31, "sps")
    op = O.Particles(x, z, x, z)
    ot.track(op, 100, 0.0, barrier=10.0)
    op.add_ghost(1e-6, "py")
    assert np.array_equal(p.get_py(), op.py, equal_nan=True)
    assert np.array_equal(p.get_steps(), op.steps)
    t.track(p, 50, 0.0, barrier=10.0)
    ot.track(op, 50, 0.0, barrier=10.0)
    assert np.array_equal(p.get_x(), op.x, equal_nan=True)
    assert np.array_equal(p.get_displacement_direction(), op.displacement_direction(), equal_nan=True)


def _grid(samples, lo, hi):
    v = np.linspace(lo, hi, samples)
    xx, yy = (a.flatten() for a in np.meshgrid(v, v))
    return xx, np.zeros_like(xx), yy, np.zeros_like(xx)


def test_alignment_indices_match_oracle(hm):
    """hm_ghost_alignment at a tracked state: SALI bit for bit with the
    reference harness's numpy formula, GALI (|det| here, SVD product there) to
    rounding; NaN for lost orbits."""
    x, px, y, py = _grid(12, 0.05, 0.5)
    t = hm.henon_tracker(3000, 0.168, 0.201, "sps", epsilon=32.0)
    p = hm.particles(x, px, y, py)
    ot = O.Tracker(3000, 0.168, 0.201, "sps", epsilon=32.0)
    op = O.Particles(x, px, y, py)
    for d in ("x", "y", "px", "py"):
        p.add_ghost(1e-8, d)
        op.add_ghost(1e-8, d)
    t.track(p, 1500, 0.5, barrier=10.0)
    ot.track(op, 1500, 0.5, barrier=10.0)
    s, g = p.particles.get_sali(), p.particles.get_gali()
    assert np.array_equal(s, O.sali(op), equal_nan=True)
    want = O.gali(op)
    assert np.array_equal(np.isnan(g), np.isnan(want))
    ok = ~np.isnan(want)
    np.testing.assert_allclose(g[ok], want[ok], rtol=1e-9, atol=1e-14)
    assert np.isnan(s).any() and (~np.isnan(s)).any()  # lost and surviving orbits both covered


def test_sali_gali_fli_functions(hm):
    from henon_map_cpp_amd import dynamic_indicators as DI

    x, px, y, py = _grid(8, 0.05, 0.45)
    t = hm.henon_tracker(5000, 0.168, 0.201, "sps", epsilon=32.0)
    ts, tau = [100, 1000, 2000], 10
    sali = DI.smallest_alignment_index(t, x, px, y, py, ts, tau, 1e-8, mu=0.5)
    gali = DI.global_alignment_index(t, x, px, y, py, ts, tau, 1e-8, mu=0.5)
    fli = DI.fast_lyapunov_indicator(t, x, px, y, py, ts, 1e-8, mu=0.5)
    assert list(sali.index) == ts and sali.shape == (3, x.size)
    # the running minima never increase
    for frame in (sali, gali):
        a = frame.to_numpy()
        fin = np.isfinite(a).all(axis=0)
        assert np.all(np.diff(a[:, fin], axis=0) <= 0)
    # the SALI loop, restated with the oracle: bit for bit
    ot = O.Tracker(5000, 0.168, 0.201, "sps", epsilon=32.0)
    op = O.Particles(x, px, y, py)
    op.add_ghost(1e-8, "x")
    op.add_ghost(1e-8, "y")
    acc, rows, j = np.full(x.size, np.sqrt(2.0)), [], 0
    for i in range(tau, ts[-1] + tau - 1, tau):
        ot.track(op, tau, 0.5, barrier=10.0)
        acc = np.minimum(acc, O.sali(op))
        while j < len(ts) and i >= ts[j]:
            rows.append(acc.copy())
            j += 1
    assert np.array_equal(sali.to_numpy(), np.array(rows), equal_nan=True)
    # FLI from the shadow particle: bit for bit with the oracle's ghost
    op = O.Particles(x, px, y, py)
    op.add_ghost(1e-8, "random")
    done, want = 0, []
    for tt in ts:
        ot.track(op, tt - done, 0.5, barrier=10.0)
        done = tt
        want.append(np.log10(op.displacement_module()[:, 0] / 1e-8) / tt)
    assert np.array_equal(fli.to_numpy(), np.array(want), equal_nan=True)
    # chaotic orbits: SALI and GALI collapse where the FLI grows
    last = np.isfinite(fli.to_numpy()[-1]) & np.isfinite(sali.to_numpy()[-1])
    chaotic = last & (fli.to_numpy()[-1] > 0.01)
    if chaotic.any():
        assert np.median(sali.to_numpy()[-1][chaotic]) < 1e-3
        assert np.median(gali.to_numpy()[-1][chaotic]) < 1e-6
