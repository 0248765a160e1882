"""The drop-in API (henon_map_cpp_amd, over the pybind11 henon_map_engine and
the C ABI) on the GPU: the reference notebooks' call sequences and printed
values, error behaviour, and parity with the oracle."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hm():
    import henon_map_cpp_amd as hm

    assert hm.gpu_available()
    return hm


def test_construct_notebook_sequence(hm, golden):
    # tests/test_construct.ipynb cells 4-10 and 27-30
    t = hm.henon_tracker(1000, 0.168, 0.201, "sps")
    pp = np.array([0.1, 0.1, 0.1])
    p = hm.particles(pp, np.zeros(3), pp, np.zeros(3))
    m1 = t.get_tangent_matrix(p, 0.0)
    for m in m1:
        np.testing.assert_allclose(m, golden["tangent_matrix_initial"]["value"], atol=0.6e-8)
    t.track(p, 1, 0.0)
    np.testing.assert_allclose(p.get_x(), golden["x_after_one_turn"]["value"], atol=0.6e-8)
    for m in t.get_tangent_matrix(p, 0.0):
        np.testing.assert_allclose(m, golden["tangent_matrix_after_one_turn"]["value"], atol=0.6e-8)
    np.testing.assert_allclose(hm.birkhoff_weights(10), golden["birkhoff_weights_10"]["value"], rtol=1e-8,
                               atol=1e-12)


@pytest.mark.parametrize("mu,kind,eps", [(0.0, "sps", 0.0), (0.5, "sps", 32.0), (-0.3, "basic", 5.0)])
def test_api_track_bit_exact(hm, mu, kind, eps):
    rng = np.random.default_rng(12)
    x, y = rng.uniform(0, 0.5, 3000), rng.uniform(0, 0.5, 3000)
    z = np.zeros(3000)
    omega_0 = 0.02 if kind == "basic" else np.nan
    t = hm.henon_tracker(4000, 0.28, 0.31, kind, omega_0=omega_0, epsilon=eps)
    p = hm.particles(x, z, y, z)
    t.track(p, 1500, mu, barrier=10.0)
    t.track(p, 1500, mu, barrier=10.0)
    ot = O.Tracker(4000, 0.28, 0.31, kind, omega_0=omega_0, epsilon=eps)
    op = O.Particles(x, z, y, z)
    ot.track(op, 3000, mu, barrier=10.0)
    assert np.array_equal(p.get_x(), op.x, equal_nan=True)
    assert np.array_equal(p.get_py(), op.py, equal_nan=True)
    assert np.array_equal(p.get_steps(), op.steps)
    # readback statistics: the reference's NaN-aware mean/std over GPU observables
    assert p.get_radius_mean() == pytest.approx(O.nan_mean(np.sqrt(op.x**2 + op.px**2 + op.y**2 + op.py**2)),
                                                rel=1e-13)
    assert p.get_action_x_std() == pytest.approx(O.nan_std(op.x**2 + op.px**2), rel=1e-12)


def test_turn_limits_and_errors(hm):
    t = hm.henon_tracker(100, 0.168, 0.201, "sps")
    p = hm.particles(np.array([0.1]), np.array([0.0]), np.array([0.1]), np.array([0.0]))
    with pytest.raises(RuntimeError, match="too large"):
        t.track(p, 101, 0.0)
    t.track(p, 60, 0.0)
    with pytest.raises(RuntimeError, match="too large"):
        t.track(p, 41, 0.0)
    with pytest.raises(RuntimeError, match="too large"):
        t.track(p, 61, 0.0, inverse=True)
    t.track(p, 60, 0.0, inverse=True)
    np.testing.assert_allclose(p.get_x(), [0.1], atol=1e-14)
    with pytest.raises(RuntimeError, match="Unknown modulation kind"):
        hm.henon_tracker(100, 0.1, 0.2, "triangle")
    with pytest.raises(NotImplementedError):  # the kick is taken by track only
        t.megno(p, 1, 0.0, kick_module=1e-3)
    with pytest.raises(TypeError):
        hm.particles([0.1], [0.0], [0.1], [0.0])


def test_reset_keeps_global_steps(hm):
    """henon.cu:727-735: reset() restores the coordinates and steps, not the
    tracker-side turn counter, so later tracks index the table from 0 again
    while the range check counts from the old value."""
    t = hm.henon_tracker(100, 0.168, 0.201, "sps", epsilon=16.0)
    x = np.array([0.1, 0.2])
    z = np.zeros(2)
    p = hm.particles(x, z, x, z)
    t.track(p, 70, 0.3)
    p.reset()
    assert np.all(p.get_steps() == 0) and np.array_equal(p.get_x(), x)
    with pytest.raises(RuntimeError, match="too large"):
        t.track(p, 31, 0.3)
    t.track(p, 30, 0.3)
    q = hm.particles(x, z, x, z)
    t.track(q, 30, 0.3)
    assert np.array_equal(p.get_x(), q.get_x(), equal_nan=True)


def test_fast_mode_through_api(hm):
    t = hm.henon_tracker(1000, 0.168, 0.201, "sps", epsilon=32.0)
    x = np.linspace(0.01, 0.1, 500)
    z = np.zeros(500)
    a, b = hm.particles(x, z, x, z), hm.particles(x, z, x, z)
    t.track(a, 1000, 0.5, mode="exact")
    t.track(b, 1000, 0.5, mode="fast")
    np.testing.assert_allclose(b.get_x(), a.get_x(), rtol=0, atol=1e-12)
    with pytest.raises(ValueError):
        t.track(b, 0, 0.5, mode="quick")


def test_construct_notebook_matrices_vectors_lyapunov(hm, golden):
    # tests/test_construct.ipynb cells 3-17, 23-31
    t = hm.henon_tracker(1000, 0.168, 0.201, "sps")
    pp = np.array([0.1, 0.1, 0.1])
    p = hm.particles(pp, np.zeros(3), pp, np.zeros(3))
    m, m1, m2 = hm.matrix_4d_vector(3), hm.matrix_4d_vector(3), hm.matrix_4d_vector(3)
    m.set_with_tracker(t, p, 0.0)
    m1.set_with_tracker(t, p, 0.0)
    for mm in m1.get_matrix():
        np.testing.assert_allclose(mm, golden["tangent_matrix_initial"]["value"], atol=0.6e-8)
    st = hm.storage_gpu(3, 2)
    st.store(p)
    t.track(p, 1, 0.0)
    st.store(p)
    assert st.get_x().shape == (2, 3)
    np.testing.assert_allclose(st.get_x()[1], golden["x_after_one_turn"]["value"], atol=0.6e-8)
    m2.set_with_tracker(t, p, 0.0)
    for mm in m2.get_matrix():
        np.testing.assert_allclose(mm, golden["tangent_matrix_after_one_turn"]["value"], atol=0.6e-8)
    # megno_construct.add(m2, m1): closed form from the reference formula
    mg = hm.megno_construct(3)
    mg.add(m2, m1)
    a, b = m2.get_matrix()[0], m1.get_matrix()[0]
    want = np.log10(np.sqrt(np.sum(a * a)) / np.sqrt(np.sum(b * b)))
    np.testing.assert_allclose(mg.get_values(), want, rtol=1e-13)
    m1.explicit_copy(m2)
    assert np.array_equal(m1.get_matrix(), m2.get_matrix())
    # vector_4d: v = J e0, ln|v| printed with all digits, normalised v
    g = golden["log_norm_J_e0"]
    v = hm.vector_4d(np.tile([1.0, 0.0, 0.0, 0.0], (3, 1)))
    v.multiply(m)
    np.testing.assert_allclose(v.get_vectors()[0], g["v"], atol=0.6e-8)
    assert np.log(np.linalg.norm(v.get_vectors()[0])) == pytest.approx(g["ln_norm"], abs=2e-16)
    v.normalize()
    np.testing.assert_allclose(v.get_vectors()[0], g["v_normalized"], atol=0.6e-8)
    # lyapunov_birkhoff_construct_multi(3, [5, 10, 15]) fed v = (1, 1, 0, 0), 3 x 5 adds
    gl = golden["lyapunov_multi"]
    v2 = hm.vector_4d(np.tile([1.0, 1.0, 0.0, 0.0], (3, 1)))
    c = hm.lyapunov_birkhoff_construct_multi(3, [5, 10, 15])
    for block in range(3):
        for _ in range(5):
            c.add(v2)
        np.testing.assert_allclose(c.get_values_raw()[:, 0], gl["raw"][block], atol=0.6e-8)
        np.testing.assert_allclose(c.get_values_b()[:, 0], gl["b"][block], atol=0.6e-8)
    c1 = hm.lyapunov_birkhoff_construct(3, 10)
    c1.change_weights(10)
    np.testing.assert_allclose(c1.get_weights(), golden["birkhoff_weights_10"]["value"], rtol=1e-8, atol=1e-12)


def test_tune_all_notebook(hm, golden):
    # tests/new_tune.ipynb cells 1-3 (two copies of the particle instead of 90000)
    e = golden["new_tune_all_tunes"]["exact"]
    pp = np.array([0.1, 0.1])
    p = hm.particles(pp, np.zeros_like(pp), pp, np.zeros_like(pp), force_CPU=True)
    t = hm.henon_tracker(10000, 0.168, 0.201, "sps", force_CPU=True)
    df = t.tune_all(p, 10000, 0.0, from_idx=np.array([0, 500]), to_idx=np.array([500, 1000]))
    assert list(df["from"]) == [0, 500, 0] and list(df["to"]) == [500, 1000, 10000]
    assert df["tune_x_birkhoff"][0][0] == pytest.approx(e["birkhoff_window_0_500_x"], abs=1e-14)
    assert df["tune_x_birkhoff"][1][0] == pytest.approx(e["birkhoff_window_500_1000_x"], abs=1e-14)
    # the last row is the full signal (samples 1..n for the FFT, all for Birkhoff)
    assert df["tune_x_fft"][2][0] == pytest.approx(e["fft_full_x"], abs=4e-15)
    assert df["tune_y_fft"][2][0] == pytest.approx(e["fft_full_y"], abs=4e-15)
    assert df["tune_x_birkhoff"][2][0] == pytest.approx(e["birkhoff_full_x"], abs=4e-15)
    # the appended (0, n_turns) window sits just before it in the engine's rows
    t2 = hm.henon_tracker(10000, 0.168, 0.201, "sps")
    p2 = hm.particles(pp, np.zeros_like(pp), pp, np.zeros_like(pp))
    rb, rf = t2.tracker.all_tunes(p2.particles, 10000, 0.0, 10.0, float("nan"), False, [], [])
    assert rf.shape == (2, 4)
    assert rf[0, 0] == pytest.approx(e["fft_window_0_10000_x"], abs=4e-15)
    assert rf[0, 1] == pytest.approx(e["fft_window_0_10000_y"], abs=4e-15)
    assert rb[0, 0] == pytest.approx(e["birkhoff_full_x"], abs=4e-15)


def test_api_megno_and_birkhoff_tunes_match_oracle(hm):
    rng = np.random.default_rng(41)
    x, y = rng.uniform(0, 0.4, 256), rng.uniform(0, 0.4, 256)
    z = np.zeros(256)
    t = hm.henon_tracker(800, 0.01, 0.012, "sps", epsilon=4.0)
    got = t.megno(hm.particles(x, z, y, z), 700, 0.3, turn_samples=np.array([10, 300, 699]),
                  carry_off_diagonal=False)
    ot = O.Tracker(800, 0.01, 0.012, "sps", epsilon=4.0)
    want = ot.megno(O.Particles(x, z, y, z), 700, 0.3, barrier=10.0, turn_samples=[10, 300, 699],
                    carry_off_diagonal=False)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14, equal_nan=True)
    t2 = hm.henon_tracker(800, 0.28, 0.31, "sps", epsilon=8.0)
    df = t2.tune_birkhoff(hm.particles(x, z, y, z), 800, 0.3, from_idx=np.array([0, 100]), to_idx=np.array([400, 800]))
    ot2 = O.Tracker(800, 0.28, 0.31, "sps", epsilon=8.0)
    want2 = ot2.birkhoff_tunes(O.Particles(x, z, y, z), 800, 0.3, barrier=10.0, from_idx=[0, 100], to_idx=[400, 800])
    np.testing.assert_allclose(np.stack(df["tune_x"].to_list()), want2[:, [0, 2, 4]].T, atol=2e-15, equal_nan=True)
    np.testing.assert_allclose(np.stack(df["tune_y"].to_list()), want2[:, [1, 3, 5]].T, atol=2e-15, equal_nan=True)


@pytest.mark.parametrize("n_threads,inverse", [(1, False), (7, False), (16, True), (300, False)])
def test_api_megno_per_thread_carry_matches_oracle(hm, n_threads, inverse):
    """henon.cu:2782-2803: the reference's threads take particles t,
    t + n_threads, ... and start each from the previous one's tangent matrix
    with only the diagonal reset, so results depend on n_threads and a lost
    particle turns the later ones of its thread NaN; with inverse=True the map
    steps backwards while J keeps the forward form at table index k. The GPU
    chains the per-particle products on the host (the running product up to
    rounding): equal to the oracle's literal restatement within 1e-10, NaN at
    the same places. Small tunes and 40 turns keep many traces positive (the
    trace form's logarithm is NaN otherwise)."""
    rng = np.random.default_rng(43)
    n = 257
    x, y = rng.uniform(0, 0.05, n), rng.uniform(0, 0.05, n)
    x[100] = y[100] = 3.0  # lost at once
    px = rng.uniform(-0.01, 0.01, n)
    z = np.zeros(n)
    t = hm.henon_tracker(1200, 0.01, 0.012, "sps", epsilon=4.0)
    p = hm.particles(x, px, y, z)
    ot = O.Tracker(1200, 0.01, 0.012, "sps", epsilon=4.0)
    op = O.Particles(x, px, y, z)
    if inverse:  # move forward first so that the inverse turns exist
        t.track(p, 40, 0.3)
        ot.track(op, 40, 0.3, barrier=10.0)
    samples = [5, 15, 39, 40]
    got = t.megno(p, 40, 0.3, inverse=inverse, turn_samples=np.array(samples), n_threads=n_threads)
    want = ot.megno(op, 40, 0.3, barrier=10.0, inverse=inverse, turn_samples=samples, n_threads=n_threads,
                    carry_off_diagonal=True)
    assert np.array_equal(np.isnan(got), np.isnan(want))
    assert np.isfinite(want).sum() > 100
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-12, equal_nan=True)
    assert np.array_equal(p.get_x(), op.x, equal_nan=True) and np.array_equal(p.get_steps(), op.steps)
    if n_threads < n and not inverse:
        ident = t.megno(hm.particles(x, px, y, z), 40, 0.3, turn_samples=np.array(samples), n_threads=n_threads,
                        carry_off_diagonal=False)
        assert not np.array_equal(ident, got, equal_nan=True)  # the carry changes the results


def test_get_tunes_and_storage(hm):
    t = hm.henon_tracker(3000, 0.31, 0.32, "sps", epsilon=16.0)
    x = np.array([0.05, 0.1, 0.2])
    z = np.zeros(3)
    p = hm.particles(x, z, x, z)
    st = hm.storage(3)
    st.store(p)
    for _ in range(2048):
        t.track(p, 1, 0.4)
        st.store(p)
    hx, hpx, hy, hpy = st.get_x(), st.get_px(), st.get_y(), st.get_py()
    assert hx.shape == (3, 2049)
    for i in range(3):
        np.testing.assert_allclose(hm.get_tunes(hx[i], hpx[i]), O.get_tunes(hx[i], hpx[i]), atol=1e-12)
    df = st.tune_fft(np.array([0, 1]), np.array([2047, 2048]), 2048)
    for i in range(3):
        want = O.fft_tune_vec(hx[i], hpx[i], hy[i], hpy[i], [0, 1], [2047, 2048])
        np.testing.assert_allclose([df["tune_x"][0][i], df["tune_y"][0][i], df["tune_x"][2][i]],
                                   [want[0], want[1], want[4]], atol=1e-12)
    db = st.tune_birkhoff(np.array([0]), np.array([1000]), 2048)
    for i in range(3):
        want = O.birkhoff_tune_vec(hx[i], hpx[i], hy[i], hpy[i], [0], [1000])
        np.testing.assert_allclose([db["tune_x"][0][i], db["tune_x"][1][i]], [want[0], want[2]], atol=2e-15)


def test_api_indicators_and_reversibility(hm):
    x = np.linspace(0.02, 0.4, 64)
    z = np.zeros(64)
    t = hm.henon_tracker(4000, 0.168, 0.201, "sps", epsilon=32.0)
    r = t.tangent_indicators(hm.particles(x, z, x, z), 3000, 0.5, renorm_every=100, samples=[100, 3000],
                             birkhoff=[(0, 3000)])
    ot = O.Tracker(4000, 0.168, 0.201, "sps", epsilon=32.0)
    w = O.indicators(O.Particles(x, z, x, z), ot.t, 3000, 0.5, 100.0, 100, [100, 3000], birkhoff=[(0, 3000)])
    np.testing.assert_allclose(r["megno"], w["megno"], rtol=1e-12, atol=1e-13, equal_nan=True)
    np.testing.assert_allclose(r["fli"], w["fli"], rtol=1e-12, atol=1e-13, equal_nan=True)
    np.testing.assert_allclose(r["birkhoff"], w["birkhoff"], atol=1e-13, equal_nan=True)
    assert np.array_equal(r["matrix"], w["matrix"], equal_nan=True)
    # reversibility: forward t then inverse t returns to the start within rounding
    p = hm.particles(x, z, x, z)
    re = t.reversibility_error(p, [10, 1000, 100], 0.5)
    assert re.shape == (3, 64)
    ok = np.isfinite(re[2])
    assert ok.sum() > 20 and np.all(re[0][np.isfinite(re[0])] < 1e-12)
    # the oracle's forward/backward pair for t = 1000
    q = O.Particles(x, z, x, z)
    ot.track(q, 1000, 0.5, barrier=100.0)
    ot.track(q, 1000, 0.5, barrier=100.0, inverse=True)
    d = np.sqrt((q.x - x) ** 2 + (q.px - z) ** 2 + (q.y - x) ** 2 + (q.py - z) ** 2)
    np.testing.assert_array_equal(re[2], d)


def test_engine_track_batch_and_autotune():
    """track_batch over trackers with different tables equals per-job track();
    autotune leaves the particles untouched."""
    import henon_map_cpp_amd as hm
    rng = np.random.default_rng(7)
    jobs, ref = [], []
    for k in range(6):
        n = int(rng.integers(10, 400))
        x = rng.uniform(0.05, 0.5, n)
        y = rng.uniform(0.05, 0.5, n)
        z = np.zeros(n)
        t = hm.henon_tracker(3000, 0.16 + 0.02 * k, 0.2, "sps", epsilon=8.0 * k)
        turns = int(rng.integers(100, 3000))
        mu = 0.0 if k == 0 else 0.3
        jobs.append((t, hm.particles(x, z, y, z), turns, mu, 10.0 ** 0.5))
        p = hm.particles(x, z, y, z)
        t.track(p, turns, mu, 10.0 ** 0.5)
        ref.append(p)
    hm.track_batch(jobs)
    for (t, p, *_), r in zip(jobs, ref):
        for g in ("get_x", "get_px", "get_y", "get_py", "get_steps"):
            assert np.array_equal(getattr(p, g)(), getattr(r, g)(), equal_nan=True)
    t, p = jobs[1][0], hm.particles(np.linspace(0.01, 0.4, 5000), np.zeros(5000), np.linspace(0.01, 0.4, 5000),
                                    np.zeros(5000))
    best = t.autotune(p, 0.3, warm_turns=200, trial_turns=400)
    assert best["block"] in (64, 256) and best["trial_ms"] > 0
    assert np.all(p.get_steps() == 0)
    assert np.array_equal(p.get_x(), np.linspace(0.01, 0.4, 5000))
    hm._capi.check(hm._capi.lib().hm_autotune_clear())


def test_particle_arrays_in_and_out(hm):
    """Constructors take numpy arrays (any dtype, converted to float64) and
    lists alike; getters return numpy arrays of the dtypes the reference's
    `py::cast(vector)` gives (src/bind.cpp:256-275): float64, and int64 for the
    integer vectors; a coordinate array that is not one-dimensional is a
    TypeError."""
    from henon_map_cpp_amd import henon_map_engine as E

    x = np.linspace(0.01, 0.2, 7)
    z = np.zeros(7)
    a = E.particles_4d_gpu(x, z, x, z)
    b = E.particles_4d_gpu(list(x), list(z), list(x), list(z))
    c = E.particles_4d_gpu(x.astype(np.float32), z, x, z)
    assert a.get_x().dtype == np.float64 and a.get_steps().dtype == np.int64 and a.get_valid().dtype == np.int64
    assert np.array_equal(a.get_x(), x) and np.array_equal(b.get_x(), x) and np.array_equal(b.get_py(), z)
    assert np.array_equal(c.get_x(), x.astype(np.float32).astype(np.float64))
    assert np.array_equal(a.get_steps(), np.zeros(7)) and np.array_equal(a.get_valid(), np.ones(7))
    with pytest.raises(TypeError):
        E.particles_4d_gpu(x.reshape(7, 1), z, x, z)
    with pytest.raises(ValueError):
        E.particles_4d_gpu(x, z[:3], x, z)


def test_vector_arrays_in_and_out(hm):
    """vector_4d_gpu takes an (n, 4) array or a list of 4-vectors; set_vectors
    takes either, or one 4-vector for every particle (src/bind.cpp:411-420)."""
    from henon_map_cpp_amd import henon_map_engine as E

    v = np.arange(12, dtype=np.float64).reshape(3, 4)
    a = E.vector_4d_gpu(v)
    b = E.vector_4d_gpu([list(r) for r in v])
    assert np.array_equal(a.get_vectors(), v) and np.array_equal(b.get_vectors(), v)
    a.set_vectors([1.0, 2.0, 3.0, 4.0])
    assert np.array_equal(a.get_vectors(), np.tile([1.0, 2.0, 3.0, 4.0], (3, 1)))
    a.set_vectors(v[::-1].copy())
    assert np.array_equal(a.get_vectors(), v[::-1])
    with pytest.raises(ValueError):
        a.set_vectors(np.zeros((2, 4)))
    with pytest.raises(ValueError):
        a.set_vectors(np.zeros((3, 3)))
