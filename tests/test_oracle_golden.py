"""Pins the oracle (CPU restatement of the reference path) before anything is
checked against it: the reference notebooks' printed values, the reference's own
modulation.cpp (oracle/_ref), and what the mathematics fixes."""
import numpy as np
import pytest

import oracle as O


def test_tangent_matrix_goldens(golden):
    t = O.Tracker(1000, 0.168, 0.201, "sps")
    p = O.Particles([0.1] * 3, [0.0] * 3, [0.1] * 3, [0.0] * 3)
    g = golden["tangent_matrix_initial"]
    for m in t.get_tangent_matrix(p, 0.0):
        np.testing.assert_allclose(m, g["value"], atol=0.6e-8)
    t.track(p, 1, 0.0, barrier=10.0)
    np.testing.assert_allclose(p.x, golden["x_after_one_turn"]["value"], atol=0.6e-8)
    for m in t.get_tangent_matrix(p, 0.0):
        np.testing.assert_allclose(m, golden["tangent_matrix_after_one_turn"]["value"], atol=0.6e-8)


def test_log_norm_golden(golden):
    t = O.Tracker(1000, 0.168, 0.201, "sps")
    p = O.Particles([0.1], [0.0], [0.1], [0.0])
    J = t.get_tangent_matrix(p, 0.0)[0]
    v = O.matvec(J.reshape(1, 16), np.array([[1.0, 0, 0, 0]]))
    g = golden["log_norm_J_e0"]
    np.testing.assert_allclose(v[0], g["v"], atol=0.6e-8)
    # printed with all 17 significant digits by the notebook
    assert np.log(np.sqrt(np.sum(v[0] * v[0]))) == pytest.approx(g["ln_norm"], abs=1e-17)
    np.testing.assert_allclose(O.normalize(v)[0], g["v_normalized"], atol=0.6e-8)


def test_birkhoff_weights_golden(golden):
    g = golden["birkhoff_weights_10"]
    np.testing.assert_allclose(O.birkhoff_weights(10), g["value"], rtol=1e-8, atol=1e-12)
    w = O.birkhoff_weights(1000)
    assert w.sum() == pytest.approx(1.0, abs=1e-14)
    # the bump t -> exp(-1/(t(1-t))) is symmetric about t = 1/2: w[i] = w[n-i]
    np.testing.assert_allclose(w[1:], w[1:][::-1], rtol=1e-12)


def test_lyapunov_multi_golden(golden):
    # lyapunov_birkhoff_construct_multi(3, [5, 10, 15]) fed v = (1, 1, 0, 0)
    g = golden["lyapunov_multi"]
    v = np.tile([1.0, 1.0, 0.0, 0.0], (3, 1))
    nws = [5, 10, 15]
    vals = [np.zeros(3) for _ in nws]
    vals_b = [np.zeros(3) for _ in nws]
    ws = [O.birkhoff_weights(n) for n in nws]
    idx = 0
    for block in range(3):
        for _ in range(5):
            for i, n in enumerate(nws):
                if idx < n:  # henon.cu:2173-2184
                    O.lyap_birk_add(vals[i], vals_b[i], v, ws[i], n, idx)
            idx += 1
        np.testing.assert_allclose([a[0] for a in vals], g["raw"][block], atol=0.6e-8)
        np.testing.assert_allclose([a[0] for a in vals_b], g["b"][block], atol=0.6e-8)


@pytest.fixture(scope="module")
def new_tune_run():
    t = O.Tracker(10000, 0.168, 0.201, "sps")
    p = O.Particles([0.1, 0.1], [0.0, 0.0], [0.1, 0.1], [0.0, 0.0])
    ob, of = t.all_tunes(p, 10000, 0.0, barrier=10.0, from_idx=[0, 500], to_idx=[500, 1000])
    return ob, of


def test_all_tunes_golden(golden, new_tune_run):
    """Seven numbers printed by tests/new_tune.ipynb, reproduced to the last digit."""
    ob, of = new_tune_run
    e = golden["new_tune_all_tunes"]["exact"]
    # layout: windows (0,500), (500,1000), appended (0,10000), then full signal
    assert of[0, 6] == e["fft_full_x"]
    assert of[0, 7] == e["fft_full_y"]
    assert of[0, 4] == e["fft_window_0_10000_x"]
    assert of[0, 5] == e["fft_window_0_10000_y"]
    assert ob[0, 0] == e["birkhoff_window_0_500_x"]
    assert ob[0, 2] == e["birkhoff_window_500_1000_x"]
    assert ob[0, 6] == e["birkhoff_full_x"]
    assert np.array_equal(ob[0], ob[1]) and np.array_equal(of[0], of[1])


def test_modulation_matches_reference_source():
    """oracle/_ref is the reference's own src/modulation.cpp, compiled from source."""
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    import ctypes

    for tune, eps, start, end in [(0.168, 0.0, 0, 500), (0.31, 32.0, 0, 2000), (0.201, 64.0, 1234, 3000)]:
        r = np.empty(end - start)
        R.ref_sps_modulation(tune, eps, start, end, r.ctypes.data_as(ctypes.c_void_p))
        assert np.array_equal(O.sps_modulation(tune, eps, start, end), r)
        R.ref_basic_modulation(tune, 0.01, eps, start, end, r.ctypes.data_as(ctypes.c_void_p))
        assert np.array_equal(O.basic_modulation(tune, 0.01, eps, start, end), r)
    c, f = O.sps_lines()
    assert np.array_equal(O.lines_modulation(0.168, 32.0, c, f, 5, 900), O.sps_modulation(0.168, 32.0, 5, 900))


def test_unmodulated_table_is_exact_rotation():
    # epsilon = 0: omega = 2 pi tune exactly (used for the unmodulated DA scan)
    w = O.sps_modulation(0.168, 0.0, 0, 100)
    assert np.all(w == 2 * np.pi * 0.168)


def _fd_jacobian(t, x0, mu, h=1e-6):
    J = np.zeros((4, 4))
    for c in range(4):
        cols = []
        for sgn in (+1, -1):
            z = np.array(x0, dtype=float)
            z[c] += sgn * h
            p = O.Particles([z[0]], [z[1]], [z[2]], [z[3]])
            t.track(p, 1, mu, barrier=1e3)
            cols.append(np.array([p.x[0], p.px[0], p.y[0], p.py[0]]))
        J[:, c] = (cols[0] - cols[1]) / (2 * h)
    return J


@pytest.mark.parametrize("mu", [0.0, 0.5, -1.0])
def test_tangent_matrix_is_jacobian_and_symplectic(mu):
    t = O.Tracker(50, 0.168, 0.201, "sps", epsilon=32.0)
    rng = np.random.default_rng(3)
    for _ in range(5):
        z = rng.uniform(-0.3, 0.3, 4)
        J = O.tangent_matrix(*z, t.t.sx[0], t.t.cx[0], t.t.sy[0], t.t.cy[0], mu)
        np.testing.assert_allclose(J, _fd_jacobian(t, z, mu), atol=1e-8)
        assert np.linalg.det(J) == pytest.approx(1.0, abs=1e-12)
        Om = np.array([[0, 1, 0, 0], [-1, 0, 0, 0], [0, 0, 0, 1], [0, 0, -1, 0]], dtype=float)
        np.testing.assert_allclose(J.T @ Om @ J, Om, atol=1e-12)


@pytest.mark.parametrize("mu", [0.0, 0.5])
def test_inverse_tangent_matrix(mu):
    t = O.Tracker(50, 0.31, 0.32, "sps", epsilon=16.0)
    rng = np.random.default_rng(5)
    z = rng.uniform(-0.2, 0.2, 4)
    p = O.Particles([z[0]], [z[1]], [z[2]], [z[3]])
    J = O.tangent_matrix(*z, t.t.sx[0], t.t.cx[0], t.t.sy[0], t.t.cy[0], mu)
    t.track(p, 1, mu, barrier=10.0)
    Ji = O.inverse_tangent_matrix(p.x[0], p.px[0], p.y[0], p.py[0], t.t.sx[0], t.t.cx[0], t.t.sy[0], t.t.cy[0], mu)
    np.testing.assert_allclose(Ji @ J, np.eye(4), atol=1e-12)


def test_track_inverse_roundtrip_and_resume():
    t = O.Tracker(3000, 0.28, 0.31, "sps", epsilon=16.0)
    rng = np.random.default_rng(1)
    x = rng.uniform(0, 0.15, 64)
    y = rng.uniform(0, 0.15, 64)
    p = O.Particles(x, np.zeros(64), y, np.zeros(64))
    t.track(p, 1000, 0.3, barrier=10.0)
    t.track(p, 1000, 0.3, barrier=10.0)
    q = O.Particles(x, np.zeros(64), y, np.zeros(64))
    t.track(q, 2000, 0.3, barrier=10.0)
    assert np.array_equal(p.x, q.x, equal_nan=True) and np.array_equal(p.steps, q.steps)
    t.track(q, 2000, 0.3, barrier=10.0, inverse=True)
    ok = q.steps == 0
    assert ok.sum() > 10
    np.testing.assert_allclose(q.x[ok], x[ok], atol=1e-10)
    np.testing.assert_allclose(q.y[ok], y[ok], atol=1e-10)
    with pytest.raises(RuntimeError):
        t.track(q, 1, 0.3, inverse=True)
    with pytest.raises(RuntimeError):
        t.track(q, 3001, 0.3)


def test_nan_and_barrier_semantics():
    t = O.Tracker(100, 0.168, 0.201, "sps")
    p = O.Particles([np.nan, 0.1, 5.0, 0.0], [0, 0, 0, 0], [0.1, 0.1, 0.0, 0.0], [0, 0, 0, 0])
    t.track(p, 100, 0.0, barrier=1.0)
    assert np.isnan(p.x[0]) and p.steps[0] == 0          # NaN in: never stepped
    assert p.steps[1] == 100 and not np.isnan(p.x[1])    # inside
    assert np.isnan(p.x[2]) and p.steps[2] == 0          # lost on the first turn
    assert p.steps[3] == 100 and p.x[3] == 0.0           # fixed point at the origin


def test_linear_limit_tunes():
    # tiny amplitude, no modulation: z -> e^{-i w} z, tune = w / 2 pi
    t = O.Tracker(4096, 0.168, 0.201, "sps")
    p = O.Particles([1e-9], [0.0], [1e-9], [0.0])
    h = [(p.x[0], p.px[0], p.y[0], p.py[0])]
    for _ in range(2048):
        t.track(p, 1, 0.0, barrier=10.0)
        h.append((p.x[0], p.px[0], p.y[0], p.py[0]))
    h = np.array(h)
    assert O.birkhoff_tune(h[:, 0], h[:, 1]) == pytest.approx(0.168, abs=1e-12)
    assert O.birkhoff_tune(h[:, 2], h[:, 3]) == pytest.approx(0.201, abs=1e-12)
    assert O.fft_tune(h[1:, 0], h[1:, 1]) == pytest.approx(0.168, abs=1e-9)
    assert O.fft_tune(h[1:, 2], h[1:, 3]) == pytest.approx(0.201, abs=1e-9)


def test_dft_matches_numpy_fft():
    rng = np.random.default_rng(7)
    for n in (7, 64, 501, 2048):
        re, im = rng.normal(size=n), rng.normal(size=n)
        h = 2 * np.sin(np.pi * np.arange(n) / n) ** 2
        ref = np.abs(np.fft.fft((re + 1j * im) * h))
        np.testing.assert_allclose(O.fft_magnitude(re, im), ref, rtol=1e-12, atol=1e-12 * ref.max())


def test_interpolation_edges():
    assert O.interpolation(np.array([3.0, 1.0, 0.5])) == 1.0
    assert O.interpolation(np.array([0.5, 1.0, 3.0])) == 0.0


def test_cpu_megno_linear_closed_form():
    """At the origin with mu=0 the tangent matrix is the constant rotation R and
    tr M_k = 2 cos(k wx) + 2 cos(k wy); the trace-form MEGNO follows from it."""
    wx, wy = 2 * np.pi * 0.168, 2 * np.pi * 0.201
    n = 200
    t = O.Tracker(n, 0.168, 0.201, "sps")
    p = O.Particles([0.0], [0.0], [0.0], [0.0])
    samples = [10, 50, 100, 199]
    got = t.megno(p, n, 0.0, turn_samples=samples, n_threads=1)[0]
    k = np.arange(0, n + 1)
    tr = 2 * np.cos(k * wx) + 2 * np.cos(k * wy)
    with np.errstate(invalid="ignore", divide="ignore"):
        a = np.arange(1, n + 1) * np.log10(tr[1:] / tr[:-1])
    s1 = np.cumsum(a)
    b = s1 / np.arange(1, n + 1)
    s2 = np.cumsum(b)
    want = np.array([s2[s] / s for s in samples])
    np.testing.assert_allclose(got, want, rtol=1e-9, equal_nan=True)
    # default turn_samples = {n_turns} reads a slot that is never written (0)
    p2 = O.Particles([0.0], [0.0], [0.0], [0.0])
    assert t.megno(p2, n, 0.0, n_threads=1)[0, 0] == 0.0


def test_cpu_megno_thread_carry_quirk():
    """henon.cu:2794-2803 resets only the diagonal between the particles one
    thread owns: the second particle of a thread starts from a non-identity."""
    # slow tunes keep the trace positive (finite log10) over the sampled turns
    t = O.Tracker(100, 0.01, 0.012, "sps")
    x = [0.1, 0.2]
    a = t.megno(O.Particles(x, [0, 0], x, [0, 0]), 12, 0.0, turn_samples=[10], n_threads=1)
    b = t.megno(O.Particles(x, [0, 0], x, [0, 0]), 12, 0.0, turn_samples=[10], n_threads=2)
    c = t.megno(O.Particles(x, [0, 0], x, [0, 0]), 12, 0.0, turn_samples=[10], n_threads=1,
                carry_off_diagonal=False)
    assert np.all(np.isfinite(a)) and np.all(np.isfinite(c))
    assert np.array_equal(a[0], b[0], equal_nan=True)
    assert np.array_equal(b[1], c[1], equal_nan=True)
    assert not np.array_equal(a[1], c[1], equal_nan=True)


def test_ghost_golden(golden):
    """Shadow particles (henon.cu:737-952) against test_new_standard.ipynb: the
    displaced coordinates, the random directions of the default-seeded particle
    generator, unit modules after renormalize(1)."""
    g = golden["ghost_3x5"]
    p = O.Particles(np.ones(3), np.ones(3), np.ones(3), np.ones(3))
    for d in ("x", "px", "y", "py", "random"):
        p.add_ghost(1e-4, d)
    assert p.n == 18 and p.n_ghosts == 5
    np.testing.assert_allclose(p.x, g["x"], rtol=0, atol=0.6 * 10.0 ** -g["x_digits"])
    direction = p.displacement_direction()
    np.testing.assert_allclose(direction[:, :, 4], g["random_direction"], rtol=0,
                               atol=0.6 * 10.0 ** -g["direction_digits"])
    # axis ghosts point along their axis; every ghost sits 1e-4 from its base
    np.testing.assert_allclose(direction[:, :, :4], np.broadcast_to(np.eye(4)[:, None, :], (4, 3, 4)), atol=1e-11)
    np.testing.assert_allclose(p.displacement_module(), g["module"], rtol=1e-11)
    np.testing.assert_allclose(np.sum(direction ** 2, axis=0), 1.0, rtol=1e-12)
    p.renormalize(1.0)
    np.testing.assert_allclose(p.displacement_module(), g["module_after_renormalize_1"], rtol=1e-14)
    np.testing.assert_allclose(p.displacement_direction(), direction, atol=1e-11)
    # the base particles never move
    np.testing.assert_array_equal(p.x[:3], np.ones(3))


def test_ghost_random_kicks_continue_the_generator():
    """A second random set continues the generator: it differs from the first,
    and a fresh particle set replays the same sequence."""
    a = O.Particles(np.zeros(4), np.zeros(4), np.zeros(4), np.zeros(4))
    a.add_ghost(1.0, "random")
    a.add_ghost(1.0, "random")
    b = O.Particles(np.zeros(8), np.zeros(8), np.zeros(8), np.zeros(8))
    b.add_ghost(1.0, "random")
    np.testing.assert_array_equal(a.x[4:], b.x[8:])
    assert not np.array_equal(a.x[4:8], a.x[8:12])
    with pytest.raises(RuntimeError):
        a.add_ghost(1.0, "z")


def test_alignment_indices_closed_forms():
    """SALI and GALI as the reference's harness computes them
    (dynamic_indicators.py:154-366) on an unmodulated, nearly linear orbit: the
    x- and y-plane deviations stay orthogonal (SALI = sqrt 2) and the normalised
    deviation matrix stays a rotation (GALI = 1); on a chaotic orbit both
    collapse. GALI is cross-checked against |det|."""
    t = O.Tracker(4000, 0.168, 0.201, "sps", epsilon=0.0)
    p = O.Particles([1e-6], [0.0], [1e-6], [0.0])
    for d in ("x", "y", "px", "py"):
        p.add_ghost(1e-9, d)
    t.track(p, 1000, 0.0, barrier=10.0)
    assert abs(O.sali(p)[0] - np.sqrt(2.0)) < 1e-5
    assert abs(O.gali(p)[0] - 1.0) < 1e-5
    c = O.Particles([0.3], [0.0], [0.2], [0.0])
    for d in ("x", "y", "px", "py"):
        c.add_ghost(1e-9, d)
    O.Tracker(4000, 0.168, 0.201, "sps", epsilon=32.0).track(c, 2000, 0.5, barrier=10.0)
    assert np.all(c.steps == 2000)
    dev = [O._deviations(c, k) for k in range(4)]
    m = np.array([[dev[k][r][0] for k in range(4)] for r in range(4)])
    assert O.gali(c)[0] == pytest.approx(abs(np.linalg.det(m)), abs=1e-14)
    assert O.gali(c)[0] < 1e-9 and O.sali(c)[0] < 0.1


def test_philox_known_answers():
    """The kick generator: Philox-4x32-10's published known-answer vectors
    (Random123 distribution, kat_vectors: philox4x32_10)."""
    cases = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
             ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
             ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
              (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, want in cases:
        assert tuple(int(v) for v in O.philox4x32_10(ctr, key)) == want


def test_kick_normals_and_kicked_step():
    """Box-Muller normals from the generator: standard normal moments; the kick
    moves a point by exactly kick_module before the map (inverse map undoes the
    map, leaving the kick)."""
    z = np.array([O.kick_normals(7, i, d) for i in range(500) for d in range(40)]).ravel()
    assert abs(z.mean()) < 0.02 and abs(z.var() - 1.0) < 0.03 and abs((z ** 4).mean() - 3.0) < 0.15
    t = O.Tracker(100, 0.168, 0.201, "sps", epsilon=8.0)
    x0 = np.array([0.1, 0.2, 0.05])
    p = O.Particles(x0, x0 / 2, x0, -x0)
    t.track_kicked(p, 1, 0.3, 10.0, 1e-3, seed=11, draw=0)
    t.track(p, 1, 0.3, barrier=10.0, inverse=True)
    d = np.stack([p.x - x0, p.px - x0 / 2, p.y - x0, p.py + x0])
    np.testing.assert_allclose(np.sqrt((d ** 2).sum(axis=0)), 1e-3, rtol=1e-10)
    n = np.array([O.kick_normals(11, i, 0) for i in range(3)]).T
    np.testing.assert_allclose(d / 1e-3, n / np.sqrt((n ** 2).sum(axis=0)), atol=1e-10)
