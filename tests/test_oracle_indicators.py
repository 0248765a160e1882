"""Pins the oracle's fused indicators (an algorithm with no reference
counterpart) to the reference's per-turn construct formulas, to closed forms,
and to renormalisation invariance."""
import numpy as np
import pytest

import oracle as O


def construct_loop(x0, tables, n_turns, mu, samples, barrier=10.0):
    """The reference's construct sequence, one turn at a time: J at the
    pre-step point (henon.cu:198-222), M <- J M (multiply_matrices,
    henon.cu:251-265), megno_construct.add(M_new, M_old) (henon.cu:2253-2284),
    then one map step (henon.cu:420-480)."""
    n = len(x0[0])
    p = O.Particles(*x0)
    t = O.Tracker.from_tables(tables)
    M = np.tile(np.eye(4), (n, 1, 1))
    l1, l2, l3 = np.zeros(n), np.zeros(n), np.zeros(n)
    out = {}
    for k in range(n_turns):
        J = t.get_tangent_matrix(p, mu)  # table index global_steps = k
        Mn = np.stack([O.multiply_matrices(J[i], M[i]) for i in range(n)])
        O.megno_construct_add(l1, l2, l3, Mn.reshape(n, 16), M.reshape(n, 16), k + 1)
        M = Mn
        t.track(p, 1, mu, barrier=barrier)
        if k + 1 in samples:
            out[k + 1] = l3.copy()
    return out, M, p


def test_fused_megno_equals_construct_formulas():
    tables = O.Tables(O.sps_modulation(0.28, 16.0, 0, 400), O.sps_modulation(0.31, 16.0, 0, 400))
    rng = np.random.default_rng(2)
    x0 = (rng.uniform(0, 0.3, 6), rng.uniform(-0.05, 0.05, 6), rng.uniform(0, 0.3, 6), np.zeros(6))
    samples = [1, 7, 100, 399]
    ref, M, pr = construct_loop(x0, tables, 399, 0.4, samples)
    p = O.Particles(*x0)
    r = O.indicators(p, tables, 399, 0.4, samples=samples)
    for q, s in enumerate(samples):
        np.testing.assert_array_equal(r["megno"][q], ref[s])
    np.testing.assert_array_equal(r["matrix"], M)
    assert np.array_equal(p.x, pr.x, equal_nan=True) and np.array_equal(p.steps, pr.steps)


def test_rotation_closed_forms():
    # At the origin (mu=0) J is the rotation R(wx) x R(wy): |M|_F = 2 for every
    # turn, so MEGNO and FLI vanish; Birkhoff recovers nothing (no phase).
    tables = O.Tables(O.sps_modulation(0.168, 0.0, 0, 500), O.sps_modulation(0.201, 0.0, 0, 500))
    p = O.Particles([0.0], [0.0], [0.0], [0.0])
    r = O.indicators(p, tables, 500, 0.0, samples=[500], fli_dir=(0.3, 0.1, -0.2, 0.5))
    assert abs(r["megno"][0, 0]) < 1e-14 and abs(r["fli"][0, 0]) < 1e-14
    # tiny amplitude: Birkhoff tunes are the linear tunes
    p = O.Particles([1e-9], [0.0], [1e-9], [0.0])
    r = O.indicators(p, tables, 500, 0.0, birkhoff=[(0, 500), (100, 300)])
    np.testing.assert_allclose(r["birkhoff"][:, 0, 0], [0.168, 0.168], atol=1e-12)
    np.testing.assert_allclose(r["birkhoff"][:, 1, 0], [0.201, 0.201], atol=1e-12)


def test_birkhoff_window_equals_reference_tune_of_history():
    tables = O.Tables(O.sps_modulation(0.168, 32.0, 0, 600), O.sps_modulation(0.201, 32.0, 0, 600))
    x0 = ([0.11, 0.05], [0.0, 0.01], [0.09, 0.07], [0.0, 0.0])
    # history the way henon_tracker::birkhoff_tunes stores it (henon.cu:2926-2943)
    p = O.Particles(*x0)
    t = O.Tracker.from_tables(tables)
    h = [np.stack([p.x, p.px, p.y, p.py])]
    for _ in range(600):
        t.track(p, 1, 0.5, barrier=10.0)
        h.append(np.stack([p.x, p.px, p.y, p.py]))
    h = np.array(h)  # (601, 4, n)
    q = O.Particles(*x0)
    r = O.indicators(q, tables, 600, 0.5, birkhoff=[(0, 600), (37, 415)])
    for i in range(2):
        want = O.birkhoff_tune_vec(h[:, 0, i], h[:, 1, i], h[:, 2, i], h[:, 3, i], [37], [415])
        np.testing.assert_allclose(r["birkhoff"][1, :, i], want[0:2], rtol=0, atol=1e-15)
        np.testing.assert_allclose(r["birkhoff"][0, :, i], want[2:4], rtol=0, atol=1e-15)


def test_renormalisation_invariance():
    tables = O.Tables(O.sps_modulation(0.31, 0.0, 0, 3000), O.sps_modulation(0.32, 0.0, 0, 3000))
    x0 = ([0.05, 0.2, 0.35], [0.0] * 3, [0.05, 0.2, 0.3], [0.0] * 3)
    samples = [1000, 2000, 3000]
    a = O.indicators(O.Particles(*x0), tables, 3000, 0.0, samples=samples)
    b = O.indicators(O.Particles(*x0), tables, 3000, 0.0, samples=samples, renorm_every=100)
    ok = np.isfinite(a["megno"])
    np.testing.assert_allclose(b["megno"][ok], a["megno"][ok], rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(b["fli"][ok], a["fli"][ok], rtol=1e-9, atol=1e-12)


def test_fli_matches_finite_difference_growth():
    # FLI(t) = log10 |M_t v| / t: M_t v is the linearised growth of a small
    # displacement along v, checked against two oracle trajectories.
    tables = O.Tables(O.sps_modulation(0.28, 8.0, 0, 200), O.sps_modulation(0.31, 8.0, 0, 200))
    v = np.array([0.6, 0.0, 0.8, 0.0])
    x0 = np.array([0.2, 0.01, 0.15, 0.0])
    r = O.indicators(O.Particles(*[[c] for c in x0]), tables, 200, 0.3, samples=[200], fli_dir=v)
    h = 1e-7
    a = O.Particles(*[[c] for c in x0 + h * v])
    b = O.Particles(*[[c] for c in x0 - h * v])
    t = O.Tracker.from_tables(tables)
    t.track(a, 200, 0.3, barrier=10.0)
    t.track(b, 200, 0.3, barrier=10.0)
    d = np.array([a.x[0] - b.x[0], a.px[0] - b.px[0], a.y[0] - b.y[0], a.py[0] - b.py[0]]) / (2 * h)
    assert r["fli"][0, 0] == pytest.approx(np.log10(np.linalg.norm(d)) / 200, abs=1e-6)


def test_megno_thread_carry_properties():
    """The reference CPU megno (henon.cu:2782-2803) keeps one tangent matrix per
    thread and resets only its diagonal between the thread's particles: with a
    thread per particle nothing is carried (equal to the identity start); with
    one thread a lost particle's NaN matrix makes every later particle NaN, and
    a chain after a surviving particle differs from the identity start. (Small
    tunes and few turns keep the traces of the trace form positive, so its
    logarithms are finite.)"""
    tables = O.Tables(O.sps_modulation(0.01, 4.0, 0, 400), O.sps_modulation(0.012, 4.0, 0, 400))
    x0 = ([0.02, 1.3, 0.03, 0.04], [0.0] * 4, [0.02, 1.3, 0.025, 0.01], [0.0] * 4)  # particle 1 is lost
    t = O.Tracker.from_tables(tables)

    def run(nt, carry):
        return t.megno(O.Particles(*x0), 20, 0.3, barrier=2.0, turn_samples=[5, 15], n_threads=nt,
                       carry_off_diagonal=carry)

    ident = run(4, False)
    np.testing.assert_array_equal(run(4, True), ident)
    assert np.isnan(ident[1]).all() and np.isfinite(ident[[0, 2, 3]]).all()
    one = run(1, True)
    np.testing.assert_array_equal(one[0], ident[0])
    assert np.isnan(one[1:]).all()
    two = run(2, True)  # chains {0, 2} and {1, 3}
    np.testing.assert_array_equal(two[:2], ident[:2])
    assert np.isfinite(two[2, 0]) and abs(two[2, 0] - ident[2, 0]) > 1e-3
    assert np.isnan(two[3]).all()
