"""Multi-window accumulators of the fused indicator kernel through the drop-in
API: the construct-form turn-sample windows (megno_birkhoff_construct,
tune_birkhoff_construct; src/henon.cu:2302-2655) summed inside the tracking
kernel must equal the construct classes fed turn by turn (which the closed-form
tests pin) and the oracle's constructs; more samples and windows than one
launch takes run as further passes with the same results."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hm():
    import henon_map_cpp_amd as hm

    assert hm.gpu_available()
    return hm


def start(n, seed, hi=0.4):
    rng = np.random.default_rng(seed)
    return [rng.uniform(0.02, hi, n), rng.uniform(-0.03, 0.03, n), rng.uniform(0.02, hi, n), np.zeros(n)]


def test_fused_turn_sample_windows_equal_constructs(hm):
    n, T, mu = 64, 400, 0.5
    mb_samples, tb_samples = [1, 50, 128, 450], [16, 100, 200, 333]
    x0 = start(n, 51, hi=0.6)  # 11 of the 64 are lost within the 400 turns
    t = hm.henon_tracker(T + 1, 0.168, 0.201, "sps", epsilon=32.0)
    fused = t.tangent_indicators(hm.particles(*x0), T, mu, renorm_every=0, megno_birkhoff=mb_samples,
                                 tune_birkhoff=tb_samples)
    assert fused["megno_birkhoff"].shape == (4, n) and fused["tune_birkhoff"].shape == (4, 4, n)
    # the construct classes, one add per turn (GPU)
    p = hm.particles(*x0)
    m_old, m_new = hm.matrix_4d_vector(n), hm.matrix_4d_vector(n)
    mb = hm.megno_birkhoff_construct(n, mb_samples)
    tb = hm.tune_birkhoff_construct(n, tb_samples)
    t2 = hm.henon_tracker(T + 1, 0.168, 0.201, "sps", epsilon=32.0)
    tb.first_add(p)
    for _ in range(T):
        m_new.explicit_copy(m_old)
        m_new.structured_multiply(t2, p, mu)
        mb.add(m_new, m_old)
        m_old.explicit_copy(m_new)
        t2.track(p, 1, mu)
        tb.add(p)
    np.testing.assert_allclose(fused["megno_birkhoff"], mb.get_values(), rtol=1e-12, atol=1e-14, equal_nan=True)
    got_t = fused["tune_birkhoff"]
    for c, arr in enumerate((tb.get_tune1_x(), tb.get_tune2_x(), tb.get_tune1_y(), tb.get_tune2_y())):
        np.testing.assert_allclose(got_t[:, c, :], np.asarray(arr), rtol=0, atol=1e-12, equal_nan=True)
    # and the oracle's constructs on the same turns
    tables = O.Tables(O.sps_modulation(0.168, 32.0, 0, T + 1), O.sps_modulation(0.201, 32.0, 0, T + 1))
    ot, op = O.Tracker.from_tables(tables), O.Particles(*x0)
    om = np.tile(np.eye(4), (n, 1, 1)).reshape(n, 16)
    omb, otb = O.MegnoBirkhoffConstruct(n, mb_samples), O.TuneBirkhoffConstruct(n, tb_samples)
    otb.first_add(op.x, op.px, op.y, op.py)
    for k in range(T):
        om_new = O.structured_multiply(om, op.x, op.px, op.y, op.py, tables, k, mu)
        omb.add(om_new, om)
        om = om_new
        ot.track(op, 1, mu, barrier=10.0)
        otb.add(op.x, op.px, op.y, op.py)
    np.testing.assert_allclose(fused["megno_birkhoff"], omb.get_values(), rtol=1e-12, atol=1e-14, equal_nan=True)
    for c, arr in enumerate((otb.t1x, otb.t2x, otb.t1y, otb.t2y)):
        np.testing.assert_allclose(got_t[:, c, :], arr, rtol=0, atol=1e-12, equal_nan=True)
    lost = np.isnan(op.x)
    assert lost.any() and (~lost).any()


def test_fused_windows_with_renormalisation(hm):
    """Renormalising M leaves every norm ratio, and so L1 and the
    megno_birkhoff sums, unchanged (to rounding)."""
    n, T, mu = 128, 3000, 0.0
    x0 = start(n, 52)
    t = hm.henon_tracker(T, 0.31, 0.32, "sps")
    a = t.tangent_indicators(hm.particles(*x0), T, mu, renorm_every=0, megno_birkhoff=[1000, 2000, 3000])
    b = t.tangent_indicators(hm.particles(*x0), T, mu, renorm_every=100, megno_birkhoff=[1000, 2000, 3000])
    ok = np.isfinite(a["megno_birkhoff"])
    assert ok.any()
    np.testing.assert_allclose(b["megno_birkhoff"][ok], a["megno_birkhoff"][ok], rtol=1e-9, atol=1e-12)


def test_more_samples_and_windows_than_one_launch(hm):
    """20 MEGNO/FLI samples, 6 Birkhoff windows and 6 turn samples of each
    construct kind: several passes of the kernel; every output equals the one
    of a call that asks for its chunk alone, and the samples the oracle's."""
    n, T, mu = 200, 2000, 0.5
    x0 = start(n, 53)
    t = hm.henon_tracker(T, 0.168, 0.201, "sps", epsilon=16.0)
    samples = list(range(100, 2001, 100))
    birk = [(0, 300), (100, 900), (0, 2000), (500, 1500), (1000, 2000), (7, 1234)]
    mbs, tbs = [30, 100, 400, 800, 1500, 2000], [20, 64, 300, 500, 700, 999]
    big = t.tangent_indicators(hm.particles(*x0), T, mu, renorm_every=100, samples=samples, birkhoff=birk,
                               megno_birkhoff=mbs, tune_birkhoff=tbs)
    assert big["megno"].shape == (20, n) and big["birkhoff"].shape == (6, 2, n)
    assert big["megno_birkhoff"].shape == (6, n) and big["tune_birkhoff"].shape == (6, 4, n)
    for lo in (0, 4):
        part = t.tangent_indicators(hm.particles(*x0), T, mu, renorm_every=100,
                                    samples=samples[4 * lo:4 * lo + 16], birkhoff=birk[lo:lo + 4],
                                    megno_birkhoff=mbs[lo:lo + 4], tune_birkhoff=tbs[lo:lo + 4])
        k = len(part["birkhoff"])
        assert np.array_equal(part["birkhoff"], big["birkhoff"][lo:lo + k], equal_nan=True)
        assert np.array_equal(part["megno_birkhoff"], big["megno_birkhoff"][lo:lo + k], equal_nan=True)
        assert np.array_equal(part["tune_birkhoff"], big["tune_birkhoff"][lo:lo + k], equal_nan=True)
        s = len(part["megno"])
        assert np.array_equal(part["megno"], big["megno"][4 * lo:4 * lo + s], equal_nan=True)
        assert np.array_equal(part["matrix"], big["matrix"], equal_nan=True)
    tables = O.Tables(O.sps_modulation(0.168, 16.0, 0, T), O.sps_modulation(0.201, 16.0, 0, T))
    want = O.indicators(O.Particles(*x0), tables, T, mu, 10.0, 100, samples[:16], birkhoff=birk[:4])
    np.testing.assert_allclose(big["megno"][:16], want["megno"], rtol=1e-11, atol=1e-12, equal_nan=True)
    np.testing.assert_allclose(big["birkhoff"][:4], want["birkhoff"], rtol=0, atol=1e-13, equal_nan=True)
