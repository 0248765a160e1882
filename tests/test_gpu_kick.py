"""Tracking with the random 4D kick (henon.cu:362-405, 420-542) on the GPU:
the kicked step against the oracle (same counter-based generator; normals
equal to rounding of log/sin/cos), the kick's length and isotropy, and the
generator stream's reproducibility and continuation."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hm():
    import henon_map_cpp_amd as hm

    assert hm.gpu_available()
    return hm


def _points(n, r, seed):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(4, n))
    return tuple(r * v / np.sqrt((v ** 2).sum(axis=0)))


@pytest.mark.parametrize("mu", [0.0, 0.4])
def test_kicked_tracking_matches_oracle(hm, mu):
    x, px, y, py = _points(400, 0.15, 3)
    t = hm.henon_tracker(1000, 0.168, 0.201, "sps", epsilon=16.0)
    p = hm.particles(x, px, y, py)
    p.particles.set_kick_seed(1234)
    ot = O.Tracker(1000, 0.168, 0.201, "sps", epsilon=16.0)
    op = O.Particles(x, px, y, py)
    for draw in (0, 300):  # the second call continues the particle object's stream
        t.track(p, 300, mu, barrier=10.0, kick_module=1e-5)
        ot.track_kicked(op, 300, mu, 10.0, 1e-5, seed=1234, draw=draw)
    assert np.array_equal(p.get_steps(), op.steps)
    for got, want in ((p.get_x(), op.x), (p.get_px(), op.px), (p.get_y(), op.y), (p.get_py(), op.py)):
        np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-13)
    # inverse tracking with the kick: kick after the inverse map
    t.track(p, 200, mu, barrier=10.0, kick_module=1e-5, inverse=True)
    ot.track_kicked(op, 200, mu, 10.0, 1e-5, seed=1234, draw=600, inverse=True)
    assert np.array_equal(p.get_steps(), op.steps)
    np.testing.assert_allclose(p.get_x(), op.x, rtol=1e-9, atol=1e-13)


def test_kick_length_and_isotropy(hm):
    """One kicked turn undone by one unkicked inverse turn leaves the kick: its
    length is kick_module and its direction is uniform on the 3-sphere
    (component means 0, second moments 1/4 on the diagonal, 0 off it)."""
    n, module = 200_000, 1e-3
    z = np.zeros(n)
    t = hm.henon_tracker(10, 0.168, 0.201, "sps")
    p = hm.particles(z, z, z, z)
    p.particles.set_kick_seed(99)
    t.track(p, 1, 0.0, barrier=10.0, kick_module=module)
    t.track(p, 1, 0.0, barrier=10.0, inverse=True)
    d = np.stack([p.get_x(), p.get_px(), p.get_y(), p.get_py()]) / module
    np.testing.assert_allclose(np.sqrt((d ** 2).sum(axis=0)), 1.0, rtol=1e-9)
    assert np.all(np.abs(d.mean(axis=1)) < 5 * 0.5 / np.sqrt(n))
    cov = d @ d.T / n
    np.testing.assert_allclose(np.diag(cov), 0.25, atol=0.01)
    assert np.all(np.abs(cov - np.diag(np.diag(cov))) < 0.01)


def test_kick_stream_reproducible_and_continued(hm):
    x, px, y, py = _points(64, 0.1, 5)
    t = hm.henon_tracker(400, 0.28, 0.31, "sps")

    def run(seed, calls):
        p = hm.particles(x, px, y, py)
        p.particles.set_kick_seed(seed)
        for n in calls:
            t.track(p, n, 0.0, barrier=10.0, kick_module=1e-4)
        return p.get_x()

    a, b = run(5, [100]), run(5, [100])
    assert np.array_equal(a, b)
    assert not np.array_equal(a, run(6, [100]))
    # two calls of 50 continue one stream: equal to one call of 100
    assert np.array_equal(run(5, [50, 50]), a)
