"""Pins the oracle's restatement of the reference's construct formulas
(megno_construct henon.cu:2253-2284, megno_birkhoff_construct :2366-2402,
tune_birkhoff_construct :2533-2584) to closed forms at every idx, and checks
that a plausible mistake in each would be caught."""
import math

import numpy as np
import pytest

import oracle as O
from tests import construct_cases as CC

N = 5


def run_megno(kind, n_adds, construct):
    s = CC.megno_scales(kind, n_adds)
    out = []
    for i in range(1, n_adds + 1):
        construct.add(CC.identities(s[i], N), CC.identities(s[i - 1], N))
        out.append(construct.get_values())
    return out


@pytest.mark.parametrize("kind", ["const2", "harmonic"])
def test_megno_construct_closed_form(kind):
    vals = run_megno(kind, 300, O.MegnoConstruct(N))
    for k in (1, 2, 3, 10, 57, 300):
        np.testing.assert_allclose(vals[k - 1], CC.megno_expected(kind, k), rtol=1e-13)


def test_megno_construct_reset():
    c = O.MegnoConstruct(N)
    run_megno("const2", 20, c)
    c.reset()
    vals = run_megno("harmonic", 40, c)
    np.testing.assert_allclose(vals[-1], CC.megno_expected("harmonic", 40), rtol=1e-13)


@pytest.mark.parametrize("kind", ["const2", "harmonic"])
def test_megno_birkhoff_construct_closed_form(kind):
    ts = [1, 10, 37, 100, 256]
    c = O.MegnoBirkhoffConstruct(N, ts)
    run_megno(kind, 300, c)
    got = c.get_values()
    assert got.shape == (len(ts), N)
    for j, t in enumerate(ts):
        if t == 1:  # birkhoff_weights(1) = [0] / 0: NaN, as the reference's
            assert np.all(np.isnan(got[j]))
        else:
            np.testing.assert_allclose(got[j], CC.megno_birkhoff_expected(kind, t), rtol=1e-13)


def test_tune_birkhoff_construct_closed_form():
    ts = [16, 50, 128, 200]
    n_adds = 2 * max(ts) + 7
    x, px, y, py = CC.tune_points(N, n_adds)
    c = O.TuneBirkhoffConstruct(N, ts)
    c.first_add(x[0], px[0], y[0], py[0])
    for m in range(1, n_adds + 1):
        c.add(x[m], px[m], y[m], py[m])
    for j, t in enumerate(ts):
        np.testing.assert_allclose(c.t1x[j], CC.tune_expected(CC.TUNE_X, t, 1), rtol=0, atol=1e-13)
        np.testing.assert_allclose(c.t2x[j], CC.tune_expected(CC.TUNE_X, t, 2), rtol=0, atol=1e-13)
        np.testing.assert_allclose(c.t1y[j], CC.tune_expected(CC.TUNE_Y, t, 1), rtol=0, atol=1e-13)
        np.testing.assert_allclose(c.t2y[j], CC.tune_expected(CC.TUNE_Y, t, 2), rtol=0, atol=1e-13)


def test_tune_birkhoff_construct_linear_rotation():
    """Unmodulated map at tiny amplitude (mu = 0): a rotation by omega per turn,
    so the construct's tune 1 gives 1 - tune1/(2 pi) = nu."""
    nu_x, nu_y, T = 0.168, 0.201, 400
    tables = O.Tables(O.sps_modulation(nu_x, 0.0, 0, 2 * T), O.sps_modulation(nu_y, 0.0, 0, 2 * T))
    t = O.Tracker.from_tables(tables)
    p = O.Particles([1e-7, 2e-7], [0.0, 1e-7], [1e-7, 0.0], [0.0, 1e-7])
    c = O.TuneBirkhoffConstruct(2, [T])
    c.first_add(p.x, p.px, p.y, p.py)
    for _ in range(2 * T):
        t.track(p, 1, 0.0, barrier=10.0)
        c.add(p.x, p.px, p.y, p.py)
    for arr, nu in ((c.t1x, nu_x), (c.t2x, nu_x), (c.t1y, nu_y), (c.t2y, nu_y)):
        np.testing.assert_allclose(1.0 - arr[0] / (2 * math.pi), nu, rtol=0, atol=1e-12)


def test_closed_forms_catch_plausible_mistakes():
    """The closed forms separate the reference recursion from its nearest
    mistakes, evaluated here in numpy on the same inputs."""
    k = 50
    s = CC.megno_scales("const2", k)
    r = np.log10(s[1:] / s[:-1])
    idx = np.arange(1, k + 1)

    def megno(weight_l1, inv_l2, inv_l3):
        l1 = np.cumsum(r * weight_l1(idx))
        l2 = np.cumsum(l1 / idx if inv_l2 else l1)
        return l2[-1] / k if inv_l3 else l2[-1]

    want = CC.megno_expected("const2", k)
    assert megno(lambda i: i, True, True) == pytest.approx(want, rel=1e-13)
    for wrong in ((lambda i: i * i, True, True), (lambda i: i, False, True), (lambda i: i, True, False),
                  (lambda i: np.ones_like(i), True, True)):
        assert abs(megno(*wrong) - want) > 0.1 * want
    # megno_birkhoff: w[idx-1] against the shifted w[idx]
    t = 37
    w = O.birkhoff_weights(t)
    l1 = np.cumsum(r * idx)[:t]
    right = np.sum(l1 * w[idx[:t] - 1] / idx[:t])
    shifted = np.sum(l1[:-1] * w[idx[:t - 1]] / idx[:t - 1])
    assert right == pytest.approx(CC.megno_birkhoff_expected("const2", t), rel=1e-13)
    assert abs(shifted - right) > 0.1
