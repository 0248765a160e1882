"""Device adjoint (whole-orbit reverse-mode derivatives) against what the
mathematics fixes: the forward sweep equals the oracle's tracking bit for bit,
dL/dx0 equals the transposed product of the oracle's one-turn tangent matrices
applied to the seeds, and the parameter derivatives equal central finite
differences of the oracle's tracking."""
import numpy as np
import pytest
import torch

import oracle as O
from henon_map_cpp_amd import _capi as C
from henon_map_cpp_amd import adjoint as A
from henon_map_cpp_amd import workloads as W

pytestmark = pytest.mark.gpu

T = 300
QX, QY, EPS, MU = 0.168, 0.201, 8.0, 0.3


def oracle_orbit(x0, n_turns, qx, qy, eps, mu, barrier=10.0, jac=False):
    """Turn-by-turn oracle tracking: final state, B = sum w_k x_{k+1}^2 and
    (when jac) the tangent matrices M_k (M_0 = I, M_{k+1} = J_k M_k)."""
    tables = O.Tables(O.sps_modulation(qx, eps, 0, n_turns), O.sps_modulation(qy, eps, 0, n_turns))
    tr = O.Tracker.from_tables(tables)
    p = O.Particles(*x0)
    w = O.birkhoff_weights(n_turns)
    b = np.zeros(p.n)
    mats = [np.tile(np.eye(4), (p.n, 1, 1))] if jac else None
    xs = []
    for k in range(n_turns):
        if jac:
            J = tr.get_tangent_matrix(p, mu)
            mats.append(J @ mats[-1])
        tr.track(p, 1, mu, barrier=barrier, n_threads=1)
        b = b + w[k] * p.x * p.x
        xs.append(p.x.copy())
    return np.stack([p.x, p.px, p.y, p.py]), b, mats, np.array(xs), w


def loss(state, b, a, c):
    return float(np.sum(a * state) + np.sum(c * b))


@pytest.fixture(scope="module")
def case():
    rng = np.random.default_rng(5)
    x0 = np.stack(W.ball(48, 0.25, seed=41))
    a = rng.normal(size=(4, 48))
    c = rng.normal(size=48)
    return x0, a, c


def run_adjoint(x0, a, c, every, mode="exact", qx=QX, qy=QY, eps=EPS, mu=MU):
    adj = A.OrbitAdjoint(T, qx, qy, eps, mu, barrier=10.0, every=every, mode=mode)
    st, b, alive, ck = adj.forward(torch.as_tensor(x0, device="cuda"))
    gx, gp = adj.backward(ck, alive, torch.as_tensor(a, device="cuda"), torch.as_tensor(c, device="cuda"))
    torch.cuda.synchronize()
    return st.cpu().numpy(), b.cpu().numpy(), alive.cpu().numpy(), gx.cpu().numpy(), gp.cpu().numpy()


@pytest.mark.parametrize("every", [40, 8, 300, 7])
def test_adjoint_forward_and_grad_x0(case, every):
    x0, a, c = case
    st, b, alive, gx, gp = run_adjoint(x0, a, c, every)
    want_st, want_b, mats, xs, w = oracle_orbit(x0, T, QX, QY, EPS, MU, jac=True)
    assert np.all(alive == 1)
    assert np.array_equal(st, want_st)                    # exact mode: bit-exact with tracking
    np.testing.assert_allclose(b, want_b, rtol=1e-13)     # fma accumulation of the same products
    # dL/dx0 = M_T^T a + sum_k 2 c w_k x_{k+1} M_{k+1}^T e_x
    want = np.einsum("nij,in->jn", mats[T], a)
    for k in range(T):
        want += (2.0 * c * w[k] * xs[k])[None, :] * mats[k + 1][:, 0, :].T
    scale = np.abs(want).max(axis=0, keepdims=True)
    assert np.all(np.abs(gx - want) <= 1e-10 * scale)


def test_adjoint_param_grads_vs_central_differences(case):
    x0, a, c = case
    _, _, _, _, gp = run_adjoint(x0, a, c, 40)
    g = gp.sum(axis=1)  # d/d eps, d/d mu, d/d omega_x0, d/d omega_y0
    base = dict(qx=QX, qy=QY, eps=EPS, mu=MU)
    checks = [("eps", 1e-3, g[0]), ("mu", 1e-6, g[1]),
              ("qx", 1e-8, 2 * np.pi * g[2]), ("qy", 1e-8, 2 * np.pi * g[3])]
    for name, h, got in checks:
        lo, hi = dict(base), dict(base)
        lo[name] -= h
        hi[name] += h
        sl, bl, *_ = oracle_orbit(x0, T, lo["qx"], lo["qy"], lo["eps"], lo["mu"])
        sh, bh, *_ = oracle_orbit(x0, T, hi["qx"], hi["qy"], hi["eps"], hi["mu"])
        fd = (loss(sh, bh, a, c) - loss(sl, bl, a, c)) / (2 * h)
        assert got == pytest.approx(fd, rel=2e-5), name


def test_adjoint_lost_particles_get_zero_gradient():
    x0 = np.array([[0.1, 2.5, 0.2], [0.0, 0.0, 0.0], [0.1, 2.5, 0.1], [0.0, 0.0, 0.0]])
    seed = np.full((4, 3), np.nan)  # seeds of lost particles are ignored
    seed[:, 0] = seed[:, 2] = 1.0
    st, b, alive, gx, gp = run_adjoint(x0, seed, np.ones(3), 40)
    want_st, _, _, _, _ = oracle_orbit(x0, T, QX, QY, EPS, MU)
    assert list(alive) == [1, 0, 1]
    assert np.array_equal(st, want_st, equal_nan=True)
    assert np.all(np.isnan(st[:, 1])) and np.isnan(b[1])
    assert np.all(gx[:, 1] == 0) and np.all(gp[:, 1] == 0)
    assert np.all(np.isfinite(gx[:, [0, 2]]))


def test_adjoint_fast_mode_close_to_exact(case):
    x0, a, c = case
    e = run_adjoint(x0, a, c, 40, mode="exact")
    f = run_adjoint(x0, a, c, 40, mode="fast")
    for u, v in zip(e[3:], f[3:]):
        np.testing.assert_allclose(v, u, rtol=1e-8, atol=1e-8 * np.abs(u).max())


def test_track_orbit_autograd(case):
    x0, a, c = case
    xt = torch.tensor(x0, device="cuda", requires_grad=True)
    par = torch.tensor([EPS, MU, QX, QY], dtype=torch.float64, device="cuda", requires_grad=True)
    st, b, alive = A.track_orbit(xt, par, T, every=40)
    at, ct = torch.as_tensor(a, device="cuda"), torch.as_tensor(c, device="cuda")
    (torch.sum(at * st) + torch.sum(ct * b)).backward()
    _, _, _, gx, gp = run_adjoint(x0, a, c, 40)
    np.testing.assert_array_equal(xt.grad.cpu().numpy(), gx)
    want = gp.sum(axis=1) * np.array([1, 1, 2 * np.pi, 2 * np.pi])
    np.testing.assert_allclose(par.grad.cpu().numpy(), want, rtol=1e-12)


@pytest.mark.parametrize("every,segments", [(40, 1), (40, 3), (300, 1), (17, 2)])
def test_adjoint_windows_match_one_pass(case, every, segments):
    """A scratch buffer of a few segments sweeps in windows of segments, carrying
    the adjoint between them: the gradients are the one-pass gradients bit for bit."""
    x0, a, c = case
    adj = A.OrbitAdjoint(T, QX, QY, EPS, MU, barrier=10.0, every=every)
    st, b, alive, ck = adj.forward(torch.as_tensor(x0, device="cuda"))
    at, ct = torch.as_tensor(a, device="cuda"), torch.as_tensor(c, device="cuda")
    gx, gp = adj.backward(ck, alive, at, ct)
    seg = ((every + 7) // 8 - 1) * 4 * x0.shape[1] * 8
    gxw, gpw = adj.backward(ck, alive, at, ct, scratch_bytes=segments * seg)
    torch.cuda.synchronize()
    assert torch.equal(gx, gxw) and torch.equal(gp, gpw)
    with pytest.raises(Exception, match="scratch"):
        adj.backward(ck, alive, at, ct, scratch_bytes=seg - 8)


@pytest.fixture(scope="module")
def config_subset():
    """The full adjoint workload (65 536 points, r ~ U[0.05, 0.4], 1e4 turns,
    eps 16, mu 0.5, seeded 7-line modulation): the 64 most regular of its first
    512 particles by the oracle's FLI, their final tangent matrices M_T and
    random seeds."""
    c = W.ADJOINT_GRADIENTS
    T = c["n_turns"]
    x0 = np.stack(W.shell_uniform_radius(c["n"], c["r_min"], c["r_max"], c["seed"]))
    coef, freq = W.seeded_line_coefficients(), W.line_frequencies()

    def tables(tx, ty, eps):
        return O.Tables(C.lines_modulation(tx, eps, coef, freq, 0, T), C.lines_modulation(ty, eps, coef, freq, 0, T))

    tb = tables(c["tune_x"], c["tune_y"], c["epsilon"])
    p = O.Particles(*[np.ascontiguousarray(v[:512]) for v in x0])
    fli = O.indicators(p, tb, T, c["mu"], c["barrier"], 100, [T], n_threads=16)["fli"][0]
    sub = np.sort(np.argsort(np.where(np.isfinite(p.x), fli, np.inf))[:64])
    p = O.Particles(*[np.ascontiguousarray(v[sub]) for v in x0])
    M = O.indicators(p, tb, T, c["mu"], c["barrier"], 0, [T], n_threads=16)["matrix"]
    a = np.random.default_rng(7).normal(size=(4, sub.size))
    return dict(c=c, T=T, x0=x0, coef=coef, freq=freq, tables=tables, sub=sub, M=M, final=p, a=a)


@pytest.mark.parametrize("every", [64, 256, 1024])
def test_adjoint_full_workload_vs_jacobian_and_differences(config_subset, every):
    """At the workload's full size and checkpoint intervals: the forward equals
    the oracle's tracking, dL/dx0 equals M_T^T a within 1e-8 of its largest
    component, and the parameter gradients equal central differences of the
    oracle's tracking (regular orbits; steps mu 1e-6, eps 1e-3, tunes 1e-8,
    whose own truncation and rounding, measured by changing the step threefold,
    stay below 2e-6 relative for mu and eps and 2e-4 for the tunes)."""
    s = config_subset
    c, T, sub, a = s["c"], s["T"], s["sub"], s["a"]
    adj = A.OrbitAdjoint(T, c["tune_x"], c["tune_y"], c["epsilon"], c["mu"], c["barrier"], coef=s["coef"],
                         freq=s["freq"], every=every)
    st, birk, alive, ck = adj.forward(torch.as_tensor(s["x0"], device="cuda"))
    seed = np.zeros((4, c["n"]))
    seed[:, sub] = a
    gx, gp = adj.backward(ck, alive, torch.as_tensor(seed, device="cuda"))
    st, alive, gx, gp = st.cpu().numpy(), alive.cpu().numpy(), gx.cpu().numpy(), gp.cpu().numpy()
    f = s["final"]
    assert np.all(alive[sub] == 1)
    assert np.array_equal(st[:, sub], np.stack([f.x, f.px, f.y, f.py]))
    want = np.einsum("nij,in->jn", s["M"], a)
    scale = np.abs(want).max(axis=0, keepdims=True)
    assert np.all(np.abs(gx[:, sub] - want) <= 1e-8 * scale)
    others = np.setdiff1d(np.arange(c["n"]), sub)
    assert np.all(gx[:, others] == 0) and np.all(gp[:, others] == 0)
    g = gp[:, sub].sum(axis=1)  # d/d eps, d/d mu, d/d omega_x0, d/d omega_y0
    base = dict(mu=c["mu"], eps=c["epsilon"], tx=c["tune_x"], ty=c["tune_y"])

    def loss(**kw):
        k = dict(base)
        k.update(kw)
        q = O.Particles(*[np.ascontiguousarray(v[sub]) for v in s["x0"]])
        O.Tracker.from_tables(s["tables"](k["tx"], k["ty"], k["eps"])).track(q, T, k["mu"], barrier=c["barrier"],
                                                                            n_threads=16)
        return float(np.sum(a * np.stack([q.x, q.px, q.y, q.py])))

    for name, h, got, rel in (("eps", 1e-3, g[0], 2e-5), ("mu", 1e-6, g[1], 2e-5),
                              ("tx", 1e-8, 2 * np.pi * g[2], 1e-3), ("ty", 1e-8, 2 * np.pi * g[3], 1e-3)):
        fd = (loss(**{name: base[name] + h}) - loss(**{name: base[name] - h})) / (2 * h)
        assert got == pytest.approx(fd, rel=rel), name
