"""The synthetic inputs of the standard workloads are seeded, deterministic and
have the stated shapes and ranges (CPU only)."""
import numpy as np

from henon_map_cpp_amd import workloads as W


def test_da_grid():
    x, px, y, py = W.da_grid(100)
    assert x.size == 10_000 and np.all(px == 0) and np.all(py == 0)
    assert x.min() > 0 and np.isclose(x.max(), 0.6) and np.array_equal(np.unique(x), np.unique(y))
    # x runs fastest, as the reference notebooks flatten their meshgrid
    assert x[1] > x[0] and y[1] == y[0]


def test_line_modulation_coefficients():
    a = W.seeded_line_coefficients()
    assert np.isclose(a.sum(), 1e-4) and np.all(a > 0)
    assert np.array_equal(a, W.seeded_line_coefficients())
    f = W.line_frequencies()
    assert np.allclose(f / (2 * np.pi / 868.12), [1, 2, 3, 6, 7, 10, 12])


def test_indicator_points():
    c = W.CHAOS_INDICATORS
    x, px, y, py = W.indicator_points()
    assert x.size == c["n_side"] ** 2 + c["n_ball"]
    r = np.sqrt(x ** 2 + px ** 2 + y ** 2 + py ** 2)[c["n_side"] ** 2:]
    assert r.max() < c["r_ball"]
    # uniform in the 4D ball: P(r < R/2) = 1/16
    assert abs(np.mean(r < c["r_ball"] / 2) - 1 / 16) < 0.01


def test_shell_uniform_radius():
    c = W.ADJOINT_GRADIENTS
    x, px, y, py = W.shell_uniform_radius(20_000, c["r_min"], c["r_max"], c["seed"])
    r = np.sqrt(x ** 2 + px ** 2 + y ** 2 + py ** 2)
    assert r.min() >= c["r_min"] and r.max() <= c["r_max"]
    assert abs(np.median(r) - (c["r_min"] + c["r_max"]) / 2) < 0.01  # uniform in radius


def test_batch_job_params():
    p = W.batch_job_params(2000)
    assert p == W.batch_job_params(2000)
    n = np.array([q[0] for q in p])
    t = np.array([q[1] for q in p])
    assert n.min() >= 64 and n.max() <= 65_536 and t.min() >= 1_000 and t.max() <= 1_000_000
    # log-uniform: the median sits near the geometric mean of the range
    assert abs(np.log(np.median(n)) - np.log(np.sqrt(64 * 65_536))) < 0.3
    assert all(0.15 <= q[2] <= 0.32 and 0.15 <= q[3] <= 0.32 and 0 <= q[4] <= 64 and -1 <= q[5] <= 1 for q in p)
