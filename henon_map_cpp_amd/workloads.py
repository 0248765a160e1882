"""Synthetic inputs of the standard workloads: dynamic-aperture grids, the
seeded 7-line tune modulation, random initial conditions in 4D balls/shells.

Every generator is seeded and deterministic, so the tests, the benchmark and
the CPU baseline see identical inputs.
"""
from __future__ import annotations

import numpy as np

TWO_PI = 2.0 * np.pi

# Line multipliers of the SPS-like modulation spectrum: f_k = m_k / 868.12 turns^-1
# (the multipliers of src/modulation.cpp:22-28).
LINE_MULTIPLIERS = np.array([1, 2, 3, 6, 7, 10, 12], dtype=np.float64)
LINE_PERIOD = 868.12
MODULATION_SEED = 20240607


def line_frequencies() -> np.ndarray:
    """Angular frequencies m_k * (2 pi / 868.12), as the reference computes them."""
    return np.array([m * (2 * np.pi / 868.12) for m in LINE_MULTIPLIERS])


def seeded_line_coefficients(seed: int = MODULATION_SEED) -> np.ndarray:
    """a_k ~ U(0.1, 1) from `seed`, normalised to sum to 1, times 1e-4
    (the coefficient scale of the SPS lines)."""
    a = np.random.default_rng(seed).uniform(0.1, 1.0, 7)
    return a / a.sum() * 1e-4


def da_grid(n_side: int, lo: float = 0.0, hi: float = 0.6):
    """n_side x n_side grid over (lo, hi]^2 in (x, y), px = py = 0 (x fastest)."""
    v = np.linspace(lo, hi, n_side + 1)[1:]
    xx, yy = np.meshgrid(v, v)
    x = xx.flatten()
    y = yy.flatten()
    z = np.zeros_like(x)
    return x, z.copy(), y, z.copy()


def ball(n: int, r_max: float, seed: int, r_min: float = 0.0):
    """n points in the 4D shell r_min <= r < r_max: Gaussian directions and
    radii with density ~ r^3 (uniform in the 4D volume)."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 4))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    u = rng.uniform(size=n)
    r = (r_min ** 4 + u * (r_max ** 4 - r_min ** 4)) ** 0.25
    p = d * r[:, None]
    return p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), p[:, 3].copy()


# Named workloads (sizes as stated in their names)
UNMODULATED_DA = dict(name="unmodulated DA scan 100x100, 1e4 turns, mu=0", n_side=100, n_turns=10_000,
                      tune_x=0.168, tune_y=0.201, epsilon=0.0, mu=0.0, barrier=float(np.sqrt(10.0)), kind="sps")
MODULATED_DA = dict(name="modulated DA scan 1000x1000, 1e5 turns, mu=0.5, eps=32", n_side=1000, n_turns=100_000,
                    tune_x=0.168, tune_y=0.201, epsilon=32.0, mu=0.5, barrier=float(np.sqrt(10.0)), kind="custom")
CHAOS_INDICATORS = dict(name="chaos indicators: 256x256 grid + 65536 ball points r<0.5, 1e5 turns, mu=0.5, eps=32, "
                             "renormalisation every 100 turns",
                        n_side=256, n_ball=65_536, r_ball=0.5, n_turns=100_000, renorm_every=100,
                        tune_x=0.168, tune_y=0.201, epsilon=32.0, mu=0.5, barrier=float(np.sqrt(10.0)),
                        samples=(1_000, 10_000, 100_000), seed=3)
ADJOINT_GRADIENTS = dict(name="adjoint gradients: 65536 points, r ~ U[0.05, 0.4], 1e4 turns, mu=0.5, eps=16",
                         n=65_536, r_min=0.05, r_max=0.4, n_turns=10_000, tune_x=0.168, tune_y=0.201, epsilon=16.0,
                         mu=0.5, barrier=float(np.sqrt(10.0)), every=(8, 64, 256, 1024), seed=4)
BATCHED_JOBS = dict(name="batched uneven jobs: N ~ logU[64, 65536], T ~ logU[1e3, 1e6], tunes U[0.15, 0.32], "
                         "eps U[0, 64], mu U[-1, 1], r < 0.6, barrier r^2 > 10",
                    n_jobs=4096, n_range=(64, 65_536), t_range=(1_000, 1_000_000), tune_range=(0.15, 0.32),
                    eps_range=(0.0, 64.0), mu_range=(-1.0, 1.0), r_max=0.6, barrier=float(np.sqrt(10.0)), seed=5,
                    megno_renorm_every=100)


def batch_megno_job(k: int) -> bool:
    """Jobs k with k mod 10 < 3 (30 % of the mix) also accumulate the MEGNO."""
    return k % 10 < 3


def indicator_points(c=CHAOS_INDICATORS):
    """The DA grid followed by uniform points of the 4D ball."""
    g = da_grid(c["n_side"])
    b = ball(c["n_ball"], c["r_ball"], seed=c["seed"])
    return tuple(np.concatenate([u, v]) for u, v in zip(g, b))


def shell_uniform_radius(n: int, r_min: float, r_max: float, seed: int):
    """n points with radius ~ U[r_min, r_max] and a uniform direction on S^3."""
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 4))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = rng.uniform(r_min, r_max, n)
    p = d * r[:, None]
    return p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), p[:, 3].copy()


def batch_job_params(n_jobs: int, c=BATCHED_JOBS):
    """Seeded per-job parameters: (n, n_turns, tune_x, tune_y, epsilon, mu, seed)."""
    rng = np.random.default_rng(c["seed"])
    lo_n, hi_n = np.log(c["n_range"][0]), np.log(c["n_range"][1])
    lo_t, hi_t = np.log(c["t_range"][0]), np.log(c["t_range"][1])
    out = []
    for k in range(n_jobs):
        n = int(round(np.exp(rng.uniform(lo_n, hi_n))))
        t = int(round(np.exp(rng.uniform(lo_t, hi_t))))
        qx, qy = rng.uniform(*c["tune_range"], 2)
        out.append((n, t, float(qx), float(qy), float(rng.uniform(*c["eps_range"])),
                    float(rng.uniform(*c["mu_range"])), int(rng.integers(1 << 31))))
    return out
