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
