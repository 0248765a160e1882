"""Chaos indicators from shadow particles, on the GPU.

Counterparts of ``henon_map_cpp/dynamic_indicators.py``: the reference computes
these in numpy from displaced copies it tracks next to the orbits (through an
engine class of an older tracker API); here the copies are ghost sets of one
``particles`` object, tracked by the same kernels, and the per-particle indices
come from ``hm_ghost_alignment``. Results are pandas DataFrames indexed by the
sample turn ``t`` with one column per particle.
"""
from __future__ import annotations

import numpy as np
import pandas as pd


def _frame(ts, rows):
    return pd.DataFrame(np.array(rows), index=pd.Index(ts, name="t"))


def _alignment(tracker, p, t_list, tau, mu, barrier, mode, get, start):
    """henon_map_cpp/dynamic_indicators.py:173-224 and 266-366: blocks of tau
    turns up to t_list[-1]; after each block the running minimum of the index;
    recorded at the first block end at or past each t in t_list."""
    t_list = [int(t) for t in t_list]
    if not t_list or tau <= 0:
        raise ValueError("t_list must be non-empty and tau positive")
    acc = np.full(p.particles.get_n_particles(), start)
    ts, rows, j = [], [], 0
    for i in range(tau, t_list[-1] + tau - 1, tau):
        tracker.track(p, tau, mu, barrier, mode=mode)
        acc = np.minimum(acc, get())
        while j < len(t_list) and i >= t_list[j]:
            ts.append(t_list[j])
            rows.append(acc.copy())
            j += 1
    return _frame(ts, rows)


def smallest_alignment_index(tracker, x, px, y, py, t_list, tau, mod_d, mu=0.0, barrier=10.0, mode="exact"):
    """SALI (henon_map_cpp/dynamic_indicators.py:154-224): the orbits and two
    shadow sets displaced by mod_d along x and along y; per block,
    min(|d1 + d2|, |d1 - d2|) of the normalised deviations, starting from
    sqrt(2). Regular orbits keep SALI away from 0; chaotic ones drive it to 0."""
    from . import particles

    p = particles(*(np.ascontiguousarray(a, dtype=np.float64) for a in (x, px, y, py)))
    p.add_ghost(mod_d, "x")
    p.add_ghost(mod_d, "y")
    return _alignment(tracker, p, t_list, tau, mu, barrier, mode, p.particles.get_sali, np.sqrt(2.0))


def global_alignment_index(tracker, x, px, y, py, t_list, tau, mod_d, mu=0.0, barrier=10.0, mode="exact"):
    """GALI_4 (henon_map_cpp/dynamic_indicators.py:227-366): four shadow sets
    displaced by mod_d along x, y, px and py; per block, the product of the
    singular values of the matrix of normalised deviations (its |det|),
    starting from 1."""
    from . import particles

    p = particles(*(np.ascontiguousarray(a, dtype=np.float64) for a in (x, px, y, py)))
    for d in ("x", "y", "px", "py"):
        p.add_ghost(mod_d, d)
    return _alignment(tracker, p, t_list, tau, mu, barrier, mode, p.particles.get_gali, 1.0)


def fast_lyapunov_indicator(tracker, x, px, y, py, t_list, mod_d, mu=0.0, barrier=10.0, mode="exact"):
    """FLI from a shadow particle (henon_map_cpp/dynamic_indicators.py:89-133):
    one ghost set displaced by mod_d along a random 4D direction (the particle
    generator's, default-seeded), tracked with the orbits;
    log10(|displacement| / mod_d) / t at each t in t_list. (The tangent-map FLI
    of henon_tracker.tangent_indicators needs no shadow particle.)"""
    from . import particles

    p = particles(*(np.ascontiguousarray(a, dtype=np.float64) for a in (x, px, y, py)))
    p.add_ghost(mod_d, "random")
    done, rows = 0, []
    for t in t_list:
        t = int(t)
        if t < done:
            raise ValueError("t_list must be non-decreasing")
        if t > done:
            tracker.track(p, t - done, mu, barrier, mode=mode)
            done = t
        rows.append(np.log10(p.get_displacement_module()[:, 0] / mod_d) / t)
    return _frame([int(t) for t in t_list], rows)
