"""MI355X-native 4D modulated Henon-map tracker.

Drop-in for the reference's ``henon_map_cpp`` package (henon_map_cpp/__init__.py):
the same classes, methods and defaults, over the ``henon_map_engine`` pybind11
module, which calls the HIP kernels of ``libhenon_hip.so`` through the C ABI in
``include/henon_hip.h``.

Differences a user of the reference sees:
  * every class is GPU-backed. ``force_CPU``/``force_cpu`` are accepted and
    change nothing: the exact mode reproduces the reference CPU path bit for
    bit, so there is no separate CPU engine (and no silent CPU fallback: without
    a HIP device the constructors raise).
  * ``gpu_available()`` asks the HIP runtime, not numba.
  * ``track(..., mode="exact"|"fast")``: "fast" uses fused multiply-adds.
"""
from __future__ import annotations

import numpy as np

from . import _capi

try:
    _capi.lib()  # loads the HIP runtime first (shared with torch when present)
    from .henon_map_engine import birkhoff_weights as cpp_birkhoff_weights
    from .henon_map_engine import henon_tracker as cpp_henon_tracker  # noqa: F401
    from .henon_map_engine import henon_tracker_gpu as cpp_henon_tracker_gpu
    from .henon_map_engine import particles_4d, particles_4d_gpu  # noqa: F401
    from .henon_map_engine import get_tunes as cpp_get_tunes
    from .henon_map_engine import lyapunov_birkhoff_construct as lbc
    from .henon_map_engine import lyapunov_birkhoff_construct_multi as lbcm
    from .henon_map_engine import matrix_4d_vector as cpp_matrix_4d_vector
    from .henon_map_engine import matrix_4d_vector_gpu as cpp_matrix_4d_vector_gpu
    from .henon_map_engine import megno_birkhoff_construct as cpp_megno_birkhoff
    from .henon_map_engine import megno_construct as cpp_megno
    from .henon_map_engine import storage_4d, vector_4d_gpu  # noqa: F401
    from .henon_map_engine import storage_4d_gpu as cpp_storage_4d_gpu
    from .henon_map_engine import tune_birkhoff_construct as cpp_tune_birkhoff
    from .henon_map_engine import track_batch as cpp_track_batch
    from .henon_map_engine import autotune_batch as cpp_autotune_batch
    _ENGINE_ERROR = None
except (ImportError, OSError, AttributeError) as _e:  # not built (or stale): entry points raise on use
    _ENGINE_ERROR = _e

    def _missing(*args, **kwargs):
        raise ImportError(f"henon_map_cpp_amd: native library not built ({_ENGINE_ERROR}); "
                          "run `python -m henon_map_cpp_amd.build`")

    cpp_birkhoff_weights = cpp_henon_tracker = cpp_henon_tracker_gpu = particles_4d = particles_4d_gpu = _missing
    cpp_get_tunes = lbc = lbcm = cpp_matrix_4d_vector = cpp_matrix_4d_vector_gpu = _missing
    cpp_megno_birkhoff = cpp_megno = storage_4d = vector_4d_gpu = cpp_storage_4d_gpu = cpp_tune_birkhoff = _missing
    cpp_track_batch = cpp_autotune_batch = _missing

import pandas as pd  # noqa: E402

__all__ = ["gpu_available", "birkhoff_weights", "get_tunes", "particles", "henon_tracker", "matrix_4d_vector",
           "vector_4d", "lyapunov_birkhoff_construct", "lyapunov_birkhoff_construct_multi", "megno_construct",
           "megno_birkhoff_construct", "tune_birkhoff_construct", "storage", "storage_gpu", "particles_4d",
           "particles_4d_gpu", "storage_4d", "vector_4d_gpu", "track_batch", "autotune_batch"]


def track_batch(jobs, mode="exact"):
    """Track many independent jobs in one batched launch sequence. jobs: list of
    (henon_tracker, particles, n_turns, mu, barrier); each job advances exactly
    as tracker.track(particles, n_turns, mu, barrier) would."""
    cpp_track_batch([(t.tracker, p.particles, int(n), float(mu), float(b)) for t, p, n, mu, b in jobs], mode)


def autotune_batch(jobs, mode="exact", trial_turns=4000):
    """Times batched launch configurations on copies of the jobs' particles
    (each job cut to trial_turns turns; the particles are not advanced) and
    makes later track_batch calls of the same job mix use the fastest; returns
    it as a dict. Results do not depend on the configuration."""
    return dict(cpp_autotune_batch([(t.tracker, p.particles, int(n), float(mu), float(b)) for t, p, n, mu, b in jobs],
                                   mode, int(trial_turns)))


def gpu_available():
    """henon_map_cpp/__init__.py:22-23, through the HIP runtime."""
    return _capi.device_count() > 0


def birkhoff_weights(n):
    return np.asarray(cpp_birkhoff_weights(n))


def get_tunes(x, px):
    """src/dynamic_indicator.cpp:333-361: [fft tune, Birkhoff tune] of x + i px."""
    return np.asarray(cpp_get_tunes(np.asarray(x, dtype=np.float64), np.asarray(px, dtype=np.float64)))


class particles:
    """henon_map_cpp/__init__.py:30-127."""

    def __init__(self, x_0, px_0, y_0, py_0, force_CPU=False):
        if not (isinstance(x_0, np.ndarray) and isinstance(px_0, np.ndarray) and isinstance(y_0, np.ndarray)
                and isinstance(py_0, np.ndarray)):
            raise TypeError("Arguments must be numpy arrays.")
        if not (x_0.ndim == 1 and px_0.ndim == 1 and y_0.ndim == 1 and py_0.ndim == 1):
            raise TypeError("Arguments must be 1D arrays.")
        if not (x_0.shape == px_0.shape == y_0.shape == py_0.shape):
            raise ValueError("Arguments must be of the same shape.")
        self.particles = particles_4d_gpu(x_0, px_0, y_0, py_0)

    def reset(self):
        self.particles.reset()

    def add_ghost(self, module, displacement_kind):
        self.particles.add_ghost(module, displacement_kind)

    def renormalize(self, module):
        self.particles.renormalize(module)

    def get_displacement_module(self):
        return np.asarray(self.particles.get_displacement_module())

    def get_displacement_direction(self):
        return np.asarray(self.particles.get_displacement_direction())

    def get_x(self):
        return self.particles.get_x()

    def get_px(self):
        return np.asarray(self.particles.get_px())

    def get_y(self):
        return np.asarray(self.particles.get_y())

    def get_py(self):
        return np.asarray(self.particles.get_py())

    def get_steps(self):
        return np.asarray(self.particles.get_steps())

    def get_radius(self):
        return np.asarray(self.particles.get_radius())

    def get_radius_mean(self):
        return np.asarray(self.particles.get_radius_mean())

    def get_radius_std(self):
        return np.asarray(self.particles.get_radius_std())

    def get_action(self):
        return np.asarray(self.particles.get_action())

    def get_action_mean(self):
        return np.asarray(self.particles.get_action_mean())

    def get_action_std(self):
        return np.asarray(self.particles.get_action_std())

    def get_action_x(self):
        return np.asarray(self.particles.get_action_x())

    def get_action_x_mean(self):
        return np.asarray(self.particles.get_action_x_mean())

    def get_action_x_std(self):
        return np.asarray(self.particles.get_action_x_std())

    def get_action_y(self):
        return np.asarray(self.particles.get_action_y())

    def get_action_y_mean(self):
        return np.asarray(self.particles.get_action_y_mean())

    def get_action_y_std(self):
        return np.asarray(self.particles.get_action_y_std())

    def get_angle_x(self):
        return np.asarray(self.particles.get_angles_x())

    def get_angle_x_mean(self):
        return np.asarray(self.particles.get_angles_x_mean())

    def get_angle_x_std(self):
        return np.asarray(self.particles.get_angles_x_std())

    def get_angle_y(self):
        return np.asarray(self.particles.get_angles_y())

    def get_angle_y_mean(self):
        return np.asarray(self.particles.get_angles_y_mean())

    def get_angle_y_std(self):
        return np.asarray(self.particles.get_angles_y_std())


class henon_tracker:
    """henon_map_cpp/__init__.py:303-500."""

    def __init__(self, N, omega_x, omega_y, modulation_kind, omega_0=np.nan, epsilon=0.0, offset=0,
                 force_CPU=False, coefficients=None, frequencies=None):
        coef = [] if coefficients is None else list(np.asarray(coefficients, dtype=np.float64))
        freq = [] if frequencies is None else list(np.asarray(frequencies, dtype=np.float64))
        self.tracker = cpp_henon_tracker_gpu(N, omega_x, omega_y, modulation_kind, omega_0, epsilon, offset,
                                             coef, freq)

    def compute_a_modulation(self, n_turns, omega_x, omega_y, epsilon=0.0, modulation_kind="sps", omega_0=np.nan,
                             offset=0):
        self.tracker.compute_a_modulation(n_turns, omega_x, omega_y, modulation_kind, omega_0, epsilon, offset)

    def track(self, particles, n_turns, mu, barrier=10.0, kick_module=np.nan, inverse=False, mode="exact"):
        self.tracker.track(particles.particles, n_turns, mu, barrier, kick_module, inverse, mode)

    def autotune(self, particles, mu, barrier=10.0, mode="exact", warm_turns=2000, trial_turns=4000):
        """Times launch configurations on a copy of `particles` and makes later
        track() calls of this size class use the fastest; returns it as a dict.
        Results do not depend on the configuration."""
        return dict(self.tracker.autotune(particles.particles, mu, barrier, mode, warm_turns, trial_turns))

    def megno(self, particles, n_turns, mu, barrier=10.0, kick_module=np.nan, inverse=False,
              turn_samples=np.array([]), n_threads=-1, carry_off_diagonal=True):
        """henon_map_cpp/__init__.py megno over henon_tracker::megno. As there,
        n_threads (-1: the host's hardware threads) fixes which particles
        share a tangent matrix: each of n_threads chains takes the particles
        t, t + n_threads, ... and starts each from the previous one's matrix
        with only the diagonal reset (carry_off_diagonal=True, the reference's
        results); carry_off_diagonal=False starts every particle from the
        identity (results independent of n_threads)."""
        return np.asarray(self.tracker.megno(particles.particles, n_turns, mu, barrier, kick_module, inverse,
                                             [int(v) for v in turn_samples], n_threads, carry_off_diagonal))

    def tune_birkhoff(self, particles, n_turns, mu, barrier=10.0, kick_module=np.nan, inverse=False,
                      from_idx=np.array([]), to_idx=np.array([])):
        result = np.asarray(self.tracker.birkhoff_tunes(particles.particles, n_turns, mu, barrier, kick_module,
                                                        inverse, [int(v) for v in from_idx],
                                                        [int(v) for v in to_idx]))
        pd_result = pd.DataFrame(columns=["from", "to", "tune_x", "tune_y"])
        for i in range(len(from_idx)):
            pd_result.loc[i] = [from_idx[i], to_idx[i], result[:, i * 2], result[:, i * 2 + 1]]
        pd_result.loc[len(from_idx)] = [0, n_turns, result[:, -2], result[:, -1]]
        return pd_result

    def tune_all(self, particles, n_turns, mu, barrier=10.0, kick_module=np.nan, inverse=False,
                 from_idx=np.array([]), to_idx=np.array([])):
        result_birkhoff, result_fft = self.tracker.all_tunes(particles.particles, n_turns, mu, barrier, kick_module,
                                                             inverse, [int(v) for v in from_idx],
                                                             [int(v) for v in to_idx])
        result_birkhoff, result_fft = np.asarray(result_birkhoff), np.asarray(result_fft)
        pd_result = pd.DataFrame(columns=["from", "to", "tune_x_birkhoff", "tune_y_birkhoff", "tune_x_fft",
                                          "tune_y_fft"])
        for i in range(len(from_idx)):
            pd_result.loc[i] = [from_idx[i], to_idx[i], result_birkhoff[:, i * 2], result_birkhoff[:, i * 2 + 1],
                                result_fft[:, i * 2], result_fft[:, i * 2 + 1]]
        pd_result.loc[len(from_idx)] = [0, n_turns, result_birkhoff[:, -2], result_birkhoff[:, -1],
                                        result_fft[:, -2], result_fft[:, -1]]
        return pd_result

    def get_tangent_matrix(self, particles, mu, reverse=False):
        return self.tracker.get_tangent_matrix(particles.particles, mu, reverse)

    # ---- additions with no reference counterpart
    def tangent_indicators(self, particles, n_turns, mu, barrier=10.0, renorm_every=100, samples=(),
                           fli_direction=(1.0, 0.0, 0.0, 0.0), birkhoff=(), history=(0, 0), mode="exact",
                           megno_birkhoff=(), tune_birkhoff=()):
        """Fused tracking + tangent map + MEGNO + FLI + Birkhoff tunes (one kernel
        per 16 samples / 4 windows of each kind). megno_birkhoff and
        tune_birkhoff list turn samples T of megno_birkhoff_construct and
        tune_birkhoff_construct fed every turn of the call: the results equal
        those constructs' get_values() / get_tune1_x() ... after the same turns.
        mode: "exact", "fast", or "exact_orbit" (the orbit, losses and Birkhoff
        tunes of "exact" bit for bit, the tangent products of "fast": MEGNO and
        FLI to rounding)."""
        return self.tracker.tangent_indicators(particles.particles, n_turns, mu, barrier, renorm_every,
                                               [int(v) for v in samples], list(fli_direction),
                                               [(int(a), int(b)) for a, b in birkhoff],
                                               (int(history[0]), int(history[1])), mode,
                                               [int(v) for v in megno_birkhoff], [int(v) for v in tune_birkhoff])

    def reversibility_error(self, particles, t_list, mu, barrier=10.0, mode="exact"):
        """|x0 - B^t F^t x0| for each t (sorted), shape (len(t_list), n)."""
        return np.asarray(self.tracker.reversibility_error(particles.particles, [int(v) for v in t_list], mu,
                                                           barrier, mode))


class matrix_4d_vector:
    """henon_map_cpp/__init__.py:130-175 (GPU-backed for both flavours)."""

    def __init__(self, N, force_cpu=False):
        self.GPU = True
        self.matrix = cpp_matrix_4d_vector_gpu(N)

    def reset(self):
        self.matrix.reset()

    def multiply(self, matrices):
        self.matrix.multiply(np.asarray(matrices, dtype=np.float64))

    def structured_multiply(self, tracker, particles, mu, reverse=False):
        if reverse:
            cpp_matrix_4d_vector.structured_multiply(self.matrix, tracker.tracker, particles.particles, mu, True)
        else:
            self.matrix.structured_multiply(tracker.tracker, particles.particles, mu)

    def set_with_tracker(self, tracker, particles, mu, reverse=False):
        self.matrix.set_with_tracker(tracker.tracker, particles.particles, mu)

    def explicit_copy(self, other_matrix):
        self.matrix.explicit_copy(other_matrix.matrix)

    def get_matrix(self):
        return np.asarray(self.matrix.get_matrix())

    def get_vector(self, vector):
        return np.asarray(self.matrix.get_vector(np.asarray(vector, dtype=np.float64)))


class vector_4d:
    """henon_map_cpp/__init__.py:178-195."""

    def __init__(self, initial_vector):
        self.vector = vector_4d_gpu(np.asarray(initial_vector, dtype=np.float64))

    def set_vectors(self, vector):
        self.vector.set_vectors(np.asarray(vector, dtype=np.float64))

    def multiply(self, matrix):
        self.vector.multiply(matrix.matrix)

    def normalize(self):
        self.vector.normalize()

    def get_vectors(self):
        return np.asarray(self.vector.get_vectors())


class lyapunov_birkhoff_construct:
    """henon_map_cpp/__init__.py:198-220."""

    def __init__(self, n, n_weights):
        self.construct = lbc(n, n_weights)

    def reset(self):
        self.construct.reset()

    def change_weights(self, n_weights):
        self.construct.change_weights(n_weights)

    def add(self, vectors):
        self.construct.add(vectors.vector)

    def get_weights(self):
        return np.asarray(self.construct.get_weights())

    def get_values_raw(self):
        return np.asarray(self.construct.get_values_raw())

    def get_values_b(self):
        return np.asarray(self.construct.get_values_b())


class lyapunov_birkhoff_construct_multi:
    """henon_map_cpp/__init__.py:223-239."""

    def __init__(self, n, n_weights):
        self.construct = lbcm(n, list(n_weights))

    def reset(self):
        self.construct.reset()

    def add(self, vectors):
        self.construct.add(vectors.vector)

    def get_values_raw(self):
        return np.asarray(self.construct.get_values_raw())

    def get_values_b(self):
        return np.asarray(self.construct.get_values_b())


class megno_construct:
    """henon_map_cpp/__init__.py:242-254."""

    def __init__(self, n):
        self.construct = cpp_megno(n)

    def reset(self):
        self.construct.reset()

    def add(self, matrix_a, matrix_b):
        self.construct.add(matrix_a.matrix, matrix_b.matrix)

    def get_values(self):
        return np.asarray(self.construct.get_values())


class megno_birkhoff_construct:
    """henon_map_cpp/__init__.py:257-269."""

    def __init__(self, n, n_weights):
        self.construct = cpp_megno_birkhoff(n, list(n_weights))

    def reset(self):
        self.construct.reset()

    def add(self, matrix_a, matrix_b):
        self.construct.add(matrix_a.matrix, matrix_b.matrix)

    def get_values(self):
        return np.asarray(self.construct.get_values())


class tune_birkhoff_construct:
    """henon_map_cpp/__init__.py:272-300."""

    def __init__(self, n, n_weights):
        self.construct = cpp_tune_birkhoff(n, list(n_weights))
        self.first_called = False

    def reset(self):
        self.construct.reset()

    def first_add(self, particles):
        self.construct.first_add(particles.particles)
        self.first_called = True

    def add(self, particles):
        if not self.first_called:
            raise ValueError("First call add_first")
        self.construct.add(particles.particles)

    def get_tune1_x(self):
        return self.construct.get_tune1_x()

    def get_tune1_y(self):
        return self.construct.get_tune1_y()

    def get_tune2_x(self):
        return self.construct.get_tune2_x()

    def get_tune2_y(self):
        return self.construct.get_tune2_y()


class storage:
    """henon_map_cpp/__init__.py:503-547."""

    def __init__(self, N):
        self.storage = storage_4d(N)

    def store(self, particles):
        self.storage.store(particles.particles)

    def tune_fft(self, from_idx, to_idx, max_value):
        result = np.asarray(self.storage.tune_fft([int(v) for v in from_idx], [int(v) for v in to_idx]))
        pd_result = pd.DataFrame(columns=["from", "to", "tune_x", "tune_y"])
        for i in range(len(from_idx)):
            pd_result.loc[i] = [from_idx[i], to_idx[i], result[:, i * 2], result[:, i * 2 + 1]]
        pd_result.loc[len(from_idx)] = [0, max_value, result[:, -2], result[:, -1]]
        return pd_result

    def tune_birkhoff(self, from_idx, to_idx, max_value):
        result = np.asarray(self.storage.tune_birkhoff([int(v) for v in from_idx], [int(v) for v in to_idx]))
        pd_result = pd.DataFrame(columns=["from", "to", "tune_x", "tune_y"])
        for i in range(len(from_idx)):
            pd_result.loc[i] = [from_idx[i], to_idx[i], result[:, i * 2], result[:, i * 2 + 1]]
        pd_result.loc[len(from_idx)] = [0, max_value, result[:, -2], result[:, -1]]
        return pd_result

    def get_x(self):
        return np.asarray(self.storage.get_x())

    def get_px(self):
        return np.asarray(self.storage.get_px())

    def get_y(self):
        return np.asarray(self.storage.get_y())

    def get_py(self):
        return np.asarray(self.storage.get_py())


class storage_gpu:
    """henon_map_cpp/__init__.py:550-571."""

    def __init__(self, N, batch_size):
        self.storage = cpp_storage_4d_gpu(N, batch_size)

    def store(self, particles):
        self.storage.store(particles.particles)

    def reset(self):
        self.storage.reset()

    def get_x(self):
        return np.asarray(self.storage.get_x())

    def get_px(self):
        return np.asarray(self.storage.get_px())

    def get_y(self):
        return np.asarray(self.storage.get_y())

    def get_py(self):
        return np.asarray(self.storage.get_py())
