"""Oracle for the 4D modulated Henon-map path -- TEST INFRASTRUCTURE ONLY.

A CPU restatement of the reference's CPU path (``oracle/oracle.cpp``), loaded
through ctypes. Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg use it, always as the checker or the timed CPU baseline,
never as part of the product.

Functions with no reference counterpart (the fused renormalised indicators,
``indicators()`` below) are pinned by closed forms, finite differences and
tangent-map products, as stated there.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_REF_PATH = os.path.join(_HERE, "_ref", "libref_modulation.so")

_lib = None

D = ctypes.c_double
P = ctypes.c_void_p
SZ = ctypes.c_size_t
U = ctypes.c_uint
I = ctypes.c_int


def build(force: bool = False, with_ref: bool = False) -> None:
    """Compile liboracle.so; with `with_ref`, also oracle/_ref from the
    reference's own sources (only where the reference checkout exists: the
    build step, never a test or benchmark run)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-B", "liboracle.so"], cwd=_HERE)
    if with_ref and os.path.isdir("/root/reference/src") and (force or not os.path.exists(_REF_PATH)):
        subprocess.check_call(["make", "-s", "-B", "ref"], cwd=_HERE)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sigs = {
            "or_basic_modulation": (I, [D, D, D, I, I, P]),
            "or_sps_modulation": (I, [D, D, I, I, P]),
            "or_lines_modulation": (I, [D, D, P, P, I, I, I, P]),
            "or_sps_lines": (None, [P, P]),
            "or_sin_cos": (None, [P, SZ, P, P]),
            "or_track": (I, [P, P, P, P, P, SZ, P, P, P, P, U, D, D, D, I, I]),
            "or_tangent_matrix": (None, [D] * 9 + [P]),
            "or_inverse_tangent_matrix": (None, [D] * 9 + [P]),
            "or_multiply_matrices": (None, [P, P, P]),
            "or_get_tangent_matrix": (None, [P, P, P, P, P, SZ, P, P, P, P, U, D, I, P]),
            "or_structured_multiply": (None, [P, P, P, P, P, SZ, P, P, P, P, U, D]),
            "or_megno": (I, [P, P, P, P, P, SZ, P, P, P, P, U, D, D, I, P, SZ, I, I, P]),
            "or_birkhoff_weights": (None, [U, P]),
            "or_birkhoff_tune": (D, [P, P, SZ]),
            "or_birkhoff_tune_vec": (I, [P, P, P, P, SZ, P, P, SZ, P]),
            "or_fft_magnitude": (None, [P, P, SZ, I, P]),
            "or_interp_A": (D, [D, D, D]),
            "or_interpolation": (D, [P, SZ]),
            "or_fft_tune": (D, [P, P, SZ]),
            "or_fft_tune_vec": (I, [P, P, P, P, SZ, P, P, SZ, P]),
            "or_get_tunes": (None, [P, P, SZ, P]),
            "or_birkhoff_tunes": (I, [P, P, P, P, P, SZ, P, P, P, P, U, D, D, I, P, P, SZ, I, P]),
            "or_all_tunes": (I, [P, P, P, P, P, SZ, P, P, P, P, U, D, D, I, P, P, SZ, I, P, P]),
            "or_lyap_birk_add": (None, [P, P, P, P, SZ, SZ, SZ]),
            "or_megno_construct_add": (None, [P, P, P, P, P, SZ, SZ]),
            "or_megno_birkhoff_add": (None, [P, P, P, P, P, SZ, P, P, SZ, SZ]),
            "or_tune_birkhoff_init": (None, [P, P, P, P, P, P, SZ]),
            "or_tune_birkhoff_add": (None, [P] * 10 + [P, P, P, SZ, SZ, SZ]),
            "or_matvec": (None, [P, P, SZ]),
            "or_normalize": (None, [P, SZ]),
            "or_nan_mean": (D, [P, SZ]),
            "or_nan_std": (D, [P, SZ]),
            "or_random_kicks": (None, [ctypes.c_uint64, SZ, D, P, P, P, P]),
            "or_philox4x32_10": (None, [P, U, U]),
            "or_kick_normals": (None, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, P]),
            "or_track_kicked": (I, [P, P, P, P, P, SZ, P, P, P, P, SZ, U, D, D, D, I, ctypes.c_uint64,
                                    ctypes.c_uint64]),
            "or_ghost_renormalize": (None, [P, P, P, P, SZ, SZ, D]),
            "or_ghost_displacement": (None, [P, P, P, P, SZ, SZ, P, P]),
            "or_indicators": (I, [P, P, P, P, P, SZ, P, P, P, P, SZ, U, D, D, U, P, SZ, P, P, P, P, P, P, P, SZ,
                                  P, I]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ref_lib():
    """oracle/_ref: the reference's own modulation.cpp, or None if not built."""
    if not os.path.exists(_REF_PATH):
        return None
    L = ctypes.CDLL(_REF_PATH)
    L.ref_sps_modulation.argtypes = [D, D, I, I, P]
    L.ref_sps_modulation.restype = I
    L.ref_basic_modulation.argtypes = [D, D, D, I, I, P]
    L.ref_basic_modulation.restype = I
    L.ref_pick_a_modulation.argtypes = [U, D, D, ctypes.c_char_p, D, D, U, P, P]
    L.ref_pick_a_modulation.restype = I
    return L


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


# ---------------------------------------------------------------- modulation
def sps_modulation(tune, epsilon, start, end):
    out = np.empty(end - start)
    if lib().or_sps_modulation(tune, epsilon, start, end, _p(out)):
        raise ValueError("start must be < end")
    return out


def basic_modulation(tune, omega, epsilon, start, end):
    out = np.empty(end - start)
    if lib().or_basic_modulation(tune, omega, epsilon, start, end, _p(out)):
        raise ValueError("start must be < end")
    return out


def lines_modulation(tune, epsilon, coef, freq, start, end):
    coef, freq = _f64(coef), _f64(freq)
    out = np.empty(end - start)
    if lib().or_lines_modulation(tune, epsilon, _p(coef), _p(freq), len(coef), start, end, _p(out)):
        raise ValueError("start must be < end")
    return out


def sps_lines():
    c, f = np.empty(7), np.empty(7)
    lib().or_sps_lines(_p(c), _p(f))
    return c, f


def sin_cos(w):
    w = _f64(w)
    s, c = np.empty_like(w), np.empty_like(w)
    lib().or_sin_cos(_p(w), w.size, _p(s), _p(c))
    return s, c


def pick_a_modulation(n_turns, omega_x, omega_y, kind, omega_0=np.nan, epsilon=0.0, offset=0, coef=None, freq=None):
    """modulation.cpp:98-134 for the deterministic kinds, plus the 'custom' lines kind."""
    if kind == "sps":
        return (sps_modulation(omega_x, epsilon, offset, offset + n_turns),
                sps_modulation(omega_y, epsilon, offset, offset + n_turns))
    if kind == "basic":
        return (basic_modulation(omega_x, omega_0, epsilon, offset, offset + n_turns),
                basic_modulation(omega_y, omega_0, epsilon, offset, offset + n_turns))
    if kind == "custom":
        return (lines_modulation(omega_x, epsilon, coef, freq, offset, offset + n_turns),
                lines_modulation(omega_y, epsilon, coef, freq, offset, offset + n_turns))
    if kind == "none":
        # the reference fills empty vectors here; the well-defined reading is a
        # constant 2*pi*tune table
        return (np.full(n_turns, 2 * np.pi * omega_x), np.full(n_turns, 2 * np.pi * omega_y))
    raise RuntimeError("Unknown modulation kind")


class Tables:
    """The four sin/cos tables of henon_tracker (henon.cu:2657-2694)."""

    def __init__(self, wx, wy):
        self.omega_x = _f64(wx)
        self.omega_y = _f64(wy)
        self.sx, self.cx = sin_cos(self.omega_x)
        self.sy, self.cy = sin_cos(self.omega_y)
        self.allowed_steps = self.omega_x.size

    def ptrs(self):
        return _p(self.sx), _p(self.cx), _p(self.sy), _p(self.cy)

    def interleaved(self):
        """[T][4] = (sin wx, cos wx, sin wy, cos wy): the product's table layout."""
        return np.ascontiguousarray(np.stack([self.sx, self.cx, self.sy, self.cy], axis=1))


# ---------------------------------------------------------------- particles + tracker
class Particles:
    """particles_4d state (henon.cu:699-735), host arrays. With shadow particles
    (add_ghost, henon.cu:737-817) the arrays hold the base particles, then one
    set of ghosts per add_ghost call."""

    def __init__(self, x, px, y, py):
        self.x, self.px, self.y, self.py = (_f64(a).copy() for a in (x, px, y, py))
        self.x0, self.px0, self.y0, self.py0 = (a.copy() for a in (self.x, self.px, self.y, self.py))
        self.steps = np.zeros(self.x.size, dtype=np.uint32)
        self.valid = np.ones(self.x.size, dtype=np.uint8)
        self.global_steps = 0
        self.n_base = self.x.size
        self.n_ghosts = 0
        self._kicks = 0  # random kicks drawn from the particle generator so far

    @property
    def n(self):
        return self.x.size

    def reset(self):
        # henon.cu:727-735: global_steps is not reset
        self.x, self.px, self.y, self.py = (a.copy() for a in (self.x0, self.px0, self.y0, self.py0))
        self.steps[:] = 0
        self.valid[:] = 1

    def state(self):
        return _p(self.x), _p(self.px), _p(self.y), _p(self.py), _p(self.steps)

    def add_ghost(self, module, direction):
        """henon.cu:737-817: one ghost per base particle, from its current state."""
        nb = self.n_base
        g = [a[:nb].copy() for a in (self.x, self.px, self.y, self.py)]
        axis = {"x": 0, "px": 1, "y": 2, "py": 3}
        if direction in axis:
            g[axis[direction]] = g[axis[direction]] + module
        elif direction == "random":
            lib().or_random_kicks(self._kicks, nb, float(module), *(_p(a) for a in g))
            self._kicks += nb
        else:
            raise RuntimeError("Invalid direction.")
        self.x, self.px, self.y, self.py = (np.concatenate((a, b)) for a, b in zip((self.x, self.px, self.y, self.py), g))
        self.x0, self.px0, self.y0, self.py0 = (np.concatenate((a, b)) for a, b in
                                                zip((self.x0, self.px0, self.y0, self.py0), g))
        self.steps = np.concatenate((self.steps, self.steps[:nb]))
        self.valid = np.concatenate((self.valid, self.valid[:nb]))
        self.n_ghosts += 1

    def renormalize(self, target):
        """henon.cu:826-858."""
        lib().or_ghost_renormalize(_p(self.x), _p(self.px), _p(self.y), _p(self.py), self.n_base, self.n_ghosts,
                                   float(target))

    def displacement_module(self):
        """henon.cu:860-897: [n_base][n_ghosts]."""
        out = np.empty((self.n_base, self.n_ghosts))
        lib().or_ghost_displacement(_p(self.x), _p(self.px), _p(self.y), _p(self.py), self.n_base, self.n_ghosts,
                                    _p(out), None)
        return out

    def displacement_direction(self):
        """henon.cu:899-952: [4][n_base][n_ghosts]."""
        out = np.empty((4, self.n_base, self.n_ghosts))
        lib().or_ghost_displacement(_p(self.x), _p(self.px), _p(self.y), _p(self.py), self.n_base, self.n_ghosts,
                                    None, _p(out))
        return out


def _deviations(p: "Particles", k: int):
    # henon_map_cpp/dynamic_indicators.py:198-222: (ghost - base) / norm
    n = p.n_base
    d = [a[(k + 1) * n:(k + 2) * n] - a[:n] for a in (p.x, p.px, p.y, p.py)]
    norm = np.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2 + d[3] ** 2)
    return [c / norm for c in d]


def philox4x32_10(counter, key):
    """Philox-4x32-10 block (Salmon et al., SC'11): four 32-bit words."""
    c = np.ascontiguousarray(counter, dtype=np.uint32).copy()
    lib().or_philox4x32_10(_p(c), int(key[0]), int(key[1]))
    return c


def kick_normals(seed, i, d):
    """The four Box-Muller normals of particle i's draw d."""
    out = np.empty(4)
    lib().or_kick_normals(seed, i, d, _p(out))
    return out


def sali(p: "Particles"):
    """henon_map_cpp/dynamic_indicators.py:198-224 at the current state, ghost
    sets 0 and 1 as the two displaced copies: min(|d1 + d2|, |d1 - d2|)."""
    d1, d2 = _deviations(p, 0), _deviations(p, 1)
    plus = np.sqrt((d1[0] + d2[0]) ** 2 + (d1[1] + d2[1]) ** 2 + (d1[2] + d2[2]) ** 2 + (d1[3] + d2[3]) ** 2)
    minus = np.sqrt((d1[0] - d2[0]) ** 2 + (d1[1] - d2[1]) ** 2 + (d1[2] - d2[2]) ** 2 + (d1[3] - d2[3]) ** 2)
    return np.min([plus, minus], axis=0)


def gali(p: "Particles"):
    """henon_map_cpp/dynamic_indicators.py:300-356 at the current state, ghost
    sets 0-3: the product of the singular values of the matrix of normalised
    deviations (rows x, px, y, py), NaN where it holds a NaN."""
    d = [_deviations(p, k) for k in range(4)]
    m = np.array([[d[c][r] for c in range(4)] for r in range(4)])  # [r][c][particle]
    m = np.moveaxis(m, 2, 0)
    ok = np.all(~np.isnan(m), axis=(1, 2))
    out = np.full(p.n_base, np.nan)
    if ok.any():
        out[ok] = np.prod(np.linalg.svd(m[ok], compute_uv=False), axis=-1)
    return out


class Tracker:
    """henon_tracker (CPU): tables + track/megno/birkhoff_tunes/all_tunes."""

    def __init__(self, n_turns, omega_x, omega_y, kind="sps", omega_0=np.nan, epsilon=0.0, offset=0,
                 coef=None, freq=None):
        wx, wy = pick_a_modulation(n_turns, omega_x, omega_y, kind, omega_0, epsilon, offset, coef, freq)
        self.t = Tables(wx, wy)

    @classmethod
    def from_tables(cls, tables: "Tables"):
        t = cls.__new__(cls)
        t.t = tables
        return t

    def _check(self, p, n_turns, inverse):
        # henon.cu:2706-2715
        if inverse:
            if n_turns > p.global_steps:
                raise RuntimeError("The number of turns is too large.")
        elif n_turns + p.global_steps > self.t.allowed_steps:
            raise RuntimeError("The number of turns is too large.")

    def _advance(self, p, n_turns, inverse):
        p.global_steps += -n_turns if inverse else n_turns

    def track_kicked(self, p, n_turns, mu, barrier, kick_module, seed, draw, inverse=False):
        """henon_step with the random kick (henon.cu:420-480), the kick's normals
        from Philox-4x32-10 keyed by seed at draw index draw + k (the product's
        generator); per-particle table indices, one thread."""
        self._check(p, n_turns, inverse)
        bad = lib().or_track_kicked(*p.state(), p.n, *self.t.ptrs(), self.t.allowed_steps, n_turns, mu, barrier,
                                    kick_module, int(inverse), seed, draw)
        if bad:
            raise RuntimeError("a turn index left the table")
        self._advance(p, n_turns, inverse)

    def track(self, p, n_turns, mu, barrier=100.0, kick_module=np.nan, inverse=False, n_threads=-1):
        self._check(p, n_turns, inverse)
        rc = lib().or_track(*p.state(), p.n, *self.t.ptrs(), n_turns, mu, barrier, kick_module, int(inverse),
                            n_threads)
        if rc:
            raise NotImplementedError("random kick is not restated in the oracle")
        self._advance(p, n_turns, inverse)

    def megno(self, p, n_turns, mu, barrier=100.0, inverse=False, turn_samples=(), n_threads=-1,
              carry_off_diagonal=True):
        self._check(p, n_turns, inverse)
        ts = _u32(list(turn_samples) if len(turn_samples) else [n_turns])
        out = np.empty((p.n, ts.size))
        rc = lib().or_megno(*p.state(), p.n, *self.t.ptrs(), n_turns, mu, barrier, int(inverse), _p(ts), ts.size,
                            n_threads, int(carry_off_diagonal), _p(out))
        if rc:
            raise ValueError("turn sample beyond n_turns")
        self._advance(p, n_turns, inverse)
        return out

    def birkhoff_tunes(self, p, n_turns, mu, barrier=100.0, inverse=False, from_idx=(), to_idx=(), n_threads=-1):
        self._check(p, n_turns, inverse)
        f, t = _u32(from_idx), _u32(to_idx)
        W = f.size + 1
        out = np.empty((p.n, 2 * (W + 1)))
        rc = lib().or_birkhoff_tunes(*p.state(), p.n, *self.t.ptrs(), n_turns, mu, barrier, int(inverse), _p(f),
                                     _p(t), f.size, n_threads, _p(out))
        if rc:
            raise RuntimeError("compute_tune: from and to must be valid indices")
        self._advance(p, n_turns, inverse)
        return out

    def all_tunes(self, p, n_turns, mu, barrier=100.0, inverse=False, from_idx=(), to_idx=(), n_threads=-1):
        self._check(p, n_turns, inverse)
        f, t = _u32(from_idx), _u32(to_idx)
        W = f.size + 1
        ob = np.empty((p.n, 2 * (W + 1)))
        of = np.empty((p.n, 2 * (W + 1)))
        rc = lib().or_all_tunes(*p.state(), p.n, *self.t.ptrs(), n_turns, mu, barrier, int(inverse), _p(f), _p(t),
                                f.size, n_threads, _p(ob), _p(of))
        if rc:
            raise RuntimeError("compute_tune: from and to must be valid indices")
        self._advance(p, n_turns, inverse)
        return ob, of

    def get_tangent_matrix(self, p, mu, reverse=False):
        out = np.empty((p.n, 4, 4))
        lib().or_get_tangent_matrix(_p(p.x), _p(p.px), _p(p.y), _p(p.py), _p(p.valid), p.n, *self.t.ptrs(),
                                    p.global_steps, mu, int(reverse), _p(out))
        return out


# ---------------------------------------------------------------- free functions
def tangent_matrix(x, px, y, py, sx, cx, sy, cy, mu):
    out = np.empty((4, 4))
    lib().or_tangent_matrix(x, px, y, py, sx, cx, sy, cy, mu, _p(out))
    return out


def inverse_tangent_matrix(x, px, y, py, sx, cx, sy, cy, mu):
    out = np.empty((4, 4))
    lib().or_inverse_tangent_matrix(x, px, y, py, sx, cx, sy, cy, mu, _p(out))
    return out


def multiply_matrices(a, b):
    a, b = _f64(a), _f64(b)
    out = np.empty((4, 4))
    lib().or_multiply_matrices(_p(a), _p(b), _p(out))
    return out


def birkhoff_weights(n):
    out = np.empty(n)
    lib().or_birkhoff_weights(n, _p(out))
    return out


def birkhoff_tune(x, px):
    x, px = _f64(x), _f64(px)
    return lib().or_birkhoff_tune(_p(x), _p(px), x.size)


def fft_magnitude(re, im, hanning=True):
    re, im = _f64(re), _f64(im)
    out = np.empty(re.size)
    lib().or_fft_magnitude(_p(re), _p(im), re.size, int(hanning), _p(out))
    return out


def interpolation(mag):
    mag = _f64(mag)
    return lib().or_interpolation(_p(mag), mag.size)


def fft_tune(x, px):
    x, px = _f64(x), _f64(px)
    return lib().or_fft_tune(_p(x), _p(px), x.size)


def get_tunes(x, px):
    x, px = _f64(x), _f64(px)
    out = np.empty(2)
    lib().or_get_tunes(_p(x), _p(px), x.size, _p(out))
    return out


def birkhoff_tune_vec(x, px, y, py, from_idx=(), to_idx=()):
    x, px, y, py = (_f64(a) for a in (x, px, y, py))
    f, t = _u32(from_idx), _u32(to_idx)
    out = np.empty(2 * (f.size + 1))
    if lib().or_birkhoff_tune_vec(_p(x), _p(px), _p(y), _p(py), x.size, _p(f), _p(t), f.size, _p(out)):
        raise RuntimeError("compute_tune: from and to must be valid indices")
    return out


def fft_tune_vec(x, px, y, py, from_idx=(), to_idx=()):
    x, px, y, py = (_f64(a) for a in (x, px, y, py))
    f, t = _u32(from_idx), _u32(to_idx)
    out = np.empty(2 * (f.size + 1))
    if lib().or_fft_tune_vec(_p(x), _p(px), _p(y), _p(py), x.size, _p(f), _p(t), f.size, _p(out)):
        raise RuntimeError("compute_tune: from and to must be valid indices")
    return out


def nan_mean(v):
    v = _f64(v)
    return lib().or_nan_mean(_p(v), v.size)


def nan_std(v):
    v = _f64(v)
    return lib().or_nan_std(_p(v), v.size)


def structured_multiply(mats, x, px, y, py, tables: Tables, global_steps, mu):
    mats = np.ascontiguousarray(mats, dtype=np.float64).copy()
    x, px, y, py = (_f64(a) for a in (x, px, y, py))
    lib().or_structured_multiply(_p(mats), _p(x), _p(px), _p(y), _p(py), x.size, *tables.ptrs(), global_steps, mu)
    return mats


def lyap_birk_add(values, values_b, vectors, weights, n_weights, b_idx):
    vectors, weights = _f64(vectors), _f64(weights)
    lib().or_lyap_birk_add(_p(values), _p(values_b), _p(vectors), _p(weights), values.size, n_weights, b_idx)


def megno_construct_add(l1, l2, l3, a, b, idx):
    a, b = _f64(a), _f64(b)
    lib().or_megno_construct_add(_p(l1), _p(l2), _p(l3), _p(a), _p(b), l1.size, idx)


class _SampleWeights:
    """birkhoff_weights(turn_samples[j]) concatenated, with their offsets (the
    per-sample weight arrays of henon.cu:2302-2334 and :2428-2474)."""

    def __init__(self, turn_samples):
        self.ts = np.ascontiguousarray(turn_samples, dtype=np.uintp)
        ws = [birkhoff_weights(int(t)) for t in self.ts]
        self.off = np.ascontiguousarray(np.cumsum([0] + [len(w) for w in ws[:-1]]), dtype=np.uintp)
        self.w = np.ascontiguousarray(np.concatenate(ws) if ws else np.zeros(1))


class MegnoConstruct:
    """megno_construct (henon.cu:2223-2299): idx starts at 1, add(a, b) runs
    kernel_megno's formula, get_values() is layer 3."""

    def __init__(self, n):
        self.n = n
        self.reset()

    def reset(self):
        self.idx = 1
        self.l1, self.l2, self.l3 = np.zeros(self.n), np.zeros(self.n), np.zeros(self.n)

    def add(self, a, b):
        megno_construct_add(self.l1, self.l2, self.l3, np.reshape(a, (self.n, 16)), np.reshape(b, (self.n, 16)),
                            self.idx)
        self.idx += 1

    def get_values(self):
        return self.l3.copy()


class MegnoBirkhoffConstruct:
    """megno_birkhoff_construct (henon.cu:2302-2426): layer 1 as megno_construct,
    layer 2 per turn sample j: += L1 w_j[idx-1] / idx while idx <= T_j."""

    def __init__(self, n, turn_samples):
        self.n = n
        self.sw = _SampleWeights(turn_samples)
        self.reset()

    def reset(self):
        self.idx = 1
        self.l1 = np.zeros(self.n)
        self.l2 = np.zeros((len(self.sw.ts), self.n))

    def add(self, a, b):
        a, b = _f64(np.reshape(a, (self.n, 16))), _f64(np.reshape(b, (self.n, 16)))
        lib().or_megno_birkhoff_add(_p(self.l1), _p(self.l2), _p(self.sw.w), _p(self.sw.off), _p(self.sw.ts),
                                    len(self.sw.ts), _p(a), _p(b), self.n, self.idx)
        self.idx += 1

    def get_values(self):
        return self.l2.copy()


class TuneBirkhoffConstruct:
    """tune_birkhoff_construct (henon.cu:2428-2655): first_add stores the phase
    angles, each add accumulates the wrapped phase advance with w_j[idx] into
    tune 1 (idx < T_j) or w_j[idx - T_j] into tune 2 (T_j <= idx < 2 T_j)."""

    def __init__(self, n, turn_samples):
        self.n = n
        self.sw = _SampleWeights(turn_samples)
        self.sx, self.sy = np.zeros(n), np.zeros(n)
        self.reset()

    def reset(self):
        self.idx = 1
        ns = len(self.sw.ts)
        self.t1x, self.t2x = np.zeros((ns, self.n)), np.zeros((ns, self.n))
        self.t1y, self.t2y = np.zeros((ns, self.n)), np.zeros((ns, self.n))

    def first_add(self, x, px, y, py):
        x, px, y, py = (_f64(a) for a in (x, px, y, py))
        lib().or_tune_birkhoff_init(_p(x), _p(px), _p(y), _p(py), _p(self.sx), _p(self.sy), self.n)

    def add(self, x, px, y, py):
        x, px, y, py = (_f64(a) for a in (x, px, y, py))
        lib().or_tune_birkhoff_add(_p(x), _p(px), _p(y), _p(py), _p(self.sx), _p(self.sy), _p(self.t1x),
                                   _p(self.t2x), _p(self.t1y), _p(self.t2y), _p(self.sw.w), _p(self.sw.off),
                                   _p(self.sw.ts), len(self.sw.ts), self.n, self.idx)
        self.idx += 1


def matvec(m, v):
    m = _f64(m)
    v = _f64(v).copy()
    lib().or_matvec(_p(m), _p(v), v.shape[0])
    return v


def normalize(v):
    v = _f64(v).copy()
    lib().or_normalize(_p(v), v.shape[0])
    return v


# ---------------------------------------------------------------- fused indicators
def indicators(p: Particles, tables: Tables, n_turns, mu, barrier=10.0, renorm_every=0, samples=(),
               fli_dir=(1.0, 0.0, 0.0, 0.0), matrix=None, log_scale=None, birkhoff=(), n_threads=-1):
    """CPU restatement of the fused indicator kernel (exact mode); returns a dict
    of megno, fli [n_samples][n], matrix [n][4][4], log_scale [n], birkhoff
    [n_windows][2][n]. Advances p like track()."""
    n = p.n
    s = _u32(samples)
    megno = np.empty((s.size, n))
    fli = np.empty((s.size, n))
    m = np.ascontiguousarray(np.tile(np.eye(4), (n, 1, 1)) if matrix is None else np.array(matrix, dtype=np.float64))
    ls = np.ascontiguousarray(np.zeros(n) if log_scale is None else np.array(log_scale, dtype=np.float64))
    if m.shape != (n, 4, 4) or ls.shape != (n,):
        raise ValueError("matrix must be [n][4][4] and log_scale [n]")
    bf = _u32([w[0] for w in birkhoff])
    bt = _u32([w[1] for w in birkhoff])
    bout = np.empty((bf.size, 2, n))
    fd = _f64(fli_dir)
    rc = lib().or_indicators(*p.state(), n, *tables.ptrs(), tables.allowed_steps, n_turns, mu, barrier,
                             renorm_every, _p(s), s.size, _p(fd), _p(megno), _p(fli), _p(m), _p(ls), _p(bf),
                             _p(bt), bf.size, _p(bout), n_threads)
    if rc:
        raise RuntimeError("table index out of range")
    p.global_steps += n_turns
    return dict(megno=megno, fli=fli, matrix=m, log_scale=ls, birkhoff=bout)
