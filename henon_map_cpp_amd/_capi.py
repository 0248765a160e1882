"""ctypes binding of the C ABI in include/henon_hip.h (libhenon_hip.so).

This is the binding a Python caller of the C ABI uses directly; the pybind11
module ``henon_map_engine`` is the drop-in for the reference's module and sits
on the same C ABI. Loading fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libhenon_hip.so")

HM_OK, HM_ERR_ARG, HM_ERR_RANGE, HM_ERR_HIP, HM_ERR_UNSUPPORTED, HM_ERR_ALLOC = range(6)
MODE_EXACT, MODE_FAST, MODE_EXACT_ORBIT = 0, 1, 2  # exact orbit: hm_track_indicators only
OBS_RADIUS, OBS_ACTION, OBS_ACTION_X, OBS_ACTION_Y, OBS_ANGLE_X, OBS_ANGLE_Y = range(6)

D = ctypes.c_double
P = ctypes.c_void_p
SZ = ctypes.c_size_t
I = ctypes.c_int
U32 = ctypes.c_uint32
U64 = ctypes.c_uint64


class LaunchCfg(ctypes.Structure):
    _fields_ = [("block", ctypes.c_int32), ("compact_every", ctypes.c_int32),
                ("particles_per_lane", ctypes.c_int32), ("table_staging", ctypes.c_int32)]


class Job(ctypes.Structure):
    """hm_job (include/henon_hip.h)."""
    _fields_ = [("x", ctypes.c_void_p), ("px", ctypes.c_void_p), ("y", ctypes.c_void_p), ("py", ctypes.c_void_p),
                ("steps", ctypes.c_void_p), ("n", ctypes.c_size_t), ("table", ctypes.c_void_p),
                ("table_len", ctypes.c_size_t), ("n_turns", ctypes.c_uint32), ("renorm_every", ctypes.c_uint32),
                ("mu", ctypes.c_double), ("barrier", ctypes.c_double), ("kick_module", ctypes.c_double),
                ("megno", ctypes.c_void_p)]


class BatchCfg(ctypes.Structure):
    """hm_batch_cfg (include/henon_hip.h)."""
    _fields_ = [("block", ctypes.c_int32), ("compact_every", ctypes.c_int32), ("waves", ctypes.c_int32),
                ("use_graph", ctypes.c_int32)]


MAX_SAMPLES, MAX_BIRKHOFF = 16, 4


class IndicatorSpec(ctypes.Structure):
    """hm_indicator_spec (include/henon_hip.h)."""
    _fields_ = [("renorm_every", ctypes.c_uint32), ("n_samples", ctypes.c_uint32),
                ("samples", ctypes.c_uint32 * MAX_SAMPLES), ("fli_dir", ctypes.c_double * 4),
                ("megno", ctypes.c_void_p), ("fli", ctypes.c_void_p), ("matrix", ctypes.c_void_p),
                ("log_scale", ctypes.c_void_p), ("n_birkhoff", ctypes.c_uint32),
                ("birk_from", ctypes.c_uint32 * MAX_BIRKHOFF), ("birk_to", ctypes.c_uint32 * MAX_BIRKHOFF),
                ("birk_out", ctypes.c_void_p), ("hist_from", ctypes.c_uint32), ("hist_len", ctypes.c_uint32),
                ("hist", ctypes.c_void_p), ("n_megno_birkhoff", ctypes.c_uint32),
                ("megno_birkhoff_samples", ctypes.c_uint32 * MAX_BIRKHOFF), ("megno_birkhoff_out", ctypes.c_void_p),
                ("n_tune_birkhoff", ctypes.c_uint32), ("tune_birkhoff_samples", ctypes.c_uint32 * MAX_BIRKHOFF),
                ("tune_birkhoff_out", ctypes.c_void_p)]


# name -> (restype, argtypes); mirrors include/henon_hip.h
SIGNATURES = {
    "hm_last_error": (ctypes.c_char_p, []),
    "hm_version": (ctypes.c_char_p, []),
    "hm_device_count": (I, [ctypes.POINTER(I)]),
    "hm_mem_info": (I, [ctypes.POINTER(SZ), ctypes.POINTER(SZ)]),
    "hm_malloc": (I, [ctypes.POINTER(P), SZ]),
    "hm_free": (I, [P]),
    "hm_memcpy": (I, [P, P, SZ, I, P]),
    "hm_memset": (I, [P, I, SZ, P]),
    "hm_stream_synchronize": (I, [P]),
    "hm_take_status": (I, [ctypes.POINTER(U32), P]),
    "hm_basic_modulation": (I, [D, D, D, I, I, P]),
    "hm_sps_modulation": (I, [D, D, I, I, P]),
    "hm_lines_modulation": (I, [D, D, P, P, I, I, I, P]),
    "hm_sps_lines": (I, [P, P]),
    "hm_gaussian_modulation": (I, [D, D, I, I, U64, P]),
    "hm_uniform_modulation": (I, [D, D, I, I, U64, P]),
    "hm_build_table": (I, [P, P, SZ, P]),
    "hm_birkhoff_weights": (I, [U32, P]),
    "hm_track": (I, [P, P, P, P, P, SZ, P, SZ, U32, D, D, D, I, I, ctypes.POINTER(LaunchCfg), P]),
    "hm_track_noalloc": (I, [P, P, P, P, P, SZ, P, SZ, U32, D, D, D, I, I, ctypes.POINTER(LaunchCfg), P]),
    "hm_track_batch": (I, [ctypes.POINTER(Job), SZ, I, ctypes.POINTER(BatchCfg), P]),
    "hm_autotune_batch": (I, [ctypes.POINTER(Job), SZ, I, U32, ctypes.POINTER(BatchCfg), ctypes.POINTER(D), P]),
    "hm_autotune_batch_clear": (I, []),
    "hm_autotune_track": (I, [P, P, P, P, P, SZ, P, SZ, U32, U32, D, D, I, ctypes.POINTER(LaunchCfg),
                              ctypes.POINTER(D), P]),
    "hm_autotune_clear": (I, []),
    "hm_observable": (I, [I, P, P, P, P, SZ, P, P]),
    "hm_tangent_matrix": (I, [P, P, P, P, P, SZ, P, SZ, U32, D, I, P, P]),
    "hm_megno_trace": (I, [P, P, P, P, P, SZ, P, SZ, U32, D, D, I, P, SZ, P, P, P, P]),
    "hm_track_indicators": (I, [P, P, P, P, P, SZ, P, SZ, U32, D, D, I, ctypes.POINTER(IndicatorSpec), P]),
    "hm_tangent_apply": (I, [P, P, P, P, P, SZ, P, SZ, U32, D, I, P]),
    "hm_matrix_left_multiply": (I, [P, P, SZ, P]),
    "hm_matrix_identity": (I, [P, SZ, P]),
    "hm_matvec": (I, [P, P, P, SZ, I, P]),
    "hm_normalize": (I, [P, SZ, P]),
    "hm_lyapunov_add": (I, [P, P, P, P, SZ, U64, U64, P]),
    "hm_megno_add": (I, [P, P, P, P, P, SZ, U64, P]),
    "hm_megno_birkhoff_add": (I, [P, P, P, P, SZ, U64, P, P, P, U32, P]),
    "hm_tune_birkhoff_add": (I, [P, P, P, P, P, P, P, P, P, P, P, P, P, U32, SZ, U64, I, P]),
    "hm_track_history": (I, [P, P, P, P, P, SZ, P, SZ, U32, D, D, I, P, P]),
    "hm_history_means": (I, [P, SZ, U32, P, P]),
    "hm_birkhoff_tunes_history": (I, [P, SZ, U32, P, P, P, SZ, P, P]),
    "hm_fft_tunes": (I, [P, SZ, U32, P, P, P, SZ, P, P]),
    "hm_track_rows": (I, [P, P, P, P, P, SZ, P, SZ, U32, D, D, I, P, SZ, P, P]),
    "hm_rows_means": (I, [P, SZ, SZ, U32, P, P]),
    "hm_birkhoff_tunes_rows": (I, [P, SZ, SZ, U32, P, P, P, SZ, P, P]),
    "hm_fft_tunes_rows": (I, [P, SZ, SZ, U32, P, P, P, SZ, P, P]),
    "hm_tunes_rows": (I, [P, SZ, SZ, U32, P, P, P, SZ, P, P, P]),
    "hm_adjoint_forward": (I, [P, P, P, P, SZ, P, SZ, U32, D, D, U32, P, P, P, P, P, I, P]),
    "hm_track_kicked": (I, [P, P, P, P, P, SZ, P, SZ, U32, D, D, D, I, I, ctypes.c_uint64, ctypes.c_uint64, P]),
    "hm_ghost_renormalize": (I, [P, P, P, P, SZ, SZ, D, P]),
    "hm_ghost_displacement": (I, [P, P, P, P, SZ, SZ, P, P, P]),
    "hm_ghost_alignment": (I, [P, P, P, P, SZ, SZ, I, P, P]),
    "hm_adjoint_scratch_bytes": (I, [SZ, U32, U32, ctypes.POINTER(SZ)]),
    "hm_adjoint_backward": (I, [P, SZ, P, P, SZ, U32, D, U32, P, P, P, P, P, SZ, P, P, I, P]),
    "hm_lines_modulation_derivatives": (I, [D, D, P, P, I, I, I, P]),
    "hm_profile_enable": (I, [I]),
    "hm_profile_reset": (I, []),
    "hm_profile_read": (I, [ctypes.c_char_p, ctypes.POINTER(D), ctypes.POINTER(U64)]),
}

_lib = None


def _preload_runtime():
    # torch ships its own libamdhip64 (same SONAME); importing it first makes
    # this library bind to that runtime, so one HIP runtime serves the process.
    try:
        import torch  # noqa: F401
    except Exception:
        pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m henon_map_cpp_amd.build`")
        _preload_runtime()
        L = ctypes.CDLL(LIB_PATH)
        missing = [name for name in SIGNATURES if not hasattr(L, name)]
        if missing:
            raise ImportError(f"{LIB_PATH} is stale (missing {missing[:3]}...); rebuild it")
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class HMError(RuntimeError):
    pass


def check(rc: int) -> None:
    if rc == HM_OK:
        return
    msg = lib().hm_last_error().decode(errors="replace")
    if rc == HM_ERR_ARG:
        raise ValueError(msg)
    if rc == HM_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    if rc == HM_ERR_ALLOC:
        raise MemoryError(msg)
    raise RuntimeError(msg)


def device_count() -> int:
    c = I(0)
    check(lib().hm_device_count(ctypes.byref(c)))
    return c.value


def ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class DeviceBuffer:
    """A device allocation made through hm_malloc (numpy in, numpy out)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = P()
        check(lib().hm_malloc(ctypes.byref(p), max(self.nbytes, 1)))
        self.ptr = p

    @classmethod
    def from_array(cls, a: np.ndarray, stream=None):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a, stream)
        return b

    def upload(self, a: np.ndarray, stream=None):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        check(lib().hm_memcpy(self.ptr, ptr(a), a.nbytes, 0, stream))
        check(lib().hm_stream_synchronize(stream))

    def download(self, dtype, count, stream=None):
        out = np.empty(count, dtype=dtype)
        check(lib().hm_memcpy(ptr(out), self.ptr, out.nbytes, 1, stream))
        check(lib().hm_stream_synchronize(stream))
        return out

    def free(self):
        if self.ptr:
            check(lib().hm_free(self.ptr))
            self.ptr = P()

    def __del__(self):
        try:
            if self.ptr and _lib is not None:
                _lib.hm_free(self.ptr)
        except Exception:
            pass


# ---------------------------------------------------------------- host helpers
def sps_modulation(tune, epsilon, start, end):
    out = np.empty(end - start)
    check(lib().hm_sps_modulation(tune, epsilon, start, end, ptr(out)))
    return out


def basic_modulation(tune, omega, epsilon, start, end):
    out = np.empty(end - start)
    check(lib().hm_basic_modulation(tune, omega, epsilon, start, end, ptr(out)))
    return out


def lines_modulation(tune, epsilon, coef, freq, start, end):
    coef = np.ascontiguousarray(coef, dtype=np.float64)
    freq = np.ascontiguousarray(freq, dtype=np.float64)
    out = np.empty(end - start)
    check(lib().hm_lines_modulation(tune, epsilon, ptr(coef), ptr(freq), coef.size, start, end, ptr(out)))
    return out


def sps_lines():
    c, f = np.empty(7), np.empty(7)
    check(lib().hm_sps_lines(ptr(c), ptr(f)))
    return c, f


def build_table(omega_x, omega_y):
    wx = np.ascontiguousarray(omega_x, dtype=np.float64)
    wy = np.ascontiguousarray(omega_y, dtype=np.float64)
    out = np.empty((wx.size, 4))
    check(lib().hm_build_table(ptr(wx), ptr(wy), wx.size, ptr(out)))
    return out


def birkhoff_weights(n):
    out = np.empty(n)
    check(lib().hm_birkhoff_weights(n, ptr(out)))
    return out


# ---------------------------------------------------------------- device path
class DeviceParticles:
    """Device SoA (x, px, y, py, steps) held through the C ABI."""

    def __init__(self, x, px, y, py, steps=None):
        n = len(x)
        self.n = n
        self.bufs = [DeviceBuffer.from_array(np.asarray(a, dtype=np.float64)) for a in (x, px, y, py)]
        st = np.zeros(n, dtype=np.uint32) if steps is None else np.asarray(steps, dtype=np.uint32)
        self.steps = DeviceBuffer.from_array(st)

    def ptrs(self):
        return [b.ptr for b in self.bufs] + [self.steps.ptr]

    def get(self):
        x, px, y, py = (b.download(np.float64, self.n) for b in self.bufs)
        return x, px, y, py, self.steps.download(np.uint32, self.n)


def track(parts: DeviceParticles, table_dev: DeviceBuffer, table_len: int, n_turns: int, mu: float,
          barrier: float, inverse=False, mode=MODE_EXACT, cfg: LaunchCfg | None = None, stream=None,
          kick_module=float("nan")):
    c = ctypes.byref(cfg) if cfg is not None else None
    check(lib().hm_track(*parts.ptrs(), parts.n, table_dev.ptr, table_len, n_turns, mu, barrier, kick_module,
                         int(inverse), mode, c, stream))


def _job_array(jobs):
    arr = (Job * max(len(jobs), 1))()
    for k, job in enumerate(jobs):
        parts, tdev, tlen, n_turns, mu, barrier = job[:6]
        megno = job[6].ptr if len(job) > 6 and job[6] is not None else None
        renorm = job[7] if len(job) > 7 else 0
        x, px, y, py, st = parts.ptrs()
        arr[k] = Job(x, px, y, py, st, parts.n, tdev.ptr, tlen, n_turns, renorm, mu, barrier, float("nan"), megno)
    return arr


def track_batch(jobs, mode=MODE_EXACT, cfg: BatchCfg | None = None, stream=None):
    """jobs: list of (DeviceParticles, table DeviceBuffer, table_len, n_turns, mu, barrier[, megno
    DeviceBuffer, renorm_every]); with a megno buffer the job also accumulates the MEGNO."""
    c = ctypes.byref(cfg) if cfg is not None else None
    check(lib().hm_track_batch(_job_array(jobs), len(jobs), mode, c, stream))


def autotune_batch(jobs, mode=MODE_EXACT, trial_turns=4000, stream=None):
    """Times batch configurations on copies of the jobs (each cut to trial_turns
    turns); returns (BatchCfg, ms). Later track_batch(..., cfg=None) calls of
    the same job-mix class use it."""
    best, ms = BatchCfg(), D()
    check(lib().hm_autotune_batch(_job_array(jobs), len(jobs), mode, trial_turns, ctypes.byref(best),
                                  ctypes.byref(ms), stream))
    return best, ms.value


def autotune_track(parts: DeviceParticles, table_dev: DeviceBuffer, table_len: int, mu: float, barrier: float,
                   mode=MODE_EXACT, warm_turns=2000, trial_turns=4000, stream=None):
    """Returns (LaunchCfg, trial ms); hm_track with cfg=None then uses it."""
    best, ms = LaunchCfg(), D(0.0)
    check(lib().hm_autotune_track(*parts.ptrs(), parts.n, table_dev.ptr, table_len, warm_turns, trial_turns, mu,
                                  barrier, mode, ctypes.byref(best), ctypes.byref(ms), stream))
    return best, ms.value


def take_status(stream=None) -> int:
    f = U32(0)
    check(lib().hm_take_status(ctypes.byref(f), stream))
    return f.value


def profile_enable(on=True):
    check(lib().hm_profile_enable(int(on)))


def profile_reset():
    check(lib().hm_profile_reset())


def profile_read(name: str):
    ms, n = D(0.0), U64(0)
    check(lib().hm_profile_read(name.encode() if name else None, ctypes.byref(ms), ctypes.byref(n)))
    return ms.value, n.value


def megno_trace(parts: DeviceParticles, table_dev: DeviceBuffer, table_len: int, n_turns: int, mu: float,
                barrier: float, samples, stream=None, inverse=False, matrix_in=None, matrix_out=False):
    """Trace-form MEGNO [n][n_samples]; with matrix_out=True also the final
    matrices [n][4][4], returned as (megno, matrices)."""
    s = np.ascontiguousarray(samples, dtype=np.uint32)
    out = DeviceBuffer(8 * parts.n * max(s.size, 1))
    mi = DeviceBuffer.from_array(np.ascontiguousarray(matrix_in, dtype=np.float64)) if matrix_in is not None else None
    mo = DeviceBuffer(8 * 16 * parts.n) if matrix_out else None
    check(lib().hm_megno_trace(*parts.ptrs(), parts.n, table_dev.ptr, table_len, n_turns, mu, barrier, int(inverse),
                               ptr(s), s.size, mi.ptr if mi is not None else None, mo.ptr if mo is not None else None,
                               out.ptr, stream))
    m = out.download(np.float64, parts.n * s.size).reshape(parts.n, s.size)
    return (m, mo.download(np.float64, 16 * parts.n).reshape(parts.n, 4, 4)) if matrix_out else m


def track_indicators(parts: DeviceParticles, table_dev: DeviceBuffer, table_len: int, n_turns: int, mu: float,
                     barrier: float, mode=MODE_EXACT, renorm_every=0, samples=(), fli_dir=(1.0, 0.0, 0.0, 0.0),
                     matrix=None, log_scale=None, birkhoff=(), hist=None, stream=None, megno_birkhoff=(),
                     tune_birkhoff=()):
    """Fused indicators; returns numpy results (megno, fli [n_samples][n], matrix
    [n][4][4] when `matrix` is given (array of initial matrices or True for the
    identity), log_scale [n], birkhoff [windows][2][n], hist [len][4][n],
    megno_birkhoff [len][n], tune_birkhoff [len][4][n])."""
    n = parts.n
    spec = IndicatorSpec()
    spec.renorm_every = renorm_every
    spec.n_samples = len(samples)
    for q, v in enumerate(samples):
        spec.samples[q] = v
    for q in range(4):
        spec.fli_dir[q] = fli_dir[q]
    keep = []
    out = {}
    if len(samples):
        mg, fl = DeviceBuffer(8 * n * len(samples)), DeviceBuffer(8 * n * len(samples))
        keep += [mg, fl]
        spec.megno, spec.fli = mg.ptr, fl.ptr
    if matrix is not None:
        m0 = np.tile(np.eye(4), (n, 1, 1)) if matrix is True else np.asarray(matrix, dtype=np.float64)
        md = DeviceBuffer.from_array(np.ascontiguousarray(m0))
        ld = DeviceBuffer.from_array(np.zeros(n) if log_scale is None else np.asarray(log_scale, np.float64))
        keep += [md, ld]
        spec.matrix, spec.log_scale = md.ptr, ld.ptr
    spec.n_birkhoff = len(birkhoff)
    for q, (a, b) in enumerate(birkhoff):
        spec.birk_from[q], spec.birk_to[q] = a, b
    if len(birkhoff):
        bo = DeviceBuffer(8 * n * 2 * len(birkhoff))
        keep.append(bo)
        spec.birk_out = bo.ptr
    if hist is not None:
        h0, hl = hist
        hd = DeviceBuffer(8 * n * 4 * hl)
        keep.append(hd)
        spec.hist_from, spec.hist_len, spec.hist = h0, hl, hd.ptr
    spec.n_megno_birkhoff = len(megno_birkhoff)
    for q, v in enumerate(megno_birkhoff):
        spec.megno_birkhoff_samples[q] = v
    if len(megno_birkhoff):
        mbo = DeviceBuffer(8 * n * len(megno_birkhoff))
        keep.append(mbo)
        spec.megno_birkhoff_out = mbo.ptr
    spec.n_tune_birkhoff = len(tune_birkhoff)
    for q, v in enumerate(tune_birkhoff):
        spec.tune_birkhoff_samples[q] = v
    if len(tune_birkhoff):
        tbo = DeviceBuffer(8 * n * 4 * len(tune_birkhoff))
        keep.append(tbo)
        spec.tune_birkhoff_out = tbo.ptr
    check(lib().hm_track_indicators(*parts.ptrs(), n, table_dev.ptr, table_len, n_turns, mu, barrier, mode,
                                    ctypes.byref(spec), stream))
    if len(samples):
        out["megno"] = mg.download(np.float64, n * len(samples)).reshape(len(samples), n)
        out["fli"] = fl.download(np.float64, n * len(samples)).reshape(len(samples), n)
    if matrix is not None:
        out["matrix"] = md.download(np.float64, n * 16).reshape(n, 4, 4)
        out["log_scale"] = ld.download(np.float64, n)
    if len(birkhoff):
        out["birkhoff"] = bo.download(np.float64, n * 2 * len(birkhoff)).reshape(len(birkhoff), 2, n)
    if hist is not None:
        out["hist"] = hd.download(np.float64, n * 4 * hist[1]).reshape(hist[1], 4, n)
    if len(megno_birkhoff):
        out["megno_birkhoff"] = mbo.download(np.float64, n * len(megno_birkhoff)).reshape(len(megno_birkhoff), n)
    if len(tune_birkhoff):
        out["tune_birkhoff"] = tbo.download(np.float64, n * 4 * len(tune_birkhoff)).reshape(len(tune_birkhoff), 4, n)
    return out


def track_history(parts: DeviceParticles, table_dev: DeviceBuffer, table_len: int, n_turns: int, mu: float,
                  barrier: float, inverse=False, stream=None) -> DeviceBuffer:
    hist = DeviceBuffer(8 * parts.n * 4 * (n_turns + 1))
    check(lib().hm_track_history(*parts.ptrs(), parts.n, table_dev.ptr, table_len, n_turns, mu, barrier,
                                 int(inverse), hist.ptr, stream))
    return hist


def row_length(n_turns: int) -> int:
    """Samples per particle row of hm_track_rows: n_turns + 1 rounded up to 16."""
    return (n_turns + 1 + 15) // 16 * 16


def track_rows(parts: DeviceParticles, table_dev: DeviceBuffer, table_len: int, n_turns: int, mu: float,
               barrier: float, inverse=False, stream=None, means=False):
    """The trajectory as particle rows [n][4][row_length(n_turns)]; with
    means=True also the per-coordinate means [4][n] over samples 0..n_turns,
    returned as (rows, means)."""
    L = row_length(n_turns)
    rows = DeviceBuffer(8 * parts.n * 4 * L)
    mb = DeviceBuffer(8 * 4 * parts.n) if means else None
    check(lib().hm_track_rows(*parts.ptrs(), parts.n, table_dev.ptr, table_len, n_turns, mu, barrier, int(inverse),
                              rows.ptr, L, mb.ptr if mb is not None else None, stream))
    return (rows, mb) if means else rows


def history_tunes(hist: DeviceBuffer, n: int, n_samples: int, windows, kind="fft", subtract_mean=False,
                  stream=None, row_len=None, means: DeviceBuffer | None = None):
    """Tunes of windows [(from, to)] (inclusive) of a device history
    (sample-major, or particle rows of length row_len); returns numpy
    [n_windows][2][n]. subtract_mean computes the per-coordinate means of the
    n_samples samples first; `means` passes ones already computed (track_rows)."""
    L = lib()
    if means is None and subtract_mean:
        means = DeviceBuffer(8 * 4 * n)
        if row_len is None:
            check(L.hm_history_means(hist.ptr, n, n_samples, means.ptr, stream))
        else:
            check(L.hm_rows_means(hist.ptr, n, row_len, n_samples, means.ptr, stream))
    f = np.ascontiguousarray([w[0] for w in windows], dtype=np.uint32)
    t = np.ascontiguousarray([w[1] for w in windows], dtype=np.uint32)
    out = DeviceBuffer(8 * n * 2 * len(windows))
    mp = means.ptr if means is not None else None
    if kind == "fft":
        ln = np.ascontiguousarray(t - f + 1, dtype=np.uint32)
        if row_len is None:
            check(L.hm_fft_tunes(hist.ptr, n, n_samples, mp, ptr(f), ptr(ln), len(windows), out.ptr, stream))
        else:
            check(L.hm_fft_tunes_rows(hist.ptr, n, row_len, n_samples, mp, ptr(f), ptr(ln), len(windows), out.ptr,
                                      stream))
    else:
        if row_len is None:
            check(L.hm_birkhoff_tunes_history(hist.ptr, n, n_samples, mp, ptr(f), ptr(t), len(windows), out.ptr,
                                              stream))
        else:
            check(L.hm_birkhoff_tunes_rows(hist.ptr, n, row_len, n_samples, mp, ptr(f), ptr(t), len(windows),
                                           out.ptr, stream))
    return out.download(np.float64, n * 2 * len(windows)).reshape(len(windows), 2, n)


def tunes_rows(rows: DeviceBuffer, n: int, n_samples: int, windows, row_len: int, means: DeviceBuffer | None = None,
               stream=None):
    """Birkhoff and FFT tunes of the same windows [(from, to)] of particle
    rows in one call (hm_tunes_rows): numpy ([n_windows][2][n], [n_windows][2][n])."""
    f = np.ascontiguousarray([w[0] for w in windows], dtype=np.uint32)
    t = np.ascontiguousarray([w[1] for w in windows], dtype=np.uint32)
    ob = DeviceBuffer(8 * n * 2 * len(windows))
    of = DeviceBuffer(8 * n * 2 * len(windows))
    mp = means.ptr if means is not None else None
    check(lib().hm_tunes_rows(rows.ptr, n, row_len, n_samples, mp, ptr(f), ptr(t), len(windows), ob.ptr, of.ptr,
                              stream))
    shape = (len(windows), 2, n)
    return (ob.download(np.float64, n * 2 * len(windows)).reshape(shape),
            of.download(np.float64, n * 2 * len(windows)).reshape(shape))
