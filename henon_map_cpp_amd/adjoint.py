"""Reverse-mode derivatives of n-turn tracking (device adjoint, C ABI).

The reference differentiates one turn at a time (the tangent matrix of
src/henon.cu:175-265, products of which give MEGNO/FLI). This module gives
whole-orbit gradients of a scalar loss L(s_T, B), where s_T is the state after
T turns and B = sum_k w_k x_{k+1}^2 is the Birkhoff-weighted time average of
x^2, with respect to the initial conditions and the modulation parameters
(epsilon, mu, tune_x, tune_y), at the cost of about three forward passes
(checkpointed recomputation; the map is never inverted).

Device memory and the stream come from torch; every kernel is in
libhenon_hip.so (hm_adjoint_forward / hm_adjoint_backward).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _capi as C

MODES = {"exact": C.MODE_EXACT, "fast": C.MODE_FAST}


def lines_derivatives(tune, epsilon, coef, freq, start, end) -> np.ndarray:
    """[(d omega_k / d omega_0, d omega_k / d epsilon)] for the line modulation
    omega_k = 2 pi tune (1 + epsilon sum_j coef_j cos(freq_j (k + start)))."""
    coef = np.ascontiguousarray(coef, dtype=np.float64)
    freq = np.ascontiguousarray(freq, dtype=np.float64)
    out = np.empty((end - start, 2))
    C.check(C.lib().hm_lines_modulation_derivatives(tune, epsilon, C.ptr(coef), C.ptr(freq), coef.size, start, end,
                                                    C.ptr(out)))
    return out


def default_every(n_turns: int) -> int:
    """Checkpoint interval K ~ sqrt(8 T), a multiple of 8: ceil(T/K) checkpoint
    states per particle, and a window of one segment (K/8 states per particle)
    is the least scratch the backward sweep can run in."""
    k = int(math.sqrt(8.0 * max(n_turns, 1)))
    return max(8, (k + 7) // 8 * 8)


def _p(t: torch.Tensor | None):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class OrbitAdjoint:
    """Tables and derivative tables of one modulated lattice, on the device.

    The modulation is the line set of hm_lines_modulation (the SPS lines of
    hm_sps_lines by default, as henon_tracker's "sps" kind, src/modulation.cpp),
    evaluated for turns start..start+n_turns-1 with tune_x, tune_y, epsilon.
    """

    def __init__(self, n_turns: int, tune_x: float, tune_y: float, epsilon: float, mu: float,
                 barrier: float = 10.0, coef=None, freq=None, start: int = 0, every: int | None = None,
                 mode: str = "exact", weights=None, device=None):
        if n_turns <= 0:
            raise ValueError("n_turns must be positive")
        if mode not in MODES:
            raise ValueError(f"mode must be one of {sorted(MODES)}")
        if coef is None or freq is None:
            coef, freq = C.sps_lines()
        self.n_turns, self.mu, self.barrier = int(n_turns), float(mu), float(barrier)
        self.params = (float(epsilon), float(mu), float(tune_x), float(tune_y))
        self.mode = MODES[mode]
        self.every = int(every) if every else default_every(n_turns)
        self.device = torch.device(device if device is not None else "cuda")
        end = start + n_turns
        wx = C.lines_modulation(tune_x, epsilon, coef, freq, start, end)
        wy = C.lines_modulation(tune_y, epsilon, coef, freq, start, end)
        dx = lines_derivatives(tune_x, epsilon, coef, freq, start, end)
        dy = lines_derivatives(tune_y, epsilon, coef, freq, start, end)
        dtab = np.stack([dx[:, 0], dx[:, 1], dy[:, 0], dy[:, 1]], axis=1)
        w = C.birkhoff_weights(n_turns) if weights is None else np.asarray(weights, dtype=np.float64)
        if w.shape != (n_turns,):
            raise ValueError("weights must have n_turns entries")
        f64 = dict(dtype=torch.float64, device=self.device)
        self.table = torch.as_tensor(C.build_table(wx, wy)).to(**f64).contiguous()
        self.dtable = torch.as_tensor(dtab).to(**f64).contiguous()
        self.weights = torch.as_tensor(w).to(**f64).contiguous()

    # ------------------------------------------------------------ raw sweeps
    def forward(self, x0: torch.Tensor):
        """x0: float64 [4][n] on the device. Returns (state [4][n], birk [n],
        alive uint8 [n], checkpoints)."""
        x0 = self._state(x0)
        n = x0.shape[1]
        n_ck = -(-self.n_turns // self.every)
        ck = torch.empty((n_ck, 4, n), dtype=torch.float64, device=self.device)
        st = torch.empty((4, n), dtype=torch.float64, device=self.device)
        birk = torch.empty(n, dtype=torch.float64, device=self.device)
        alive = torch.empty(n, dtype=torch.uint8, device=self.device)
        C.check(C.lib().hm_adjoint_forward(_p(x0[0]), _p(x0[1]), _p(x0[2]), _p(x0[3]), n, _p(self.table),
                                           self.table.shape[0], self.n_turns, self.mu, self.barrier, self.every,
                                           _p(self.weights), _p(ck), _p(st), _p(birk), _p(alive), self.mode,
                                           _stream()))
        return st, birk, alive, ck

    def backward(self, ck: torch.Tensor, alive: torch.Tensor, seed_state: torch.Tensor,
                 seed_birk: torch.Tensor | None = None, scratch_bytes: int | None = None):
        """Returns (grad_x0 [4][n], grad_params [4][n]) with grad_params rows
        (d/d epsilon, d/d mu, d/d omega_x0, d/d omega_y0) per particle.
        scratch_bytes (default: hm_adjoint_scratch_bytes' size) bounds the
        sub-checkpoint buffer; a smaller one sweeps in windows of segments."""
        n = ck.shape[2]
        seed_state = self._state(seed_state)
        if seed_state.shape[1] != n:
            raise ValueError("seed_state does not match the checkpoints")
        if seed_birk is not None:
            seed_birk = seed_birk.to(dtype=torch.float64, device=self.device).contiguous()
            if seed_birk.shape != (n,):
                raise ValueError("seed_birk must have one entry per particle")
        nb = ctypes.c_size_t(0)
        C.check(C.lib().hm_adjoint_scratch_bytes(n, self.n_turns, self.every, ctypes.byref(nb)))
        nbytes = nb.value if scratch_bytes is None else int(scratch_bytes)
        scratch = torch.empty(max(nbytes // 8, 1), dtype=torch.float64, device=self.device)
        gx = torch.empty((4, n), dtype=torch.float64, device=self.device)
        gp = torch.empty((4, n), dtype=torch.float64, device=self.device)
        C.check(C.lib().hm_adjoint_backward(_p(ck), n, _p(self.table), _p(self.dtable), self.table.shape[0],
                                            self.n_turns, self.mu, self.every, _p(self.weights), _p(seed_state),
                                            _p(seed_birk), _p(alive), _p(scratch), nbytes, _p(gx), _p(gp), self.mode,
                                            _stream()))
        return gx, gp

    def _state(self, s: torch.Tensor) -> torch.Tensor:
        s = torch.as_tensor(s).to(dtype=torch.float64, device=self.device).contiguous()
        if s.dim() != 2 or s.shape[0] != 4:
            raise ValueError("states are float64 [4][n] (x, px, y, py)")
        return s


class _TrackFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, params, n_turns, barrier, every, mode, coef, freq, start):
        eps, mu, qx, qy = (float(v) for v in params.detach().cpu().tolist())
        adj = OrbitAdjoint(n_turns, qx, qy, eps, mu, barrier, coef=coef, freq=freq, start=start, every=every,
                           mode=mode, device=x0.device)
        st, birk, alive, ck = adj.forward(x0.detach())
        ctx.adj = adj
        ctx.params_dtype = params.dtype
        ctx.save_for_backward(ck, alive)
        ctx.mark_non_differentiable(alive)
        return st, birk, alive

    @staticmethod
    def backward(ctx, g_state, g_birk, _g_alive):
        ck, alive = ctx.saved_tensors
        adj: OrbitAdjoint = ctx.adj
        if g_state is None:
            g_state = torch.zeros((4, ck.shape[2]), dtype=torch.float64, device=ck.device)
        gx, gp = adj.backward(ck, alive, g_state, g_birk)
        # chain d omega_0 / d tune = 2 pi; sum the per-particle parameter terms
        gpar = gp.sum(dim=1) * torch.tensor([1.0, 1.0, 2 * math.pi, 2 * math.pi], dtype=gp.dtype,
                                            device=gp.device)
        return gx, gpar.to(ctx.params_dtype), None, None, None, None, None, None, None


def track_orbit(x0: torch.Tensor, params: torch.Tensor, n_turns: int, barrier: float = 10.0,
                every: int | None = None, mode: str = "exact", coef=None, freq=None, start: int = 0):
    """Differentiable T-turn tracking on the device.

    x0: float64 [4][n] (x, px, y, py); params: float64 [4] = (epsilon, mu,
    tune_x, tune_y). Returns (state [4][n], birk [n], alive [n]); lost particles
    come back as NaN with alive = 0 and contribute zero gradient. Gradients flow
    to x0 and params through the device adjoint.
    """
    if params.numel() != 4:
        raise ValueError("params = (epsilon, mu, tune_x, tune_y)")
    return _TrackFn.apply(x0, params, int(n_turns), float(barrier), every, mode, coef, freq, int(start))
