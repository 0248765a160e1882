#!/usr/bin/env python3
"""Benchmark: FP64 particle-turns/s of the modulated Henon-map tracker on one MI355X.

Workload (one step): the modulated dynamic-aperture scan -- a 1000 x 1000 grid of
initial conditions over (0, 0.6]^2 (px = py = 0, 1e6 particles) tracked for 1e5
turns of the 4D map with mu = 0.5 and the seeded 7-line tune modulation at
epsilon = 32 (tunes 0.168 / 0.201), loss barrier r^2 > 10. Initial conditions and
the modulation table are resident in HBM before timing; a step restores the
initial conditions (device-to-device) and tracks all turns.

Prints ONE JSON line. `value` is nominal particle-turns/s (particles x turns per
second). The roofline entry uses the live work (sum of survival turns) x 40
algorithmic FP64 flops per particle-turn, over the k_track kernel time measured
with HIP events on the launch stream.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector (FMA = 2 flops)
HBM_PEAK_GBPS = 8000.0   # MI355X HBM3E
FLOPS_PER_PT = {0.0: 26, None: 40}  # mu == 0 / mu != 0, per live particle-turn
FLOPS_TANGENT = 103  # J entries + structured J.M, per live particle-turn
FLOPS_MEGNO = 35     # Frobenius norm ratio construct, per live particle-turn
TRACK_TRAFFIC_BYTES = 110.3e6
METRIC = "FP64 particle-turns/s (track, tangent/MEGNO, adjoint) + % of FP64 VALU peak"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", choices=["fast", "exact"], default="exact")
    ap.add_argument("--n-side", type=int, default=1000)
    ap.add_argument("--turns", type=int, default=100_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-stride", type=int, default=13)
    ap.add_argument("--single-mode", action="store_true", help="time only --mode (profiling runs)")
    ap.add_argument("--compact-every", type=int, default=0)
    ap.add_argument("--legs", default="indicators,reversibility,tunes,tune_all,adjoint,batch",
                    help="extra measurements after the headline: indicators, adjoint, batch (comma list, or none)")
    ap.add_argument("--batch-jobs", type=int, default=4096)
    ap.add_argument("--batch-max-turns", type=int, default=1_000_000,
                    help="upper end of the batch leg's log-uniform turn counts")
    ap.add_argument("--batch-plain-stride", type=int, default=16,
                    help="one call per job is timed on every k-th job of the mix (a seeded random sample)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run `--steps` device-to-device copies of MIB MiB (the known-size copy that "
                         "scales the FETCH_SIZE / WRITE_SIZE counters), then exit")
    return ap.parse_args()


def copy_calibration(mib, steps):
    """Plain device-to-device copies of a known size, for counter calibration."""
    from henon_map_cpp_amd import _capi as C

    nbytes = mib << 20
    src, dst = C.DeviceBuffer(nbytes), C.DeviceBuffer(nbytes)
    C.check(C.lib().hm_memset(src.ptr, 1, nbytes, None))
    C.check(C.lib().hm_stream_synchronize(None))
    for _ in range(steps):
        C.check(C.lib().hm_memcpy(dst.ptr, src.ptr, nbytes, 2, None))
    C.check(C.lib().hm_stream_synchronize(None))
    print(json.dumps({"copy_calibration": {"bytes_per_copy": nbytes, "copies": steps}}), flush=True)


def flops_per_pt(mu):
    return FLOPS_PER_PT[0.0] if mu == 0.0 else FLOPS_PER_PT[None]


class Workload:
    def __init__(self, n_side, turns):
        from henon_map_cpp_amd import _capi as C
        from henon_map_cpp_amd import workloads as W

        self.C = C
        c = W.MODULATED_DA
        self.mu, self.barrier = c["mu"], c["barrier"]
        self.turns = turns
        self.x0 = W.da_grid(n_side)
        self.n = self.x0[0].size
        coef, freq = W.seeded_line_coefficients(), W.line_frequencies()
        self.wx = C.lines_modulation(c["tune_x"], c["epsilon"], coef, freq, 0, turns)
        self.wy = C.lines_modulation(c["tune_y"], c["epsilon"], coef, freq, 0, turns)
        self.tab = C.build_table(self.wx, self.wy)
        self.tdev = C.DeviceBuffer.from_array(self.tab)
        self.ic = [C.DeviceBuffer.from_array(a) for a in self.x0]
        self.parts = C.DeviceParticles(*self.x0)
        self.zero_steps = C.DeviceBuffer.from_array(np.zeros(self.n, dtype=np.uint32))

    def reset(self):
        C = self.C
        for src, dst in zip(self.ic, self.parts.bufs):
            C.check(C.lib().hm_memcpy(dst.ptr, src.ptr, self.n * 8, 2, None))
        C.check(C.lib().hm_memcpy(self.parts.steps.ptr, self.zero_steps.ptr, self.n * 4, 2, None))

    def step(self, mode, cfg):
        self.reset()
        self.C.track(self.parts, self.tdev, self.turns, self.turns, self.mu, self.barrier, False, mode, cfg)

    def live_pt(self):
        return int(self.parts.steps.download(np.uint32, self.n).astype(np.int64).sum())


def time_mode(w, mode, steps, warmup, cfg):
    C = w.C
    for _ in range(warmup):
        w.step(mode, cfg)
    C.check(C.lib().hm_stream_synchronize(None))
    C.profile_reset()
    C.profile_enable(True)
    _sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        w.step(mode, cfg)
    _sync()
    C.check(C.lib().hm_stream_synchronize(None))
    dt = time.perf_counter() - t0
    C.profile_enable(False)
    kms, launches = C.profile_read("k_track")
    live = w.live_pt()
    return dt, kms, launches, live


def _sync():
    try:
        import torch

        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def _lines(tune, eps, turns):
    from henon_map_cpp_amd import _capi as C
    from henon_map_cpp_amd import workloads as W

    return C.lines_modulation(tune, eps, W.seeded_line_coefficients(), W.line_frequencies(), 0, turns)


def indicators_leg(steps, warmup, mode_name="exact"):
    """Fused tangent map + MEGNO + FLI with renormalisation every 100 turns: the
    chaos-indicator workload (256x256 grid + 65536 ball points, 1e5 turns)."""
    from henon_map_cpp_amd import _capi as C
    from henon_map_cpp_amd import workloads as W

    c = W.CHAOS_INDICATORS
    x0 = W.indicator_points()
    T = c["n_turns"]
    tdev = C.DeviceBuffer.from_array(C.build_table(_lines(c["tune_x"], c["epsilon"], T),
                                                   _lines(c["tune_y"], c["epsilon"], T)))
    n = x0[0].size
    kms = 0.0
    live = 0
    for it in range(warmup + steps):
        parts = C.DeviceParticles(*x0)
        if it == warmup:
            C.profile_reset()
            C.profile_enable(True)
        C.track_indicators(parts, tdev, T, T, c["mu"], c["barrier"],
                           mode={"fast": C.MODE_FAST, "exact_orbit": C.MODE_EXACT_ORBIT}.get(mode_name, C.MODE_EXACT),
                           renorm_every=c["renorm_every"], samples=c["samples"], matrix=True)
        if it >= warmup:
            live += int(parts.get()[4].astype(np.int64).sum())
    kms, launches = C.profile_read("k_indicators")
    C.profile_enable(False)
    fpt = FLOPS_PER_PT[None] + FLOPS_TANGENT + FLOPS_MEGNO
    ks = kms / 1e3
    return {
        "workload": c["name"],
        "mode": mode_name,
        "kernel": "k_indicators",
        "particles": n,
        "turns": T,
        "value": n * T * steps / ks,
        "live_value": live / ks,
        "unit": "particle-turns/s (kernel time)",
        "flops_per_live_particle_turn": fpt,
        "achieved": live * fpt / ks / 1e12,
        "frac": live * fpt / ks / 1e12 / FP64_PEAK_TFLOPS,
        "kernel_ms_per_step": kms / steps,
    }


def reversibility_leg():
    """Reversibility error |x0 - B^t F^t x0| (henon_map_cpp/dynamic_indicators.py:
    136-151) of the chaos-indicator particle set at t = 1e3, 1e4, 1e5: one
    forward pass with snapshots, then the inverse map from each snapshot."""
    from henon_map_cpp_amd import _capi as C
    from henon_map_cpp_amd import workloads as W

    c = W.CHAOS_INDICATORS
    x0 = W.indicator_points()
    T = c["n_turns"]
    tdev = C.DeviceBuffer.from_array(C.build_table(_lines(c["tune_x"], c["epsilon"], T),
                                                   _lines(c["tune_y"], c["epsilon"], T)))
    n = x0[0].size
    parts = C.DeviceParticles(*x0)
    C.check(C.lib().hm_stream_synchronize(None))
    C.profile_reset()
    C.profile_enable(True)
    out, done = {}, 0
    for t in (1_000, 10_000, 100_000):
        C.track(parts, tdev, T, t - done, c["mu"], c["barrier"], False, C.MODE_EXACT)
        done = t
        g = parts.get()
        snap = C.DeviceParticles(*g[:4], steps=g[4])
        C.track(snap, tdev, T, t, c["mu"], c["barrier"], True, C.MODE_EXACT)
        back = snap.get()
        err = np.sqrt(sum((b - a) ** 2 for a, b in zip(x0, back[:4])))
        ok = np.isfinite(err)
        out[str(t)] = {"median": float(np.median(err[ok])) if ok.any() else None,
                       "p99": float(np.quantile(err[ok], 0.99)) if ok.any() else None,
                       "survivors": int(ok.sum())}
    ms, launches = C.profile_read("k_track")
    C.profile_enable(False)
    pt = n * (T + 1_000 + 10_000 + 100_000)
    return {"workload": f"{n} particles (chaos-indicator set), forward 1e5 turns + inverse 1e3, 1e4, 1e5 turns",
            "kernel_ms": ms, "particle_turns_per_s": pt / (ms / 1e3), "error": out}


def tunes_leg():
    """Trajectory store + 2048-sample FFT tunes and Birkhoff tunes of the
    chaos-indicator particle set (256x256 grid + 65536 ball points), as
    all_tunes runs them (both tunes from one read of the rows: hm_tunes_rows),
    next to the two stages run separately. The trajectory is stored as
    particle rows (hm_track_rows), which also sums the per-coordinate means that
    both tune stages remove; the FFT reads the rows in place. Bytes: 32 B per
    particle-sample written by the store; 32 B per particle-sample read by each
    tune stage for the particles whose four means are finite (a NaN mean, a lost
    particle, makes both tunes NaN without a read); every kernel of a stage is
    timed."""
    from henon_map_cpp_amd import _capi as C
    from henon_map_cpp_amd import workloads as W

    c = W.CHAOS_INDICATORS
    x0 = W.indicator_points()
    T = 2048
    tdev = C.DeviceBuffer.from_array(C.build_table(_lines(c["tune_x"], c["epsilon"], T),
                                                   _lines(c["tune_y"], c["epsilon"], T)))
    n = x0[0].size
    L = C.row_length(T - 1)
    best = None
    for it in range(2):
        parts = C.DeviceParticles(*x0)
        C.check(C.lib().hm_stream_synchronize(None))
        C.profile_reset()
        C.profile_enable(True)
        rows, means = C.track_rows(parts, tdev, T, T - 1, c["mu"], c["barrier"], means=True)  # samples 0..2047
        fft = C.history_tunes(rows, n, T, [(0, T - 1)], "fft", row_len=L, means=means)
        birk = C.history_tunes(rows, n, T, [(0, T - 1)], "birkhoff", row_len=L, means=means)
        birk2, fft2 = C.tunes_rows(rows, n, T, [(0, T - 1)], L, means=means)
        h_ms = C.profile_read("k_history_rows")[0]
        f_ms = sum(C.profile_read(k)[0] for k in ("k_signal_rows", "k_fft_tune_rows"))
        b_ms = C.profile_read("k_birkhoff_hist")[0]
        u_ms = C.profile_read("k_tunes_fused")[0]
        C.profile_enable(False)
        assert np.array_equal(fft2, fft, equal_nan=True)
        assert np.allclose(birk2, birk, rtol=0, atol=1e-14, equal_nan=True)
        n_read = int(np.all(np.isfinite(means.download(np.float64, 4 * n).reshape(4, n)), axis=0).sum())
        rows.free()
        means.free()
        if best is None or h_ms + f_ms + b_ms + u_ms < sum(best):
            best = (h_ms, f_ms, b_ms, u_ms)
    h_ms, f_ms, b_ms, u_ms = best
    nbytes, rbytes = 32.0 * n * T, 32.0 * n_read * T
    fft_flops = 2 * 5 * T * np.log2(T) * n_read
    return {
        "workload": f"{n} particles x 2048 samples (chaos-indicator set), mu=0.5, eps=32",
        "history_ms": h_ms, "history_GBps": nbytes / (h_ms / 1e3) / 1e9,
        "particles_read": n_read,
        "fft_ms": f_ms, "fft_GBps": rbytes / (f_ms / 1e3) / 1e9, "fft_TFLOPs": fft_flops / (f_ms / 1e3) / 1e12,
        "birkhoff_ms": b_ms, "birkhoff_GBps": rbytes / (b_ms / 1e3) / 1e9,
        "fused_ms": u_ms, "fused_GBps": rbytes / (u_ms / 1e3) / 1e9,
        "fused_vs_separate": (f_ms + b_ms) / u_ms,
        "hbm_peak_GBps": HBM_PEAK_GBPS,
        "finite_fft_tunes": float(np.mean(np.isfinite(fft))), "finite_birkhoff_tunes": float(np.mean(np.isfinite(birk))),
    }


def tune_all_leg(n=90_000, n_turns=10_000, reps=2):
    """The reference's own tune workload (tests/new_tune.ipynb cells 1-3): n
    copies of (0.1, 0, 0.1, 0), henon_tracker(1e4, 0.168, 0.201, "sps") (eps=0),
    tune_all(p, 1e4, mu=0, from_idx=[0, 500], to_idx=[500, 1000]), through the
    drop-in engine call (trajectory rows, means, Birkhoff and FFT tunes of four
    windows: 501, 501, 10001 and 10000 samples, results back on the host). The
    notebook reports 38.33 s on the reference's CPU path. FFT bytes: the two
    coordinates of a plane read once per window (16 B per sample per plane)."""
    import henon_map_cpp_amd as hm
    from henon_map_cpp_amd import _capi as C

    pp = np.full(n, 0.1)
    z = np.zeros(n)
    t = hm.henon_tracker(n_turns, 0.168, 0.201, "sps")
    best = None
    for _ in range(reps):
        p = hm.particles(pp, z, pp, z)
        C.check(C.lib().hm_stream_synchronize(None))
        C.profile_reset()
        C.profile_enable(True)
        t0 = time.perf_counter()
        rb, rf = t.tracker.all_tunes(p.particles, n_turns, 0.0, 10.0, float("nan"), False, [0, 500], [500, 1000])
        dt = time.perf_counter() - t0
        ks = {k: C.profile_read(k)[0] for k in ("k_history_rows", "k_birkhoff_hist", "k_fft_tune_rows",
                                                "k_cz_columns", "k_cz_rows", "k_cz_final", "k_cz_tune",
                                                "k_mr_columns", "k_mr_rows", "k_mr_final", "k_mr_tune",
                                                "k_wh_tune")}
        C.profile_enable(False)
        if best is None or dt < best[0]:
            best = (dt, ks, np.asarray(rb), np.asarray(rf))
    dt, ks, rb, rf = best
    fft_ms = sum(v for k, v in ks.items() if k.startswith(("k_cz", "k_mr", "k_wh")) or k == "k_fft_tune_rows")
    fft_bytes = 16.0 * n * 2 * (501 + 501 + 10001 + 10000)
    return {
        "workload": f"tune_all of tests/new_tune.ipynb: {n} particles at (0.1, 0, 0.1, 0), {n_turns} turns, "
                    "sps eps=0, mu=0, windows [0,500], [500,1000] + (0, 1e4) + full signal",
        "seconds": dt,
        "reference_cpu_seconds": 38.33,
        "speedup_vs_reference_cpu": 38.33 / dt,
        "particle_turns_per_s": n * n_turns / dt,
        "kernel_ms": ks,
        "fft_stage_ms": fft_ms,
        "fft_algorithmic_GBps": fft_bytes / (fft_ms / 1e3) / 1e9,
        "history_GBps": 32.0 * n * (n_turns + 1) / (ks["k_history_rows"] / 1e3) / 1e9,
        "fft_tune_x_full": float(rf[0, -2]), "birkhoff_tune_x_full": float(rb[0, -2]),
    }


def adjoint_leg():
    """Whole-orbit gradients: 65536 points (r ~ U[0.05, 0.4]), 1e4 turns, eps=16,
    mu=0.5, losses L1 = mean final r^2 and L2 = mean Birkhoff-averaged x^2 over the
    survivors; cost of forward + backward sweeps against forward-only tracking."""
    import torch

    from henon_map_cpp_amd import _capi as C
    from henon_map_cpp_amd import adjoint as A
    from henon_map_cpp_amd import workloads as W

    c = W.ADJOINT_GRADIENTS
    T = c["n_turns"]
    x0 = np.stack(W.shell_uniform_radius(c["n"], c["r_min"], c["r_max"], c["seed"]))
    coef, freq = W.seeded_line_coefficients(), W.line_frequencies()
    wx, wy = _lines(c["tune_x"], c["epsilon"], T), _lines(c["tune_y"], c["epsilon"], T)
    tdev = C.DeviceBuffer.from_array(C.build_table(wx, wy))
    x0t = torch.as_tensor(x0, device="cuda")

    def track_once():
        parts = C.DeviceParticles(*x0)
        C.check(C.lib().hm_stream_synchronize(None))
        C.profile_reset()
        C.profile_enable(True)
        C.track(parts, tdev, T, T, c["mu"], c["barrier"], False, C.MODE_EXACT)
        ms = C.profile_read("k_track")[0]
        C.profile_enable(False)
        return ms

    def adjoint_once(adj):
        torch.cuda.synchronize()
        C.profile_reset()
        C.profile_enable(True)
        st, birk, alive, ck = adj.forward(x0t)
        ok = alive.bool()
        ns = max(int(ok.sum()), 1)
        seed_s = torch.where(ok[None, :], 2.0 * st / ns, torch.zeros_like(st))   # dL1/ds_T
        seed_b = torch.where(ok, torch.full_like(birk, 1.0 / ns), torch.zeros_like(birk))  # dL2/dB
        adj.backward(ck, alive, seed_s, seed_b)
        torch.cuda.synchronize()
        f_ms = C.profile_read("k_adjoint_forward")[0]
        b_ms = C.profile_read("k_adjoint_backward")[0]
        C.profile_enable(False)
        return f_ms, b_ms, int(ok.sum())

    # The GPU clock drifts with load and temperature by up to ~20 % over a run,
    # so every adjoint run is timed right after a forward-only tracking run on
    # the same inputs and the cost is the median of three such pairs' ratios.
    # The tracking run is the second of two back to back: the first after other
    # work finds the table out of L2, which slows tracking at one wavefront per
    # SIMD by up to 1.8x.
    out = {"workload": c["name"], "every": {}}
    all_track = []
    for K in c["every"]:
        adj = A.OrbitAdjoint(T, c["tune_x"], c["tune_y"], c["epsilon"], c["mu"], c["barrier"], coef=coef,
                             freq=freq, every=K)
        adjoint_once(adj)  # warm-up
        runs = []
        for it in range(3):
            track_once()
            t_ms = track_once()
            f_ms, b_ms, survivors = adjoint_once(adj)
            runs.append(((f_ms + b_ms) / t_ms, t_ms, f_ms, b_ms))
            all_track.append(t_ms)
        ratio, t_ms, f_ms, b_ms = sorted(runs)[1]
        out["every"][str(K)] = {"forward_ms": f_ms, "backward_ms": b_ms, "track_ms": t_ms,
                                "cost_vs_tracking": ratio,
                                "particle_turns_per_s": x0.shape[1] * T / ((f_ms + b_ms) / 1e3)}
    out["forward_only_track_ms"] = min(all_track)
    out["survivors"] = survivors
    return out


def batch_leg(n_jobs, max_turns=1_000_000, plain_stride=16):
    """Uneven independent jobs (the batched-jobs mix: 4096 jobs, N ~ logU[64,
    65536], T ~ logU[1e3, 1e6], 30 % with MEGNO) in one batched launch sequence,
    against one plain call per job (hm_track, or hm_track_indicators for the
    MEGNO jobs; same GPU, same exact arithmetic), wall time each. The whole mix
    runs batched; the one-call-per-job rate is measured on every
    `plain_stride`-th job (a seeded random sample of the mix: about 190 s for
    all 4096 jobs), which is also run batched on its own to compare results
    and times job for job."""
    from henon_map_cpp_amd import _capi as C
    from henon_map_cpp_amd import workloads as W

    c = dict(W.BATCHED_JOBS, t_range=(W.BATCHED_JOBS["t_range"][0], max_turns))
    c["name"] = c["name"].replace("logU[1e3, 1e6]", f"logU[1e3, {max_turns:.0e}]")
    params = W.batch_job_params(n_jobs, c)
    coef, freq = W.seeded_line_coefficients(), W.line_frequencies()
    t0 = time.perf_counter()
    ics, tabs = [], []
    for k, (n, T, qx, qy, eps, mu, seed) in enumerate(params):
        ics.append(np.stack(W.ball(n, c["r_max"], seed=seed)))
        wx = C.lines_modulation(qx, eps, coef, freq, 0, T)
        wy = C.lines_modulation(qy, eps, coef, freq, 0, T)
        tabs.append(C.DeviceBuffer.from_array(C.build_table(wx, wy)))
        if k and k % 1024 == 0:
            print(f"[bench] batch: tables of {k} jobs, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    megno_job = [W.batch_megno_job(k) for k in range(n_jobs)]
    renorm = c["megno_renorm_every"]

    def batched(sel, graph=False):
        parts = [C.DeviceParticles(*ics[k]) for k in sel]
        megs = [C.DeviceBuffer(8 * params[k][0]) if megno_job[k] else None for k in sel]
        jobs = [(pp, tabs[k], params[k][1], params[k][1], params[k][5], c["barrier"], mg, renorm)
                for pp, k, mg in zip(parts, sel, megs)]
        _sync()
        t0 = time.perf_counter()
        cfg = C.BatchCfg(tuned.block, tuned.compact_every, tuned.waves, 1) if graph else None
        C.track_batch(jobs, C.MODE_EXACT, cfg)
        C.check(C.lib().hm_stream_synchronize(None))
        return time.perf_counter() - t0, parts, megs

    def plain(sel):
        parts = [C.DeviceParticles(*ics[k]) for k in sel]
        megs = []
        _sync()
        t0 = time.perf_counter()
        for q, (pp, k) in enumerate(zip(parts, sel)):
            n, T, mu = params[k][0], params[k][1], params[k][5]
            if q and q % 64 == 0:
                C.check(C.lib().hm_stream_synchronize(None))
                print(f"[bench] batch: {q} of {len(sel)} plain calls, {time.perf_counter() - t0:.1f} s",
                      file=sys.stderr, flush=True)
            if megno_job[k]:
                r = C.track_indicators(pp, tabs[k], T, T, mu, c["barrier"], mode=C.MODE_EXACT, renorm_every=renorm,
                                       samples=[T])
                megs.append(r["megno"][0])
            else:
                C.track(pp, tabs[k], T, T, mu, c["barrier"], False, C.MODE_EXACT)
                megs.append(None)
        C.check(C.lib().hm_stream_synchronize(None))
        return time.perf_counter() - t0, parts, megs

    def note(what, t):
        print(f"[bench] batch: {what} {t:.2f} s", file=sys.stderr, flush=True)

    every = list(range(n_jobs))
    sub = every[::plain_stride]
    pt = sum(params[k][0] * params[k][1] for k in every)
    pt_sub = sum(params[k][0] * params[k][1] for k in sub)
    # the autotuner picks block, wavefronts and compaction interval for this
    # job mix (trials on copies, every job cut to 4000 turns); the timed
    # batched calls below pass no configuration and so use its choice
    C.check(C.lib().hm_autotune_batch_clear())
    parts = [C.DeviceParticles(*ics[k]) for k in every]
    megs = [C.DeviceBuffer(8 * params[k][0]) if megno_job[k] else None for k in every]
    t0 = time.perf_counter()
    tuned, trial_ms = C.autotune_batch([(pp, tabs[k], params[k][1], params[k][1], params[k][5], c["barrier"], mg, renorm)
                                        for pp, k, mg in zip(parts, every, megs)], C.MODE_EXACT, trial_turns=4000)
    tune_s = time.perf_counter() - t0
    del parts, megs
    note(f"autotune (block {tuned.block}, compact_every {tuned.compact_every}, waves {tuned.waves})", tune_s)
    runs = []
    for _ in range(2):
        runs.append(batched(every)[0])
        note("batched, whole mix", runs[-1])
    tb = min(runs)
    batched(every, graph=True)  # records the launch sequence
    tg = batched(every, graph=True)[0]
    note("batched, graph replay, whole mix", tg)
    tbs, pb, mb = batched(sub)
    note(f"batched, every {plain_stride}th job", tbs)
    tp, pp, mp = plain(sub)
    note(f"one call per job, every {plain_stride}th job", tp)
    identical = all(np.array_equal(u, v, equal_nan=True) for a, b in zip(pb, pp) for u, v in zip(a.get(), b.get()))
    identical = identical and all(
        np.array_equal(a.download(np.float64, len(b)), b, equal_nan=True) for a, b in zip(mb, mp) if a is not None)
    return {"workload": c["name"] + f", 30% of the jobs with MEGNO (renormalised every {renorm} turns)",
            "jobs": n_jobs, "megno_jobs": int(sum(megno_job)), "nominal_particle_turns": pt,
            "autotuned": {"block": tuned.block, "compact_every": tuned.compact_every, "waves": tuned.waves,
                          "trial_ms": trial_ms, "autotune_s": tune_s},
            "batched_s": tb, "batched_graph_replay_s": tg, "batched_value": pt / tb,
            "plain_sample": f"every {plain_stride}th job: {len(sub)} jobs, {pt_sub:.4g} nominal particle-turns",
            "plain_s": tp, "plain_value": pt_sub / tp,
            # the sample is a uniform random draw of the jobs, so n_jobs / len(sub)
            # times its one-call-per-job time estimates the whole mix's
            "plain_whole_mix_estimate_s": tp * n_jobs / len(sub),
            "speedup": tp * n_jobs / len(sub) / tb,
            "sample_batched_s": tbs, "sample_speedup": tp / tbs,
            "identical_to_plain": identical}


def host_cpu_info():
    """What the host offers: the reference's thread count
    (std::thread::hardware_concurrency() = os.cpu_count()), the CPUs this
    process may run on, the cgroup CPU quota, the host's CPU share given in
    OMP_NUM_THREADS, and the CPU model."""
    import subprocess

    info = {"hardware_concurrency": os.cpu_count()}
    try:
        info["nproc"] = int(subprocess.run(["nproc"], capture_output=True, text=True).stdout.strip())
    except Exception:
        info["nproc"] = None
    info["affinity"] = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = int(q) / int(period)
    except Exception:
        pass
    info["cgroup_cpu_quota"] = quota
    omp = os.environ.get("OMP_NUM_THREADS", "")
    info["cpu_share"] = int(omp) if omp.isdigit() and int(omp) > 0 else None
    model = None
    try:
        for line in subprocess.run(["lscpu"], capture_output=True, text=True).stdout.splitlines():
            if line.lower().startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except Exception:
        pass
    info["model"] = model
    return info


def cpu_baseline(stride, turns):
    """The oracle (C++ restatement of the reference CPU path, compiled like the
    reference's CPU build: -O3, no FMA contraction, no -march) on every
    `stride`-th particle of the same grid, with the reference's threading model
    (one std::thread per hardware thread, particles strided over threads,
    henon.cu:2701-2752). The thread count is hardware_concurrency() limited to
    the CPUs this process can use (affinity, cgroup quota, the CPU share given
    in OMP_NUM_THREADS):
    more threads than that only time-slice. A one-thread run on a smaller sample
    gives the per-core rate; `projected_full_host_value` scales it to every
    hardware thread (linear scaling, a projection, not a measurement)."""
    import oracle as O
    from henon_map_cpp_amd import workloads as W

    O.build()
    c = W.MODULATED_DA
    info = host_cpu_info()
    limits = [v for v in (info["hardware_concurrency"], info["affinity"], info["cgroup_cpu_quota"],
                          info["cpu_share"]) if v]
    threads = max(1, int(min(limits)))
    coef, freq = W.seeded_line_coefficients(), W.line_frequencies()
    t = O.Tracker(turns, c["tune_x"], c["tune_y"], "custom", epsilon=c["epsilon"], coef=coef, freq=freq)

    def run(step, n_threads):
        x, px, y, py = (a[::step].copy() for a in W.da_grid(1000))
        p = O.Particles(x, px, y, py)
        t0 = time.perf_counter()
        t.track(p, turns, c["mu"], barrier=c["barrier"], n_threads=n_threads)
        dt = time.perf_counter() - t0
        return p.n, int(p.steps.astype(np.int64).sum()), dt

    n, live, dt = run(stride, threads)
    n1, live1, dt1 = run(stride * threads, 1)
    per_core = n1 * turns / dt1
    return {
        "value": n * turns / dt,
        "unit": "particle-turns/s",
        "cores": threads,
        "kind": "port",
        "sample": f"every {stride}th particle of the 1000x1000 grid ({n} particles) x {turns} turns, mu=0.5, "
                  f"eps=32, {threads} threads: {dt:.1f} s, live particle-turns/s {live / dt:.4g}; one thread on "
                  f"every {stride * threads}th particle ({n1} particles): {dt1:.1f} s",
        "live_value": live / dt,
        "seconds": dt,
        "one_thread_value": per_core,
        "thread_scaling": (n * turns / dt) / per_core,
        "projected_full_host_value": per_core * info["hardware_concurrency"],
        "host": info,
    }


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported "
              f"(every GPU would track its own share of the particles; only one is driven here)", file=sys.stderr)
        sys.exit(2)
    from henon_map_cpp_amd import _capi as C

    if C.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if a.copy_calibration:
        copy_calibration(a.copy_calibration, a.steps)
        return
    w = Workload(a.n_side, a.turns)
    cfg = C.LaunchCfg(0, a.compact_every, 0, 0) if a.compact_every else None
    fpt = flops_per_pt(w.mu)
    results = {}
    modes = [a.mode] + ([] if a.single_mode else [m for m in ("fast", "exact") if m != a.mode])
    for mode_name in modes:
        mode = C.MODE_FAST if mode_name == "fast" else C.MODE_EXACT
        dt, kms, launches, live = time_mode(w, mode, a.steps, a.warmup, cfg)
        per_step_live = live
        kernel_s_per_step = kms / 1e3 / a.steps
        achieved = per_step_live * fpt / kernel_s_per_step / 1e12
        results[mode_name] = dict(
            ms_per_step=dt / a.steps * 1e3,
            value=w.n * w.turns * a.steps / dt,
            live_value=per_step_live * a.steps / dt,
            kernel_ms_per_step=kernel_s_per_step * 1e3,
            launches_per_step=launches / a.steps,
            achieved_tflops=achieved,
            frac=achieved / FP64_PEAK_TFLOPS,
            live_pt=per_step_live,
        )
        print(f"[bench] mode={mode_name} {results[mode_name]}", file=sys.stderr, flush=True)
    head = results[a.mode]
    line = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "particle-turns/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"modulated DA scan {a.n_side}x{a.n_side} grid over (0,0.6]^2, {a.turns} turns, mu=0.5, "
                        f"7-line tune modulation eps=32 (seeded a_k), tunes 0.168/0.201, barrier r^2>10",
            "mode": a.mode,
            "particles": w.n,
            "turns": w.turns,
            "live_particle_turns_per_step": head["live_pt"],
            "live_value": head["live_value"],
            "flops_per_live_particle_turn": fpt,
            "other_mode": {k: {kk: v[kk] for kk in ("value", "live_value", "ms_per_step", "frac")}
                           for k, v in results.items() if k != a.mode},
        },
        "roofline": {
            "bound": "valu",
            "achieved": head["achieved_tflops"],
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": head["frac"],
            # HBM bytes per k_track launch: FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0
            # (factors from a plain copy of known size), 2e4-turn run of this scan;
            # profiles/k_track_hbm_traffic_pmc.txt
            "traffic": TRACK_TRAFFIC_BYTES,
            "traffic_unit": "bytes per launch",
            "kernel": "k_track",
            "kernel_ms_per_step": head["kernel_ms_per_step"],
            "launches_per_step": head["launches_per_step"],
        },
        "cpu_baseline": None,
    }
    legs = [x for x in a.legs.split(",") if x and x != "none"]
    if "indicators" in legs:
        line["tangent_megno"] = indicators_leg(2, 1)
        other = {m: indicators_leg(1, 0, m) for m in ("exact_orbit", "fast")}
        line["tangent_megno"]["other_mode"] = {m: {k: r[k] for k in ("value", "live_value", "frac",
                                                                     "kernel_ms_per_step")}
                                               for m, r in other.items()}
        print(f"[bench] tangent/MEGNO {line['tangent_megno']}", file=sys.stderr, flush=True)
    if "reversibility" in legs:
        line["reversibility"] = reversibility_leg()
        print(f"[bench] reversibility {line['reversibility']}", file=sys.stderr, flush=True)
    if "tunes" in legs:
        line["tunes"] = tunes_leg()
        print(f"[bench] tunes {line['tunes']}", file=sys.stderr, flush=True)
    if "tune_all" in legs:
        line["tune_all"] = tune_all_leg()
        print(f"[bench] tune_all {line['tune_all']}", file=sys.stderr, flush=True)
    if "adjoint" in legs:
        line["adjoint"] = adjoint_leg()
        print(f"[bench] adjoint {line['adjoint']}", file=sys.stderr, flush=True)
    if "batch" in legs:
        line["batch"] = batch_leg(a.batch_jobs, a.batch_max_turns, a.batch_plain_stride)
        print(f"[bench] batch {line['batch']}", file=sys.stderr, flush=True)
    if not a.no_cpu_baseline:
        try:
            line["cpu_baseline"] = cpu_baseline(a.cpu_sample_stride, a.turns)
        except Exception as e:  # reported, never silently replaced
            line["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
