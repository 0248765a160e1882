"""Batched jobs (hm_track_batch): many independent tracking jobs, each with its
own particles, table, turn count, mu and barrier, in one sequence of persistent
launches. Every job must equal a separate hm_track call of the same mode bit for
bit, and in exact mode the oracle."""
import os

import numpy as np
import pytest

import oracle as O
from henon_map_cpp_amd import _capi as C

pytestmark = pytest.mark.gpu
NT = min(os.cpu_count() or 1, 16)


def make_jobs(n_jobs, seed, n_range=(1, 3000), t_range=(50, 4000)):
    """Seeded uneven jobs: log-uniform particle and turn counts, tunes in
    [0.15, 0.32], epsilon in [0, 64], mu in [-1, 1] (some exactly 0), initial
    conditions uniform in the 4D ball r < 0.6."""
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(n_jobs):
        n = int(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]))))
        T = int(np.exp(rng.uniform(np.log(t_range[0]), np.log(t_range[1]))))
        qx, qy = rng.uniform(0.15, 0.32, 2)
        eps = rng.uniform(0, 64)
        mu = 0.0 if k % 5 == 0 else rng.uniform(-1, 1)
        d = rng.normal(size=(4, n))
        d /= np.linalg.norm(d, axis=0)
        r = 0.6 * rng.uniform(0, 1, n) ** 0.25
        x0 = d * r
        wx, wy = C.sps_modulation(qx, eps, 0, T), C.sps_modulation(qy, eps, 0, T)
        jobs.append(dict(x0=x0, T=T, mu=mu, barrier=float(rng.uniform(1.0, 10.0) ** 0.5), wx=wx, wy=wy))
    return jobs


def run_batch(jobs, mode, cfg=None):
    dev = []
    for j in jobs:
        parts = C.DeviceParticles(*j["x0"])
        tdev = C.DeviceBuffer.from_array(C.build_table(j["wx"], j["wy"]))
        dev.append((parts, tdev, j["T"], j["T"], j["mu"], j["barrier"]))
    C.track_batch(dev, mode, cfg)
    out = [d[0].get() for d in dev]
    assert C.take_status() == 0
    return out


def run_single(jobs, mode):
    out = []
    for j in jobs:
        parts = C.DeviceParticles(*j["x0"])
        tdev = C.DeviceBuffer.from_array(C.build_table(j["wx"], j["wy"]))
        C.track(parts, tdev, j["T"], j["T"], j["mu"], j["barrier"], False, mode)
        out.append(parts.get())
    return out


def same(a, b):
    return all(np.array_equal(u, v, equal_nan=True) for u, v in zip(a, b))


@pytest.mark.parametrize("mode", [C.MODE_EXACT, C.MODE_FAST])
def test_batch_equals_separate_launches(mode):
    jobs = make_jobs(48, seed=11)
    got = run_batch(jobs, mode, C.BatchCfg(0, 300, 0, 0))
    want = run_single(jobs, mode)
    bad = [k for k in range(len(jobs)) if not same(got[k], want[k])]
    assert not bad, f"jobs differing from separate launches: {bad}"
    lost = np.mean([np.mean(g[4] < j["T"]) for g, j in zip(got, jobs)])
    assert 0.05 < lost < 0.95  # the jobs straddle their apertures


def test_batch_exact_matches_oracle():
    jobs = make_jobs(12, seed=12, n_range=(50, 800), t_range=(100, 2000))
    got = run_batch(jobs, C.MODE_EXACT)
    for j, g in zip(jobs, got):
        p = O.Particles(*j["x0"])
        t = O.Tracker.from_tables(O.Tables(j["wx"], j["wy"]))
        t.track(p, j["T"], j["mu"], barrier=j["barrier"], n_threads=NT)
        assert same((p.x, p.px, p.y, p.py, p.steps), g)


def test_batch_graph_and_launch_shapes_agree():
    jobs = make_jobs(30, seed=13)
    ref = run_batch(jobs, C.MODE_EXACT)
    for cfg in (C.BatchCfg(0, 0, 0, 1), C.BatchCfg(64, 125, 97, 0), C.BatchCfg(128, 777, 2048, 1)):
        got = run_batch(jobs, C.MODE_EXACT, cfg)
        assert all(same(a, b) for a, b in zip(ref, got)), (cfg.block, cfg.compact_every, cfg.waves, cfg.use_graph)


def test_batch_empty_and_degenerate_jobs():
    jobs = make_jobs(6, seed=14)
    jobs[1]["T"] = 0
    jobs[1]["wx"], jobs[1]["wy"] = jobs[2]["wx"], jobs[2]["wy"]
    jobs[3]["x0"] = jobs[3]["x0"][:, :1]
    jobs[4]["x0"][:, :3] = np.nan  # NaN inputs are not stepped
    got = run_batch(jobs, C.MODE_EXACT)
    want = run_single([j for k, j in enumerate(jobs) if k != 1], C.MODE_EXACT)
    assert same(got[1], (*jobs[1]["x0"], np.zeros(jobs[1]["x0"].shape[1], np.uint32)))
    for g, w in zip([g for k, g in enumerate(got) if k != 1], want):
        assert same(g, w)
    assert np.all(got[4][4][:3] == 0)


@pytest.mark.parametrize("mode", [C.MODE_EXACT, C.MODE_FAST])
def test_batch_megno_jobs_equal_indicator_runs(mode):
    """Jobs that request MEGNO (about a third, mixed with tracking-only jobs)
    give the same MEGNO, coordinates and survival turns as one
    hm_track_indicators call per job (same renormalisation, one sample at the
    job's last turn)."""
    jobs = make_jobs(30, seed=15, n_range=(10, 1500), t_range=(50, 2500))
    dev, megs = [], {}
    for k, j in enumerate(jobs):
        parts = C.DeviceParticles(*j["x0"])
        tdev = C.DeviceBuffer.from_array(C.build_table(j["wx"], j["wy"]))
        if k % 3 == 1:
            megs[k] = C.DeviceBuffer(8 * parts.n)
            dev.append((parts, tdev, j["T"], j["T"], j["mu"], j["barrier"], megs[k], 100 if k % 2 else 0))
        else:
            dev.append((parts, tdev, j["T"], j["T"], j["mu"], j["barrier"]))
    C.track_batch(dev, mode, C.BatchCfg(0, 300, 0, 0))
    assert C.take_status() == 0
    for k, j in enumerate(jobs):
        got = dev[k][0].get()
        parts = C.DeviceParticles(*j["x0"])
        tdev = C.DeviceBuffer.from_array(C.build_table(j["wx"], j["wy"]))
        if k in megs:
            r = C.track_indicators(parts, tdev, j["T"], j["T"], j["mu"], j["barrier"], mode=mode,
                                   renorm_every=100 if k % 2 else 0, samples=[j["T"]], matrix=None)
            m = megs[k].download(np.float64, parts.n)
            assert np.array_equal(m, r["megno"][0], equal_nan=True), k
        else:
            C.track(parts, tdev, j["T"], j["T"], j["mu"], j["barrier"], False, mode)
        assert same(got, parts.get()), k


def test_batch_autotune_leaves_inputs_and_results_unchanged():
    """The batch autotuner times configurations on copies (the caller's
    particles stay as given); a later call without a configuration uses the
    chosen one and gives the results of the defaults bit for bit."""
    jobs = make_jobs(40, seed=16)
    dev = []
    for j in jobs:
        dev.append((C.DeviceParticles(*j["x0"]), C.DeviceBuffer.from_array(C.build_table(j["wx"], j["wy"])), j["T"],
                    j["T"], j["mu"], j["barrier"]))
    before = [d[0].get() for d in dev]
    C.check(C.lib().hm_autotune_batch_clear())
    best, ms = C.autotune_batch(dev, C.MODE_EXACT, trial_turns=500)
    assert best.block in (64, 256) and best.compact_every in (250, 1000, 4000) and best.waves > 0 and ms > 0
    assert all(same(a, d[0].get()) for a, d in zip(before, dev))
    C.track_batch(dev, C.MODE_EXACT)  # the tuned configuration
    got = [d[0].get() for d in dev]
    C.check(C.lib().hm_autotune_batch_clear())
    ref = run_batch(jobs, C.MODE_EXACT, C.BatchCfg(256, 1000, 0, 0))
    assert all(same(a, b) for a, b in zip(got, ref))


def test_batch_graph_replay_with_new_particles():
    """A recorded launch sequence is replayed for later calls with the same job
    shapes: the second call (new initial conditions, new mu and barriers, same
    particle and turn counts) must still equal separate launches."""
    jobs = make_jobs(20, seed=17)
    first = run_batch(jobs, C.MODE_EXACT, C.BatchCfg(0, 0, 0, 1))
    assert all(same(a, b) for a, b in zip(first, run_single(jobs, C.MODE_EXACT)))
    rng = np.random.default_rng(18)
    for j in jobs:
        j["x0"] = j["x0"] * rng.uniform(0.5, 1.0)
        j["mu"] = float(rng.uniform(-1, 1))
    second = run_batch(jobs, C.MODE_EXACT, C.BatchCfg(0, 0, 0, 1))
    assert all(same(a, b) for a, b in zip(second, run_single(jobs, C.MODE_EXACT)))
