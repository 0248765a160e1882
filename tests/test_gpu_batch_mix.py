"""The batched-jobs mix at its stated size (4096 jobs, N ~ logU[64, 65536],
T ~ logU[1e3, 1e6], 30 % with MEGNO): a seeded sample of its jobs that always
holds the largest particle count, the longest job, the largest job and the
largest and smallest MEGNO jobs, tracked in one batched call. Every job must equal its own
plain call (hm_track, or hm_track_indicators for MEGNO jobs) bit for bit, and
the small jobs the oracle."""
import os

import numpy as np
import pytest

import oracle as O
from henon_map_cpp_amd import _capi as C
from henon_map_cpp_amd import workloads as W

pytestmark = pytest.mark.gpu
NT = min(os.cpu_count() or 1, 16)


def sample_of_mix(n_sample=24, seed=77):
    c = W.BATCHED_JOBS
    params = W.batch_job_params(c["n_jobs"], c)
    n = np.array([p[0] for p in params])
    t = np.array([p[1] for p in params])
    megno = np.array([W.batch_megno_job(k) for k in range(len(params))])
    mk = np.flatnonzero(megno)
    forced = {int(np.argmax(n)), int(np.argmax(t)), int(np.argmax(n * t)), int(mk[np.argmax((n * t)[mk])]),
              int(mk[np.argmin((n * t)[mk])])}
    rng = np.random.default_rng(seed)
    rest = [k for k in rng.permutation(len(params)) if k not in forced][:n_sample - len(forced)]
    return c, params, megno, sorted(forced) + sorted(int(k) for k in rest)


def test_mix_sample_batched_equals_plain_and_oracle():
    c, params, megno, sel = sample_of_mix()
    assert max(params[k][0] for k in sel) > 60_000 and max(params[k][1] for k in sel) > 900_000
    coef, freq = W.seeded_line_coefficients(), W.line_frequencies()
    setup = {}
    for k in sel:
        n, T, qx, qy, eps, mu, seed = params[k]
        wx = C.lines_modulation(qx, eps, coef, freq, 0, T)
        wy = C.lines_modulation(qy, eps, coef, freq, 0, T)
        setup[k] = (np.stack(W.ball(n, c["r_max"], seed=seed)), wx, wy, C.DeviceBuffer.from_array(C.build_table(wx, wy)))
    renorm = c["megno_renorm_every"]
    parts, megs, jobs = {}, {}, []
    for k in sel:
        n, T, mu = params[k][0], params[k][1], params[k][5]
        parts[k] = C.DeviceParticles(*setup[k][0])
        megs[k] = C.DeviceBuffer(8 * n) if megno[k] else None
        jobs.append((parts[k], setup[k][3], T, T, mu, c["barrier"], megs[k], renorm))
    print(f"batched call: {len(sel)} jobs, {sum(params[k][0] * params[k][1] for k in sel):.3g} particle-turns",
          flush=True)
    C.track_batch(jobs, C.MODE_EXACT)
    assert C.take_status() == 0
    n_oracle = 0
    for k in sel:
        n, T, mu = params[k][0], params[k][1], params[k][5]
        got = parts[k].get()
        ref = C.DeviceParticles(*setup[k][0])
        if megno[k]:
            r = C.track_indicators(ref, setup[k][3], T, T, mu, c["barrier"], mode=C.MODE_EXACT, renorm_every=renorm,
                                   samples=[T])
            assert np.array_equal(megs[k].download(np.float64, n), r["megno"][0], equal_nan=True), k
        else:
            C.track(ref, setup[k][3], T, T, mu, c["barrier"], False, C.MODE_EXACT)
        want = ref.get()
        assert all(np.array_equal(a, b, equal_nan=True) for a, b in zip(got, want)), k
        if n * T <= 3e6:  # the oracle on the small jobs
            n_oracle += 1
            p = O.Particles(*setup[k][0])
            tables = O.Tables(setup[k][1], setup[k][2])
            if megno[k]:
                r = O.indicators(p, tables, T, mu, c["barrier"], renorm, [T], n_threads=NT)
                np.testing.assert_allclose(megs[k].download(np.float64, n), r["megno"][0], rtol=1e-11, atol=1e-13,
                                           equal_nan=True)
            else:
                O.Tracker.from_tables(tables).track(p, T, mu, barrier=c["barrier"], n_threads=NT)
            assert all(np.array_equal(a, b, equal_nan=True) for a, b in zip((p.x, p.px, p.y, p.py, p.steps), got)), k
    assert n_oracle >= 2
