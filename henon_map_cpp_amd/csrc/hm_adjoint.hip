// Reverse-mode (adjoint) differentiation of n-turn tracking, with checkpoints.
// No reference counterpart: the reference has no derivatives beyond the
// one-turn tangent matrix (src/henon.cu:175-265).
//
// Forward sweep (k_adjoint_forward): track from the initial points with table
// index k, write the state every K turns (checkpoints), the final state, the
// Birkhoff-weighted sum B = sum_{k=1..T} w_{k-1} x_k^2 (w = birkhoff_weights(T))
// and whether the particle survived all T turns.
//
// Backward sweep (k_adjoint_backward): per particle, for each K-turn segment in
// reverse, recompute the segment from its checkpoint writing sub-checkpoints
// every kChunk turns to scratch, then per chunk (in reverse) recompute its
// kChunk+1 states into registers and propagate the adjoint lambda <- J^T lambda
// (+ the Birkhoff seed), accumulating the parameter derivatives
//   d/dmu     : lambda_p~ . (x^3 - 3 x y^2, y^3 - 3 x^2 y)   (octupole kick)
//   d/domega_k: lambda_x' p' - lambda_p' x'                   (rotation, per plane)
// chained through d omega_k / d(omega_0, epsilon) (a per-turn table).
// Recomputation never inverts the map, which is unstable on chaotic orbits.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

#include "hm_device.h"
#include "hm_internal.h"

namespace hm {

constexpr int kChunk = 8;

struct AdjointFwdArgs {
    const double *x0, *px0, *y0, *py0;  // initial points [n]
    uint32_t n;
    const TabEntry *tab;
    uint32_t n_turns, every;  // T, K
    double mu, b2;
    const double *weights;    // [T] birkhoff_weights(T)
    double *ckpt;             // [T/K + 1][4][n]
    double *final_state;      // [4][n]
    double *birk;             // [n]
    uint8_t *alive;           // [n]
};

// Every lane runs every turn (a lost particle's lane carries on with values
// that are discarded), so the table record and the weight of a turn are
// wave-uniform scalar loads, requested ahead of use. Per K-turn segment: the
// checkpoint, then the turns four at a time (the tracking loop's shape: two
// alternating pairs of prefetched records, the loss test on a rare path), with
// the Birkhoff sum taking the carried x*x in exact mode.
template <int MODE, bool OCT>
__global__ void __launch_bounds__(256) k_adjoint_forward(AdjointFwdArgs a)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = i < a.n;
    const size_t n = a.n;
    double x = __builtin_nan(""), px = x, y = x, py = x;
    if (in) {
        x = a.x0[i]; px = a.px0[i]; y = a.y0[i]; py = a.py0[i];
    }
    bool alive = in && !any_nan(x, px, y, py);
    if (!__any(alive)) {
        if (in) {
            a.final_state[i] = x; a.final_state[n + i] = px; a.final_state[2 * n + i] = y;
            a.final_state[3 * n + i] = py;
            a.birk[i] = __builtin_nan("");
            a.alive[i] = 0;
        }
        return;
    }
    if (!alive) x = px = y = py = 0.0;  // NaN inputs and lanes past n: parked (see below)
    const ConstPtr<TabEntry> tab = as_const(a.tab);
    const ConstPtr<double> weights = as_const(a.weights);
    const double mu = a.mu, m3mu = -3.0 * mu, b2 = a.b2, b2s = screen_bound(b2);
    double xx = x * x, yy = y * y;
    double bsum = 0.0;
    // a lost particle is parked at the origin, a fixed point of the map, so
    // that its lane never again takes the loss test's rare path
    auto turn = [&](const TabEntry &t, double w) {
        if (forward_turn_lost<MODE, OCT>(x, px, y, py, xx, yy, mu, m3mu, t, b2, b2s)) {
            alive = false;
            x = px = y = py = xx = yy = 0.0;
        }
        bsum = fma(w, MODE == kExact ? xx : x * x, bsum);
    };
    const uint32_t T = a.n_turns, K = a.every;
    double *ck = a.ckpt;
    for (uint32_t s0 = 0; s0 < T; s0 += K) {
        if (in) {
            ck[i] = x; ck[n + i] = px; ck[2 * n + i] = y; ck[3 * n + i] = py;
        }
        ck += 4 * n;
        const uint32_t s1 = min(s0 + K, T);
        uint32_t k = s0;
        if (s1 - s0 >= 6) {
            ConstPtr<TabEntry> p = tab + s0;
            ConstPtr<double> pw = weights + s0;
            TabEntry a0 = load_record(p), a1 = load_record(p + 1);
            double w0 = pw[0], w1 = pw[1];
            // turns k..k+3 use (a0, a1), (b0, b1); the next quad's pair must exist
            for (uint32_t quads = (s1 - s0 - 2) / 4; quads; --quads) {
                const TabEntry b0 = load_record(p + 2), b1 = load_record(p + 3);
                const double v0 = pw[2], v1 = pw[3];
                turn(a0, w0);
                turn(a1, w1);
                a0 = load_record(p + 4);
                a1 = load_record(p + 5);
                w0 = pw[4];
                w1 = pw[5];
                turn(b0, v0);
                turn(b1, v1);
                p = p + 4;
                pw = pw + 4;
                k += 4;
            }
        }
        for (; k < s1; ++k) turn(load_record(tab + k), weights[k]);
    }
    if (!in) return;
    if (!alive) x = px = y = py = bsum = __builtin_nan("");
    a.final_state[i] = x; a.final_state[n + i] = px; a.final_state[2 * n + i] = y; a.final_state[3 * n + i] = py;
    a.birk[i] = bsum;
    a.alive[i] = alive ? 1 : 0;
}

struct AdjointBwdArgs {
    const double *ckpt;
    uint32_t n;
    const TabEntry *tab;
    const double *dtab;       // [T][4]: d wx/d wx0, d wx/d eps, d wy/d wy0, d wy/d eps
    uint32_t n_turns, every;
    double mu;
    const double *weights;    // [T]
    const double *seed_s;     // [4][n]: dL/d(final state)
    const double *seed_b;     // [n]:    dL/dB, or NULL
    const uint8_t *alive;     // [n]
    double *sub;              // scratch [2][ceil(K/kChunk) + 1][4][n]
    double *grad_x0;          // [4][n]
    double *grad_p;           // [4][n]: d/d eps, d/d mu, d/d omega_x0, d/d omega_y0
};

// One turn of the adjoint, turn k: pre-step point `pre`, post-step point
// `post`, post-step adjoint (lx, lpx, ly, lpy) -> pre-step adjoint, with the
// parameter terms accumulated.
struct AdjAcc {
    double lx, lpx, ly, lpy;
    double g_eps, g_mu, g_wx, g_wy;
};

template <bool OCT>
__device__ __forceinline__ void adjoint_turn(const double (&pre)[4], const double (&post)[4], const TabEntry &t,
                                             double bw, const double4 &dt, double mu3, double mu6, AdjAcc &a)
{
    // Birkhoff seed of sample k+1: d/dx_{k+1} of c w_k x_{k+1}^2 (bw = 2 c w_k)
    a.lx = fma(bw, post[0], a.lx);
    // rotation angles: d x'/d w = p', d p'/d w = -x'
    const double dwx = fma(a.lx, post[1], -a.lpx * post[0]);
    const double dwy = fma(a.ly, post[3], -a.lpy * post[2]);
    a.g_wx = fma(dwx, dt.x, a.g_wx);
    a.g_wy = fma(dwy, dt.z, a.g_wy);
    a.g_eps = fma(dwx, dt.y, fma(dwy, dt.w, a.g_eps));
    // through the rotations
    const double lxt = fma(t.cx, a.lx, -t.sx * a.lpx), lpt = fma(t.sx, a.lx, t.cx * a.lpx);
    const double lyt = fma(t.cy, a.ly, -t.sy * a.lpy), lqt = fma(t.sy, a.ly, t.cy * a.lpy);
    // through the kicks, at the pre-step point
    const double x = pre[0], y = pre[2], xx = x * x, yy = y * y;
    if (OCT) a.g_mu = fma(lpt, x * fma(-3.0, yy, xx), fma(lqt, y * fma(-3.0, xx, yy), a.g_mu));
    const double ja = fma(mu3, xx - yy, 2.0 * x);
    const double jb = fma(-mu6 * x, y, -2.0 * y);
    a.lx = fma(ja, lpt, fma(jb, lqt, lxt));
    a.ly = fma(jb, lpt, fma(-ja, lqt, lyt));
    a.lpx = lpt;
    a.lpy = lqt;
}

template <int MODE, bool OCT>
__device__ __forceinline__ void step4(double (&v)[4], double mu, double m3mu, const TabEntry &t)
{
    map_step<MODE, false, OCT>(v[0], v[1], v[2], v[3], mu, m3mu, t);
}

__device__ __forceinline__ void copy4(double (&d)[4], const double (&s)[4])
{
    d[0] = s[0]; d[1] = s[1]; d[2] = s[2]; d[3] = s[3];
}

// Backward sweep. The K-segments are processed in reverse. A segment's states
// every kChunk = 8 turns (sub-checkpoints, in scratch) come from recomputing it
// forward from its checkpoint; per 8-turn chunk (in reverse) the chunk's table
// records are loaded, its states S1..S7 recomputed from S0 into registers (the
// post-step point of its last turn, `nxt`, is known), and the adjoint run back
// through its turns; S0 is the next chunk's `nxt`.
// The forward recomputation of segment s-1 is interleaved with the reverse
// sweep of segment s (one 8-turn block per chunk, into the other half of the
// scratch): the two dependency chains are independent, so one wavefront per SIMD
// (small particle counts) still keeps the FP64 pipeline busy. The state after
// segment s-1 is the checkpoint of segment s.
template <int MODE, bool OCT>
__global__ void __launch_bounds__(256) k_adjoint_backward(AdjointBwdArgs a)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const size_t n = a.n;
    if (!a.alive[i]) {
        for (int c = 0; c < 4; c++) { a.grad_x0[c * n + i] = 0.0; a.grad_p[c * n + i] = 0.0; }
        return;
    }
    AdjAcc acc{a.seed_s[i], a.seed_s[n + i], a.seed_s[2 * n + i], a.seed_s[3 * n + i], 0.0, 0.0, 0.0, 0.0};
    const double bseed = 2.0 * (a.seed_b ? a.seed_b[i] : 0.0);
    const double mu = a.mu, m3mu = -3.0 * mu, mu3 = 3.0 * mu, mu6 = 6.0 * mu;
    const uint32_t T = a.n_turns, K = a.every;
    const uint32_t n_seg = (T + K - 1) / K;
    const uint32_t seg_chunks = (K + kChunk - 1) / kChunk;  // sub-checkpoints of a full segment
    // tables and weights are read-only here and indexed by the (wave-uniform)
    // turn: scalar loads
    const ConstPtr<TabEntry> tab = as_const(a.tab);
    const ConstPtr<double> dtab = as_const(a.dtab), weights = as_const(a.weights);
    auto sub_at = [&](uint32_t seg, uint32_t c) -> double * {
        return a.sub + ((size_t)(seg & 1u) * (seg_chunks + 1) + c) * 4 * n;
    };
    auto load_ckpt = [&](uint32_t seg, double (&v)[4]) {
        const size_t o = (size_t)seg * 4 * n;
        v[0] = a.ckpt[o + i]; v[1] = a.ckpt[o + n + i]; v[2] = a.ckpt[o + 2 * n + i]; v[3] = a.ckpt[o + 3 * n + i];
    };
    auto store4 = [&](double *p, const double (&v)[4]) {
        p[i] = v[0]; p[n + i] = v[1]; p[2 * n + i] = v[2]; p[3 * n + i] = v[3];
    };
    // forward recomputation of one segment, one 8-turn block per call: writes the
    // block's sub-checkpoint, then advances unless it is the segment's last block
    struct Recompute {
        double v[4];
        uint32_t seg, done, chunks, s0;
    } rc{};
    auto rc_start = [&](uint32_t seg) {
        rc.seg = seg;
        rc.done = 0;
        rc.s0 = seg * K;
        rc.chunks = (min(rc.s0 + K, T) - rc.s0 + kChunk - 1) / kChunk;
        load_ckpt(seg, rc.v);
    };
    auto rc_block = [&]() {
        if (rc.done >= rc.chunks) return;
        store4(sub_at(rc.seg, rc.done), rc.v);
        if (rc.done + 1 < rc.chunks) {
            const uint32_t k0 = rc.s0 + rc.done * kChunk;
#pragma unroll
            for (int m = 0; m < kChunk; ++m) step4<MODE, OCT>(rc.v, mu, m3mu, load_record(tab + k0 + m));
        }
        ++rc.done;
    };

    // prologue: sub-checkpoints of the last segment
    rc_start(n_seg - 1);
    while (rc.done < rc.chunks) rc_block();
    // the state after the last turn: from the last sub-checkpoint (< 8 turns)
    double nxt[4];
    {
        const uint32_t q = rc.chunks - 1;
        const double *p = sub_at(n_seg - 1, q);
        double v[4] = {p[i], p[n + i], p[2 * n + i], p[3 * n + i]};
        for (uint32_t k = (n_seg - 1) * K + q * kChunk; k < T; ++k)
            step4<MODE, OCT>(v, mu, m3mu, load_record(tab + k));
        copy4(nxt, v);
    }
    for (int seg = (int)n_seg - 1; seg >= 0; --seg) {
        const uint32_t s0 = (uint32_t)seg * K, s1 = min(s0 + K, T);
        const bool overlap = seg > 0;
        if (overlap) rc_start((uint32_t)seg - 1);
        const uint32_t n_chunk = (s1 - s0 + kChunk - 1) / kChunk;
        for (int ch = (int)n_chunk - 1; ch >= 0; --ch) {
            const uint32_t c0 = s0 + (uint32_t)ch * kChunk;
            const uint32_t len = min((uint32_t)kChunk, s1 - c0);
            const double *q = sub_at((uint32_t)seg, (uint32_t)ch);
            double S[kChunk][4];
            S[0][0] = q[i]; S[0][1] = q[n + i]; S[0][2] = q[2 * n + i]; S[0][3] = q[3 * n + i];
            // records are loaded where they are used (scalar cache hits the
            // second time): holding a chunk's eight records would crowd the SGPRs
            auto adjoint_m = [&](int m, const double (&post)[4]) {
                const uint32_t k = c0 + m;
                const ConstPtr<double> d = dtab + 4 * (size_t)k;
                const double4 dt{d[0], d[1], d[2], d[3]};
                adjoint_turn<OCT>(S[m], post, load_record(tab + k), bseed * weights[k], dt, mu3, mu6, acc);
            };
            if (len == kChunk) {  // every chunk but possibly the very last one
#pragma unroll
                for (int m = 1; m < kChunk; ++m) {
                    copy4(S[m], S[m - 1]);
                    step4<MODE, OCT>(S[m], mu, m3mu, load_record(tab + c0 + m - 1));
                }
                // the previous segment's next 8-turn block: independent of the
                // adjoint chain below, so the two interleave
                if (overlap) rc_block();
                adjoint_m(kChunk - 1, nxt);
#pragma unroll
                for (int m = kChunk - 2; m >= 0; --m) adjoint_m(m, S[m + 1]);
            } else {
#pragma unroll
                for (int m = 1; m < kChunk; ++m) {
                    copy4(S[m], S[m - 1]);
                    if ((uint32_t)m < len) step4<MODE, OCT>(S[m], mu, m3mu, load_record(tab + c0 + m - 1));
                }
                if (overlap) rc_block();
#pragma unroll
                for (int m = kChunk - 1; m >= 0; --m) {
                    if ((uint32_t)m >= len) continue;
                    if ((uint32_t)m + 1 == len) adjoint_m(m, nxt);
                    else adjoint_m(m, S[m + 1 < kChunk ? m + 1 : m]);
                }
            }
            copy4(nxt, S[0]);
        }
        if (overlap) {
            while (rc.done < rc.chunks) rc_block();  // a short last segment leaves blocks over
            load_ckpt((uint32_t)seg, nxt);           // the state after segment seg-1
        }
    }
    a.grad_x0[i] = acc.lx; a.grad_x0[n + i] = acc.lpx; a.grad_x0[2 * n + i] = acc.ly; a.grad_x0[3 * n + i] = acc.lpy;
    a.grad_p[i] = acc.g_eps; a.grad_p[n + i] = acc.g_mu; a.grad_p[2 * n + i] = acc.g_wx; a.grad_p[3 * n + i] = acc.g_wy;
}

}  // namespace hm

using namespace hm;

extern "C" int hm_adjoint_forward(const double *x0, const double *px0, const double *y0, const double *py0, size_t n,
                                  const double *table, size_t table_len, uint32_t n_turns, double mu, double barrier,
                                  uint32_t every, const double *weights, double *checkpoints, double *final_state,
                                  double *birk, uint8_t *alive, int mode, void *stream)
{
    return guarded([&] {
        if (every == 0) throw Error(HM_ERR_ARG, "adjoint: checkpoint interval must be positive");
        if (n_turns == 0 || n_turns > table_len) throw Error(HM_ERR_RANGE, "adjoint: n_turns must lie in 1..table_len");
        if (mode != HM_MODE_EXACT && mode != HM_MODE_FAST) throw Error(HM_ERR_ARG, "adjoint: unknown mode");
        if (n == 0) return;
        AdjointFwdArgs a{x0, px0, y0, py0, (uint32_t)n, reinterpret_cast<const TabEntry *>(table), n_turns, every,
                         mu, barrier * barrier, weights, checkpoints, final_state, birk, alive};
        hipStream_t s = as_stream(stream);
        const uint32_t grid = (uint32_t)((n + 255) / 256);
        const bool oct = mu != 0.0;
        LaunchTimer timer("k_adjoint_forward", s);
        if (mode == HM_MODE_EXACT) {
            if (oct) hipLaunchKernelGGL((k_adjoint_forward<kExact, true>), dim3(grid), dim3(256), 0, s, a);
            else hipLaunchKernelGGL((k_adjoint_forward<kExact, false>), dim3(grid), dim3(256), 0, s, a);
        } else {
            if (oct) hipLaunchKernelGGL((k_adjoint_forward<kFast, true>), dim3(grid), dim3(256), 0, s, a);
            else hipLaunchKernelGGL((k_adjoint_forward<kFast, false>), dim3(grid), dim3(256), 0, s, a);
        }
        launch_check("k_adjoint_forward");
    });
}

extern "C" int hm_adjoint_scratch_bytes(size_t n, uint32_t every, size_t *bytes)
{
    return guarded([&] {
        if (!bytes || every == 0) throw Error(HM_ERR_ARG, "adjoint: bad arguments");
        // two halves (the segment being swept and the one being recomputed)
        *bytes = 2 * ((size_t)(every + kChunk - 1) / kChunk + 1) * 4 * n * sizeof(double);
    });
}

extern "C" int hm_adjoint_backward(const double *checkpoints, size_t n, const double *table, const double *dtable,
                                   size_t table_len, uint32_t n_turns, double mu, uint32_t every,
                                   const double *weights, const double *seed_state, const double *seed_birk,
                                   const uint8_t *alive, double *scratch, double *grad_x0, double *grad_params,
                                   int mode, void *stream)
{
    return guarded([&] {
        if (every == 0) throw Error(HM_ERR_ARG, "adjoint: checkpoint interval must be positive");
        if (n_turns == 0 || n_turns > table_len) throw Error(HM_ERR_RANGE, "adjoint: n_turns must lie in 1..table_len");
        if (mode != HM_MODE_EXACT && mode != HM_MODE_FAST) throw Error(HM_ERR_ARG, "adjoint: unknown mode");
        if (n == 0) return;
        if (!checkpoints || !table || !dtable || !weights || !seed_state || !alive || !scratch || !grad_x0 ||
            !grad_params)
            throw Error(HM_ERR_ARG, "adjoint: NULL pointer");
        AdjointBwdArgs a{checkpoints, (uint32_t)n, reinterpret_cast<const TabEntry *>(table), dtable, n_turns,
                         every, mu, weights, seed_state, seed_birk, alive, scratch, grad_x0, grad_params};
        hipStream_t s = as_stream(stream);
        const uint32_t grid = (uint32_t)((n + 255) / 256);
        const bool oct = mu != 0.0;
        LaunchTimer timer("k_adjoint_backward", s);
        if (mode == HM_MODE_EXACT) {
            if (oct) hipLaunchKernelGGL((k_adjoint_backward<kExact, true>), dim3(grid), dim3(256), 0, s, a);
            else hipLaunchKernelGGL((k_adjoint_backward<kExact, false>), dim3(grid), dim3(256), 0, s, a);
        } else {
            if (oct) hipLaunchKernelGGL((k_adjoint_backward<kFast, true>), dim3(grid), dim3(256), 0, s, a);
            else hipLaunchKernelGGL((k_adjoint_backward<kFast, false>), dim3(grid), dim3(256), 0, s, a);
        }
        launch_check("k_adjoint_backward");
    });
}

extern "C" int hm_lines_modulation_derivatives(double tune, double epsilon, const double *coef, const double *freq,
                                               int n_lines, int start, int end, double *out)
{
    // omega_k = 2 pi tune (1 + eps S_k), S_k = sum_j coef_j cos(freq_j (k+start)):
    // out[k] = (d omega_k / d omega_0, d omega_k / d eps) with omega_0 = 2 pi tune
    return guarded([&] {
        if (!(start < end)) throw Error(HM_ERR_ARG, "modulation: start must be smaller than end");
        if (!out || (n_lines && (!coef || !freq))) throw Error(HM_ERR_ARG, "modulation: NULL pointer");
        for (int i = 0; i < end - start; i++) {
            double sum = 0.0;
            for (int j = 0; j < n_lines; j++) sum += coef[j] * std::cos(freq[j] * (double)((unsigned)i + start));
            out[2 * i] = 1 + epsilon * sum;
            out[2 * i + 1] = 2 * M_PI * tune * sum;
        }
    });
}
