// Tangent-map kernels.
//  * k_megno_trace: henon_tracker::megno (src/henon.cu:2755-2887) on the GPU, one
//    thread per particle, the 4x4 tangent matrix in registers.
//  * k_indicators : fused tracking + tangent matrix (structured J*M, scalar
//    renormalisation) + MEGNO (the construct form of src/henon.cu:2253-2284) +
//    FLI + online Birkhoff tunes + a trajectory window in HBM for FFT tunes.
//  * per-call construct kernels (src/henon.cu:267-359, 2083-2098, 2253-2284,
//    2366-2402, 2533-2584) for the reference's construct classes.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

#include "hm_device.h"
#include "hm_internal.h"
#include "hm_tangent.h"

namespace hm {

constexpr int kMaxSamples = 16;
constexpr int kMaxBirk = 4;

// ---------------------------------------------------------------- trace MEGNO
struct MegnoTraceArgs {
    double *x, *px, *y, *py;
    uint32_t *steps;
    uint32_t n;
    const TabEntry *tab;
    uint32_t tab_len, n_turns;
    double mu, b2;
    uint32_t samples[kMaxSamples];
    uint32_t n_samples;
    double *out;  // [n][n_samples], the reference's per-particle row layout
    const double *m_in;  // [n][16] initial matrices, or NULL (identity)
    double *m_out;       // [n][16] final matrices, or NULL
    uint32_t *status;
};

// Full 4x4 product in the reference's order (henon.cu:251-265): i-j-k, += from 0.0.
__device__ __forceinline__ void mat_mul_full(const double *a, double *m)
{
    double r[16];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < 4; k++) acc = acc + a[i * 4 + k] * m[k * 4 + j];
            r[i * 4 + j] = acc;
        }
#pragma unroll
    for (int i = 0; i < 16; i++) m[i] = r[i];
}

// One thread per particle. Per turn k: J at the pre-step point with table
// index k (the reference ignores the particle's turn counter here, Q4), M <- J M,
// a_k = (k+1) log10(tr M_{k+1} / tr M_k), then the map step with table index
// steps. The two prefix sums are accumulated online in the same order.
// Every particle starts from the identity (the reference reuses the
// off-diagonal entries of the previous particle its CPU thread handled).
template <bool OCT, bool INV>
__global__ void __launch_bounds__(256) k_megno_trace(MegnoTraceArgs a)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double x = a.x[i], px = a.px[i], y = a.y[i], py = a.py[i];
    uint32_t s = a.steps[i];
    double m[16];
#pragma unroll
    for (int q = 0; q < 16; q++) m[q] = a.m_in ? a.m_in[(size_t)i * 16 + q] : ((q % 5 == 0) ? 1.0 : 0.0);
    double trace_1 = 4.0, trace_2 = 4.0, s1 = 0.0, s2 = 0.0;
    double res[kMaxSamples];
#pragma unroll
    for (int q = 0; q < kMaxSamples; q++) res[q] = 0.0;  // slot n_turns stays 0 (Q4)
    const double mu3 = a.mu * 3, mu6 = a.mu * 6, m3mu = -3.0 * a.mu;
    bool alive = !any_nan(x, px, y, py);
    for (uint32_t k = 0; k < a.n_turns; ++k) {
        if (k >= a.tab_len) break;
        const TabEntry tj = a.tab[k];
        const Jac jc = jac_coeffs<kExact>(x, y, mu3, mu6);
        double jm[16];
        jac_dense_exact(jc, tj, jm);
        mat_mul_full(jm, m);
        trace_2 = trace_1;
        trace_1 = ((m[0] + m[5]) + m[10]) + m[15];
        const double ak = (double)(k + 1) * log10(trace_1 / trace_2);
        s1 = s1 + ak;
        s2 = s2 + s1 / (double)(k + 1);
#pragma unroll
        for (int q = 0; q < kMaxSamples; q++)
            if (q < (int)a.n_samples && a.samples[q] == k) res[q] = s2 / (double)a.samples[q];
        if (alive) {
            // the map's own table index: steps forward, steps - 1 inverse
            if (INV ? (s == 0 || s > a.tab_len) : s >= a.tab_len) {
                atomicOr(a.status, 1u);
                alive = false;
            } else {
                map_step<kExact, INV, OCT>(x, px, y, py, a.mu, m3mu, a.tab[INV ? s - 1 : s]);
                if (radius2<kExact>(x, px, y, py) < a.b2) {
                    s = INV ? s - 1 : s + 1;
                } else {
                    x = px = y = py = __builtin_nan("");
                    alive = false;
                }
            }
        }
    }
    a.x[i] = x; a.px[i] = px; a.y[i] = y; a.py[i] = py; a.steps[i] = s;
    if (a.m_out)
#pragma unroll
        for (int q = 0; q < 16; q++) a.m_out[(size_t)i * 16 + q] = m[q];
#pragma unroll
    for (int q = 0; q < kMaxSamples; q++)
        if (q < (int)a.n_samples) a.out[(size_t)i * a.n_samples + q] = res[q];
}

// ---------------------------------------------------------------- fused indicators
struct IndicatorArgs {
    double *x, *px, *y, *py;
    uint32_t *steps;
    uint32_t n;
    const TabEntry *tab;
    uint32_t tab_len, n_turns;
    double mu, b2;
    uint32_t renorm_every;
    uint32_t samples[kMaxSamples];  // turn counts 1..n_turns, increasing
    uint32_t n_samples;
    double fli_dir[4];
    double *megno, *fli;    // [n_samples][n]
    double *matrix;         // [n][16] in/out, or NULL (identity start, not stored)
    double *log_scale;      // [n] in/out log10 of the factored-out scale, or NULL
    uint32_t birk_from[kMaxBirk], birk_to[kMaxBirk], birk_off[kMaxBirk];
    uint32_t n_birk;
    const double *birk_w;   // concatenated weight tables
    double *birk_out;       // [n_birk][2][n] tunes
    uint32_t hist_from, hist_len;
    double *hist;           // [hist_len][4][n]
    // construct-form turn-sample windows (megno_birkhoff / tune_birkhoff)
    uint32_t mb_T[kMaxBirk], mb_off[kMaxBirk], n_mb;
    double *mb_out;         // [n_mb][n]
    uint32_t tb_T[kMaxBirk], tb_off[kMaxBirk], n_tb;
    double *tb_out;         // [n_tb][4][n]
    const double *inv_idx;  // [n_turns]: 1.0 / (k + 1), correctly rounded (host table)
    uint32_t *status;
    // Segments with loss compaction: this launch runs turns [k0, k1) of the call
    // over the live list idx_in (count on the device; NULL: every particle) and
    // appends its survivors to idx_out; between launches each live particle's
    // tangent matrix and sums wait in state[j] (ind_state_len(V) doubles). A single
    // launch has k0 = 0, k1 = n_turns, no lists and no state.
    uint32_t k0, k1;
    const uint32_t *idx_in, *count_in;
    uint32_t *idx_out, *count_out;
    double *state;
    uint32_t first, last;
};

// Per-particle state between segments: M (16), log10 scale, L1, L2, |M|_F^2;
// with Birkhoff windows the last angles (2) and the window sums (2 per window);
// with construct windows the megno_birkhoff sums (1 per window), the
// tune_birkhoff sums (4 per window) and their last angles (2).
constexpr int kIndBirkState = 20;                          // offset of the Birkhoff part
constexpr int kIndConsState = kIndBirkState + 2 + 2 * kMaxBirk;  // offset of the construct part
// doubles of state per particle in variant V (below)
constexpr int ind_state_len(int V) { return V == 2 ? kIndConsState + 5 * kMaxBirk + 2 : V == 1 ? kIndConsState : kIndBirkState; }


// One thread per particle; the table index of turn k is steps+k for both the
// tangent matrix (evaluated at the pre-step point) and the step.
//   MEGNO (construct form): with idx = k+1, L1 += log10(|M_{k+1}|_F / |M_k|_F) idx,
//     L2 += L1 / idx; recorded at sample t as L2 / t. The ratio is evaluated as
//     0.5 log10(|M_{k+1}|^2 / |M_k|^2) with fused squared norms (one division,
//     no square root), the logarithm by log10_fdl (hm_log10.h: within 2 ulps
//     of the library's, in about half its instructions) and 1/idx from a host
//     table; the result agrees with the separate-rounding form to a few ulps
//     per turn.
//   FLI: log10(|M_t v0|) / t with v0 = fli_dir / |fli_dir|, including the
//     log10 of the scale factored out by renormalisation.
//   Renormalisation every renorm_every turns: M <- M / |M|_F, after the MEGNO
//     update of that turn; the next ratio uses the norm of the scaled matrix.
//   Birkhoff windows (sample indices from < m <= to): acc += w[m-1-from] dtheta_m
//     with dtheta wrapped into [0, 2pi) (dynamic_indicator.cpp:110-117); tune =
//     1 - acc / 2pi (NaN once a sample of the window is NaN).
// Variants V: 0 MEGNO/FLI only; 1 with Birkhoff windows; 2 with any of the
// trajectory window and the construct windows as well. Every variant runs the
// straight runs of plain turns and loss-compaction segments (its sums kept in
// the state between them); a particle lost inside a run fed its sums and its
// trajectory window values after the loss, so the run's end writes NaN into its
// window from the loss on and the outputs of its sums that reach past the loss
// are NaN, as the turn-by-turn sums make them.
// tune_birkhoff_construct's add of index idx to one turn sample T (weights
// w[0..T)): idx < T into tune 1, T <= idx < 2T into tune 2 (sums x1, x2, y1,
// y2). Both halves are written every time, so the four sums stay registers.
__device__ inline void tune_birkhoff_add(double (&acc)[4], uint32_t idx, uint32_t T, const double *w, double dx,
                                         double dy)
{
    if ((uint64_t)idx >= 2ull * T) return;
    const bool lo = idx < T;
    const double wt = w[lo ? idx : idx - T];
    const double ax = dx * wt, ay = dy * wt;
    acc[0] = lo ? acc[0] + ax : acc[0];
    acc[2] = lo ? acc[2] + ay : acc[2];
    acc[1] = lo ? acc[1] : acc[1] + ax;
    acc[3] = lo ? acc[3] : acc[3] + ay;
}

template <int MODE, bool OCT, int V>
__global__ void __launch_bounds__(256) k_indicators(IndicatorArgs a)
{
    constexpr bool EXTRA = V == 2, BW = V >= 1;
    constexpr int SL = ind_state_len(V);
    constexpr int MAP = map_mode(MODE), TAN = tangent_mode(MODE);
    const uint32_t li = blockIdx.x * blockDim.x + threadIdx.x;  // position in the live list
    const uint32_t n_live = a.count_in ? *a.count_in : a.n;
    const bool in_range = li < n_live;
    const uint32_t i = in_range ? (a.idx_in ? a.idx_in[li] : li) : 0;  // the particle
    const uint32_t k0 = a.k0, k1 = a.k1;
    double x = __builtin_nan(""), px = x, y = x, py = x;
    uint32_t s = 0;
    if (in_range) {
        x = a.x[i]; px = a.px[i]; y = a.y[i]; py = a.py[i];
        s = a.steps[i];
    }
    double m[16];
    double lscale = 0.0, l1 = 0.0, l2 = 0.0, f0;
    double bacc[kMaxBirk][2];
#pragma unroll
    for (int w = 0; w < kMaxBirk; w++) bacc[w][0] = bacc[w][1] = 0.0;
    double thx = 0.0, thy = 0.0;
    if (!a.first && in_range) {
        const double *st = a.state + (size_t)i * SL;
#pragma unroll
        for (int q = 0; q < 16; q++) m[q] = st[q];
        lscale = st[16]; l1 = st[17]; l2 = st[18]; f0 = st[19];
        if (BW && a.n_birk) {
            thx = st[kIndBirkState]; thy = st[kIndBirkState + 1];
#pragma unroll
            for (int w = 0; w < kMaxBirk; w++) {
                bacc[w][0] = st[kIndBirkState + 2 + 2 * w];
                bacc[w][1] = st[kIndBirkState + 3 + 2 * w];
            }
        }
    } else {
        if (a.matrix && in_range) {
#pragma unroll
            for (int q = 0; q < 16; q++) m[q] = a.matrix[(size_t)i * 16 + q];
        } else {
#pragma unroll
            for (int q = 0; q < 16; q++) m[q] = (q % 5 == 0) ? 1.0 : 0.0;
        }
        lscale = (a.log_scale && in_range) ? a.log_scale[i] : 0.0;
        f0 = frob2_fused(m);
    }
    const double mu3 = a.mu * 3, mu6 = a.mu * 6, m3mu = -3.0 * a.mu;
    const bool alive0 = in_range && !any_nan(x, px, y, py);
    // held: some but not all coordinates NaN. The particle is not tracked (its
    // coordinates and steps stay), but its tangent matrix and sums go on at the
    // held point with the call's table records (index steps + call turn), so it
    // is carried between segments like a live particle; the straight runs,
    // which move every lane's coordinates, are off for its wavefront.
    const bool held = in_range && !alive0 && !(__builtin_isnan(x) && __builtin_isnan(px) &&
                                               __builtin_isnan(y) && __builtin_isnan(py));
    if (held) s += k0;
    bool alive = alive0;
    uint32_t done = k1;  // the turn a particle is lost in (k1: survived the launch)
    double v0[4];
    {
        const double nn = sqrt(((a.fli_dir[0] * a.fli_dir[0] + a.fli_dir[1] * a.fli_dir[1]) +
                                a.fli_dir[2] * a.fli_dir[2]) + a.fli_dir[3] * a.fli_dir[3]);
#pragma unroll
        for (int q = 0; q < 4; q++) v0[q] = a.fli_dir[q] / nn;
    }
    if (BW && a.n_birk && a.first) {
        thx = atan2(px, x);
        thy = atan2(py, y);
    }
    // construct-form accumulators: megno_birkhoff L2_j, tune_birkhoff
    // (tune1 x, tune2 x, tune1 y, tune2 y)_j and the stored angles
    double mbacc[kMaxBirk], tbacc[kMaxBirk][4];
#pragma unroll
    for (int w = 0; w < kMaxBirk; w++) {
        mbacc[w] = 0.0;
        tbacc[w][0] = tbacc[w][1] = tbacc[w][2] = tbacc[w][3] = 0.0;
    }
    double ctx = 0.0, cty = 0.0;
    if (EXTRA && (a.n_mb || a.n_tb)) {
        if (!a.first && in_range) {
            const double *st = a.state + (size_t)i * SL + kIndConsState;
#pragma unroll
            for (int w = 0; w < kMaxBirk; w++) {
                mbacc[w] = st[w];
#pragma unroll
                for (int c = 0; c < 4; c++) tbacc[w][c] = st[kMaxBirk + 4 * w + c];
            }
            ctx = st[5 * kMaxBirk];
            cty = st[5 * kMaxBirk + 1];
        } else if (a.n_tb) {
            ctx = atan2(px, x);
            cty = atan2(py, y);
        }
    }
    if (EXTRA && a.first && in_range && a.hist_len && a.hist_from == 0) {
        const size_t n = a.n;
        a.hist[0 * n + i] = x; a.hist[1 * n + i] = px; a.hist[2 * n + i] = y; a.hist[3 * n + i] = py;
    }
    // wave-uniform turn index when the live particles share one (scalar table loads)
    const uint64_t live = __ballot(alive0);
    uint32_t s0 = 0;
    if (live) s0 = __builtin_amdgcn_readlane(s, __builtin_ctzll(live));
    const bool uniform = live && __all(!alive0 || s == s0) && !__any(held) &&
                         (uint64_t)s0 + (k1 - k0) <= (uint64_t)a.tab_len;
    // record of call turn k at tp + k (s0 = the shared index at turn k0)
    const ConstPtr<TabEntry> tp = as_const(a.tab + s0) - k0;
    const ConstPtr<double> inv_idx = as_const(a.inv_idx);

    // samples already behind (taken by earlier segments)
    uint32_t sample_q = 0;  // FLI samples (taken at their turn)
    while (sample_q < a.n_samples && a.samples[sample_q] <= k0) ++sample_q;
    uint32_t sample_m = sample_q;  // MEGNO samples (taken once the turn's sums are in)
    uint32_t next_sample = sample_q < a.n_samples ? a.samples[sample_q] : 0;
    uint32_t next_megno = next_sample;
    uint32_t k = k0;
    // squared Frobenius norm of M before the turn; the MEGNO ratio
    // |M_{k+1}|/|M_k| is taken as sqrt(fa2 / fb2): one division, no square root
    double fb2 = f0;
    uint32_t renorm_in = a.renorm_every ? a.renorm_every - k0 % a.renorm_every : 0;
    double xx = x * x, yy = y * y;  // carried between turns in exact mode
    const double b2s = screen_bound(a.b2);
    // The MEGNO sums of turn k (norm of M_{k+1}, ratio, log10, L1, L2) are
    // formed at the top of turn k+1, where they do not wait on and are not
    // waited on by that turn's map and matrix product; a renormalisation turn
    // forms them at once, before it rescales M.
    bool pending = false;
    double w_prev = 0.0, inv_prev = 0.0;
    uint32_t idx_prev = 0;
    // megno_birkhoff_construct after the MEGNO sums of add index idx (turn
    // idx - 1): L2_j += L1 w_j[idx-1] / idx for idx <= T_j
    auto megno_birkhoff_turn = [&](uint32_t idx, double inv) {
#pragma unroll
        for (int w = 0; w < kMaxBirk; w++)
            if (w < (int)a.n_mb && idx <= a.mb_T[w])
                mbacc[w] = mbacc[w] + (l1 * a.birk_w[a.mb_off[w] + idx - 1]) * inv;
    };
    auto megno_update = [&]() {
        const double fa2 = frob2_fused(m);
        l1 = l1 + (0.5 * log10_fdl(fa2 / fb2)) * w_prev;
        l2 = l2 + inv_prev * l1;
        fb2 = fa2;
        if (EXTRA && a.n_mb) megno_birkhoff_turn(idx_prev, inv_prev);
        if (idx_prev == next_megno) {
            if (in_range && a.megno) a.megno[(size_t)sample_m * a.n + i] = inv_prev * l2;
            ++sample_m;
            next_megno = sample_m < a.n_samples ? a.samples[sample_m] : 0;
        }
        pending = false;
    };
    // Birkhoff windows after turn idx - 1 (the sample idx is the state now)
    auto birkhoff_turn = [&](uint32_t idx) {
        const double nx = atan2(px, x), ny = atan2(py, y);
        double dx = nx - thx, dy = ny - thy;
        if (dx < 0.0) dx += 2.0 * M_PI;
        if (dy < 0.0) dy += 2.0 * M_PI;
        thx = nx;
        thy = ny;
#pragma unroll
        for (int w = 0; w < kMaxBirk; w++) {
            if (w < (int)a.n_birk && idx > a.birk_from[w] && idx <= a.birk_to[w]) {
                const double wt = a.birk_w[a.birk_off[w] + (idx - 1 - a.birk_from[w])];
                bacc[w][0] = bacc[w][0] + wt * dx;
                bacc[w][1] = bacc[w][1] + wt * dy;
            }
        }
    };
    // tune_birkhoff_construct's add after turn idx - 1 (add index idx)
    auto tune_birkhoff_turn = [&](uint32_t idx) {
        const double nx = atan2(px, x), ny = atan2(py, y);
        double dx = nx - ctx, dy = ny - cty;
        dx = dx > 0 ? dx : dx + 2 * M_PI;
        dy = dy > 0 ? dy : dy + 2 * M_PI;
        ctx = nx;
        cty = ny;
#pragma unroll
        for (int w = 0; w < kMaxBirk; w++)
            if (w < (int)a.n_tb) tune_birkhoff_add(tbacc[w], idx, a.tb_T[w], a.birk_w + a.tb_off[w], dx, dy);
    };
    // the trajectory window's sample idx (the state after turn idx - 1)
    auto hist_store = [&](uint32_t idx, double vx, double vpx, double vy, double vpy) {
        if (in_range && idx >= a.hist_from && idx < a.hist_from + a.hist_len) {
            const size_t n = a.n, base = (size_t)(idx - a.hist_from) * 4 * n;
            a.hist[base + 0 * n + i] = vx; a.hist[base + 1 * n + i] = vpx;
            a.hist[base + 2 * n + i] = vy; a.hist[base + 3 * n + i] = vpy;
        }
    };
    for (; k < k1; ++k) {
        if (uniform && pending) {
            // Plain turns up to the next event (an FLI sample at its turn, a
            // MEGNO sample at the top of the turn after it, a renormalisation
            // turn): no per-turn checks, so the turn is one straight block in
            // which the previous turn's MEGNO sums, the matrix product and the
            // map interleave. The next turn's table record and reciprocal are
            // requested once this turn's have arrived (scalar loads complete
            // together), a turn ahead of their use.
            uint32_t stop = k1;
            if (sample_q < a.n_samples) stop = min(stop, next_sample - 1);
            if (sample_m < a.n_samples) stop = min(stop, next_megno);
            if (renorm_in) stop = min(stop, k + renorm_in - 1);
            if (stop > k) {
                // a particle lost in the run is marked (its loss turn kept) and
                // set to NaN when the run ends: nothing reads it before that
                const bool alive_start = alive;
                TabEntry t = load_record(tp + k);
                double inv_k = inv_idx[k];
                for (uint32_t kk = k; kk < stop; ++kk) {
                    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): this turn's entries
                    __builtin_amdgcn_sched_barrier(0);
                    const uint32_t kn = min(kk + 1, k1 - 1);
                    const TabEntry tn = load_record(tp + kn);
                    const double inv_n = inv_idx[kn];
                    {
                        const double fa2 = frob2_fused(m);
                        l1 = l1 + (0.5 * log10_fdl(fa2 / fb2)) * w_prev;
                        l2 = l2 + inv_prev * l1;
                        fb2 = fa2;
                    }
                    if (EXTRA && a.n_mb) megno_birkhoff_turn(idx_prev, inv_prev);
                    const Jac jc = jac_coeffs<TAN>(x, y, mu3, mu6);
                    mat_mul_structured<TAN>(jc, t, m);
                    const bool out = forward_turn_out<MAP, OCT>(x, px, y, py, xx, yy, a.mu, m3mu, t, a.b2, b2s);
                    done = (alive & out) ? kk : done;
                    alive = alive & !out;
                    if (BW && a.n_birk) birkhoff_turn(kk + 1);
                    if (EXTRA && a.n_tb) tune_birkhoff_turn(kk + 1);
                    if (EXTRA && a.hist_len && alive) hist_store(kk + 1, x, px, y, py);
                    w_prev = (double)(kk + 1);
                    inv_prev = inv_k;
                    idx_prev = kk + 1;
                    t = tn;
                    inv_k = inv_n;
                }
                if (alive_start && !alive) {
                    x = px = y = py = xx = yy = __builtin_nan("");
#pragma unroll
                    for (int q = 0; q < 16; q++) m[q] = __builtin_nan("");
                    // the run fed the lost particle's windows with its values after
                    // the loss; sample done + 1 on is NaN, as are those windows
                    if (BW && a.n_birk) {
                        thx = thy = __builtin_nan("");
#pragma unroll
                        for (int w = 0; w < kMaxBirk; w++)
                            if (w < (int)a.n_birk && a.birk_to[w] >= done + 1) bacc[w][0] = bacc[w][1] = __builtin_nan("");
                    }
                }
                if (renorm_in) renorm_in -= stop - k;
                k = stop;
                if (k == k1 || !__any(alive)) break;
            }
        }
        if (pending) megno_update();
        TabEntry t;
        if (uniform) {
            t = load_record(tp + k);
        } else {
            const uint64_t ti = (uint64_t)s + (k - k0);
            if (alive && ti >= a.tab_len) {
                atomicOr(a.status, 1u);
                alive = false;
                done = k;
            }
            t = a.tab[ti < a.tab_len ? ti : a.tab_len - 1];
        }
        const Jac jc = jac_coeffs<TAN>(x, y, mu3, mu6);
        mat_mul_structured<TAN>(jc, t, m);
        if (alive && forward_turn_lost<MAP, OCT>(x, px, y, py, xx, yy, a.mu, m3mu, t, a.b2, b2s)) {
            x = px = y = py = xx = yy = __builtin_nan("");
#pragma unroll
            for (int q = 0; q < 16; q++) m[q] = __builtin_nan("");
            alive = false;
            done = k;
        }
        const uint32_t idx = k + 1;
        pending = true;
        w_prev = (double)idx;
        inv_prev = inv_idx[k];
        idx_prev = idx;
        if (idx == next_sample) {
            const size_t o = (size_t)sample_q * a.n + i;
            if (in_range && a.fli) {
                double v[4];
#pragma unroll
                for (int r = 0; r < 4; r++)
                    v[r] = (((m[r * 4] * v0[0]) + (m[r * 4 + 1] * v0[1])) + (m[r * 4 + 2] * v0[2])) +
                           (m[r * 4 + 3] * v0[3]);
                const double nv = sqrt((((v[0] * v[0]) + (v[1] * v[1])) + (v[2] * v[2])) + (v[3] * v[3]));
                a.fli[o] = (lscale + log10(nv)) / (double)idx;
            }
            ++sample_q;
            next_sample = sample_q < a.n_samples ? a.samples[sample_q] : 0;
        }
        if (renorm_in && --renorm_in == 0) {
            renorm_in = a.renorm_every;
            megno_update();
            if (alive) {
                const double fa = sqrt(frob2<TAN>(m));
                const double inv = 1.0 / fa;
#pragma unroll
                for (int q = 0; q < 16; q++) m[q] = m[q] * inv;
                lscale = lscale + log10(fa);
                fb2 = frob2_fused(m);
            }
        }
        if (BW && a.n_birk) birkhoff_turn(idx);
        if (EXTRA && a.n_tb) tune_birkhoff_turn(idx);
        if (EXTRA && a.hist_len && alive) hist_store(idx, x, px, y, py);
        if (!__any(alive || held)) break;
    }
    // the last turn's MEGNO sums
    if (pending) megno_update();
    // survivors of a segment that is not the last carry on in the next launch
    const bool carry = !a.last && (alive || held);
    if (a.idx_out) {
        const uint64_t mlive = __ballot(in_range && carry);
        if (mlive) {
            const uint32_t lane = __lane_id();
            const int leader = __builtin_ctzll(mlive);
            uint32_t slot = 0;
            if ((int)lane == leader) slot = atomicAdd(a.count_out, (uint32_t)__popcll(mlive));
            slot = __builtin_amdgcn_readlane(slot, leader);
            const uint64_t bit = 1ull << lane;
            if (mlive & bit) a.idx_out[slot + __popcll(mlive & (bit - 1ull))] = i;
        }
    }
    if (!in_range) return;
    if (alive0) {
        a.x[i] = x; a.px[i] = px; a.y[i] = y; a.py[i] = py;
        a.steps[i] = s + (done - k0);
    }
    if (carry) {
        double *st = a.state + (size_t)i * SL;
#pragma unroll
        for (int q = 0; q < 16; q++) st[q] = m[q];
        st[16] = lscale; st[17] = l1; st[18] = l2; st[19] = fb2;
        if (BW && a.n_birk) {
            st[kIndBirkState] = thx; st[kIndBirkState + 1] = thy;
#pragma unroll
            for (int w = 0; w < kMaxBirk; w++) {
                st[kIndBirkState + 2 + 2 * w] = bacc[w][0];
                st[kIndBirkState + 3 + 2 * w] = bacc[w][1];
            }
        }
        if (EXTRA && (a.n_mb || a.n_tb)) {
            double *sc = st + kIndConsState;
#pragma unroll
            for (int w = 0; w < kMaxBirk; w++) {
                sc[w] = mbacc[w];
#pragma unroll
                for (int c = 0; c < 4; c++) sc[kMaxBirk + 4 * w + c] = tbacc[w][c];
            }
            sc[5 * kMaxBirk] = ctx;
            sc[5 * kMaxBirk + 1] = cty;
        }
        return;
    }
    // samples not reached because the particle (or its whole wavefront) was lost
    for (; sample_q < a.n_samples; ++sample_q)
        if (a.fli) a.fli[(size_t)sample_q * a.n + i] = __builtin_nan("");
    for (; sample_m < a.n_samples; ++sample_m)
        if (a.megno) a.megno[(size_t)sample_m * a.n + i] = __builtin_nan("");
    if (a.matrix) {
#pragma unroll
        for (int q = 0; q < 16; q++) a.matrix[(size_t)i * 16 + q] = m[q];
    }
    if (a.log_scale) a.log_scale[i] = lscale;
    if (EXTRA && !alive0) {
        // a particle NaN from the start is not tracked: its coordinates stay
        // as given, so every trajectory sample after the first is the start
        // and every tune_birkhoff add takes the same angles (a straight run
        // moved the registers' copies of an all-NaN particle, and its
        // wavefront may have stopped early; the arrays hold the start)
        x = a.x[i]; px = a.px[i]; y = a.y[i]; py = a.py[i];
        if (a.n_tb) {
            const double ax = atan2(px, x), ay = atan2(py, y);
            double dx = ax - ax, dy = ay - ay;
            dx = dx > 0 ? dx : dx + 2 * M_PI;
            dy = dy > 0 ? dy : dy + 2 * M_PI;
#pragma unroll
            for (int w = 0; w < kMaxBirk; w++) {
                if (w < (int)a.n_tb) {
                    const uint64_t T = a.tb_T[w];
                    tbacc[w][0] = tbacc[w][1] = tbacc[w][2] = tbacc[w][3] = 0.0;
                    const uint64_t hi = min<uint64_t>(2 * T - 1, a.n_turns);
                    for (uint64_t q = 1; q <= hi; ++q)
                        tune_birkhoff_add(tbacc[w], (uint32_t)q, (uint32_t)T, a.birk_w + a.tb_off[w], dx, dy);
                }
            }
        }
    }
    if (EXTRA && a.hist_len && (!alive0 || done < k1)) {
        // samples from the loss on (lost particles: NaN) or after the start
        // (particles NaN from the start: the start), which the turns above
        // did not store
        const uint32_t from = max(alive0 ? done + 1 : 1u, a.hist_from);
        for (uint64_t q = from; q < (uint64_t)a.hist_from + a.hist_len; ++q) hist_store((uint32_t)q, x, px, y, py);
    }
#pragma unroll
    for (int w = 0; w < kMaxBirk; w++) {
        if (BW && w < (int)a.n_birk) {
            // a window holding a NaN sample is NaN: the particle was NaN from the
            // start, or lost at turn done (sample done + 1 on), which a wavefront
            // that stopped early did not accumulate
            const bool nan_w = !alive0 || (done < k1 && a.birk_to[w] >= done + 1);
            a.birk_out[((size_t)w * 2 + 0) * a.n + i] = nan_w ? __builtin_nan("") : 1.0 - bacc[w][0] / (2 * M_PI);
            a.birk_out[((size_t)w * 2 + 1) * a.n + i] = nan_w ? __builtin_nan("") : 1.0 - bacc[w][1] / (2 * M_PI);
        }
        // the construct sums of a particle lost at turn done: NaN once they take
        // a sample from done + 1 on (megno_birkhoff: idx <= T; tune_birkhoff:
        // 1 <= idx < T for the first half, T <= idx < 2T, idx <= n_turns for
        // the second); a particle all NaN from the start has NaN MEGNO sums
        // from the first turn on, a held one its own, and both the
        // tune_birkhoff sums formed above
        const bool lost = alive0 && done < k1;
        if (EXTRA && w < (int)a.n_mb) {
            const bool nan_m = (!alive0 && !held) || (lost && a.mb_T[w] >= done + 1);
            a.mb_out[(size_t)w * a.n + i] = nan_m ? __builtin_nan("") : mbacc[w];
        }
        if (EXTRA && w < (int)a.n_tb) {
            const uint64_t T = a.tb_T[w];
            const bool nan1 = lost && (uint64_t)done + 1 < T;
            const bool nan2 = lost && (uint64_t)done + 1 < 2 * T && T <= a.n_turns;
#pragma unroll
            for (int c = 0; c < 4; c++)
                a.tb_out[((size_t)w * 4 + c) * a.n + i] = ((c % 2 == 0) ? nan1 : nan2) ? __builtin_nan("") : tbacc[w][c];
        }
    }
}

// ---------------------------------------------------------------- construct kernels
// henon.cu:267-313: M <- J(table[ti]) M (multiply=1) or M <- J (multiply=0),
// full product in the reference's order.
__global__ void k_tangent_apply(double *__restrict__ mats, const double *__restrict__ x, const double *__restrict__ px,
                                const double *__restrict__ y, const double *__restrict__ py, uint32_t n,
                                const TabEntry *__restrict__ tab, uint32_t ti, double mu, int multiply)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    (void)px; (void)py;
    const Jac jc = jac_coeffs<kExact>(x[i], y[i], mu * 3, mu * 6);
    double jm[16], m[16];
    jac_dense_exact(jc, tab[ti], jm);
    if (multiply) {
#pragma unroll
        for (int q = 0; q < 16; q++) m[q] = mats[(size_t)i * 16 + q];
        mat_mul_full(jm, m);
    } else {
#pragma unroll
        for (int q = 0; q < 16; q++) m[q] = jm[q];
    }
#pragma unroll
    for (int q = 0; q < 16; q++) mats[(size_t)i * 16 + q] = m[q];
}

// matrix_4d_vector::multiply (henon.cu:1721-1741): M <- L M, L given per particle
__global__ void k_left_multiply(double *__restrict__ mats, const double *__restrict__ left, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double l[16], m[16];
#pragma unroll
    for (int q = 0; q < 16; q++) { l[q] = left[(size_t)i * 16 + q]; m[q] = mats[(size_t)i * 16 + q]; }
    mat_mul_full(l, m);
#pragma unroll
    for (int q = 0; q < 16; q++) mats[(size_t)i * 16 + q] = m[q];
}

__global__ void k_identity(double *__restrict__ mats, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int q = 0; q < 16; q++) mats[(size_t)i * 16 + q] = (q % 5 == 0) ? 1.0 : 0.0;
}

// henon.cu:316-338 (in place, out == vecs) and matrix_4d_vector::get_vector
// (henon.cu:1758-1787: out = M rv, a sum of four products per row).
__global__ void k_matvec(const double *__restrict__ mats, const double *__restrict__ vecs, double *__restrict__ out,
                         uint32_t n, int reference_order)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *m = mats + (size_t)i * 16;
    const double v0 = vecs[(size_t)i * 4], v1 = vecs[(size_t)i * 4 + 1], v2 = vecs[(size_t)i * 4 + 2],
                 v3 = vecs[(size_t)i * 4 + 3];
    double r[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (reference_order) {  // tmp = 0; tmp += m*v ... (henon.cu:328-332)
            double t = 0.0;
            t = t + m[q * 4] * v0; t = t + m[q * 4 + 1] * v1; t = t + m[q * 4 + 2] * v2; t = t + m[q * 4 + 3] * v3;
            r[q] = t;
        } else {  // m0 v0 + m1 v1 + m2 v2 + m3 v3 (henon.cu:1777-1780)
            r[q] = (((m[q * 4] * v0) + (m[q * 4 + 1] * v1)) + (m[q * 4 + 2] * v2)) + (m[q * 4 + 3] * v3);
        }
    }
#pragma unroll
    for (int q = 0; q < 4; q++) out[(size_t)i * 4 + q] = r[q];
}

// henon.cu:341-359
__global__ void k_normalize(double *__restrict__ v, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double norm = 0.0;
#pragma unroll
    for (int q = 0; q < 4; q++) norm = norm + v[(size_t)i * 4 + q] * v[(size_t)i * 4 + q];
    norm = sqrt(norm);
#pragma unroll
    for (int q = 0; q < 4; q++) v[(size_t)i * 4 + q] = v[(size_t)i * 4 + q] / norm;
}

// henon.cu:2083-2098
__global__ void k_lyap_birk(double *__restrict__ values, double *__restrict__ values_b, const double *__restrict__ v,
                            const double *__restrict__ weights, uint32_t n, uint64_t n_weights, uint64_t b_idx)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double norm = 0.0;
#pragma unroll
    for (int q = 0; q < 4; q++) norm = norm + v[(size_t)i * 4 + q] * v[(size_t)i * 4 + q];
    norm = sqrt(norm);
    values[i] = values[i] + log(norm) / (double)n_weights;
    values_b[i] = values_b[i] + log(norm) * weights[b_idx];
}

__device__ __forceinline__ double frob_ref(const double *m)
{
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < 16; q++) t = t + m[q] * m[q];
    return sqrt(t);
}

// henon.cu:2253-2284 (layer3 may be NULL) and :2366-2402 (layer2 per sample)
__global__ void k_megno_add(double *__restrict__ l1, double *__restrict__ l2, double *__restrict__ l3,
                            const double *__restrict__ ma, const double *__restrict__ mb, uint32_t n, uint64_t idx,
                            const double *__restrict__ w_flat, const uint64_t *__restrict__ w_off,
                            const uint64_t *__restrict__ turn_samples, uint32_t n_samples)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double ta = frob_ref(ma + (size_t)i * 16), tb = frob_ref(mb + (size_t)i * 16);
    l1[i] = l1[i] + log10(ta / tb) * (double)idx;
    if (!w_flat) {
        l2[i] = l2[i] + (1.0 / (double)idx) * l1[i];
        l3[i] = (1.0 / (double)idx) * l2[i];
    } else {
        for (uint32_t j = 0; j < n_samples; j++)
            if (idx <= turn_samples[j])
                l2[(size_t)j * n + i] = l2[(size_t)j * n + i] + l1[i] * w_flat[w_off[j] + idx - 1] / (double)idx;
    }
}

// henon.cu:2533-2584; init=1 stores the angles only
__global__ void k_tune_birkhoff_add(const double *__restrict__ x, const double *__restrict__ px,
                                    const double *__restrict__ y, const double *__restrict__ py,
                                    double *__restrict__ store_x, double *__restrict__ store_y,
                                    double *__restrict__ t1x, double *__restrict__ t2x, double *__restrict__ t1y,
                                    double *__restrict__ t2y, const double *__restrict__ w_flat,
                                    const uint64_t *__restrict__ w_off, const uint64_t *__restrict__ turn_samples,
                                    uint32_t n_samples, uint32_t n, uint64_t idx, int init)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double ax = atan2(px[i], x[i]), ay = atan2(py[i], y[i]);
    if (init) {
        store_x[i] = ax;
        store_y[i] = ay;
        return;
    }
    double dx = ax - store_x[i], dy = ay - store_y[i];
    dx = dx > 0 ? dx : dx + 2 * M_PI;
    dy = dy > 0 ? dy : dy + 2 * M_PI;
    store_x[i] = ax;
    store_y[i] = ay;
    for (uint32_t j = 0; j < n_samples; j++) {
        const size_t o = (size_t)j * n + i;
        const uint64_t T = turn_samples[j];
        if (idx < T) {
            t1x[o] = t1x[o] + dx * w_flat[w_off[j] + idx];
            t1y[o] = t1y[o] + dy * w_flat[w_off[j] + idx];
        } else if (idx < 2 * T) {
            t2x[o] = t2x[o] + dx * w_flat[w_off[j] + idx - T];
            t2y[o] = t2y[o] + dy * w_flat[w_off[j] + idx - T];
        }
    }
}

}  // namespace hm
using namespace hm;

namespace {
inline uint32_t grid_of(size_t n, uint32_t block = 256) { return (uint32_t)((n + block - 1) / block); }

void host_birkhoff_weights(uint32_t n, double *out)
{
    // src/dynamic_indicator.cpp:63-79
    for (uint32_t i = 0; i < n; i++) {
        double t = (double)i / (double)(n);
        out[i] = std::exp(-1 / (t * (1.0 - t)));
    }
    double sum = 0.0;
    for (uint32_t i = 0; i < n; i++) sum = sum + out[i];
    for (uint32_t i = 0; i < n; i++) out[i] = out[i] / sum;
}
}  // namespace

extern "C" int hm_megno_trace(double *x, double *px, double *y, double *py, uint32_t *steps, size_t n,
                              const double *table, size_t table_len, uint32_t n_turns, double mu, double barrier,
                              int inverse, const uint32_t *turn_samples, size_t n_samples, const double *matrix_in,
                              double *matrix_out, double *out, void *stream)
{
    return guarded([&] {
        if (n_samples > (size_t)kMaxSamples) throw Error(HM_ERR_ARG, "megno: at most 16 turn samples");
        for (size_t q = 0; q < n_samples; q++)
            if (turn_samples[q] > n_turns) throw Error(HM_ERR_ARG, "megno: turn sample beyond n_turns");
        if (n_turns > table_len) throw Error(HM_ERR_RANGE, "megno: the table is shorter than n_turns");
        if (n == 0) return;
        if (!x || !px || !y || !py || !steps || !table || !out) throw Error(HM_ERR_ARG, "megno: NULL pointer");
        MegnoTraceArgs a{};
        a.x = x; a.px = px; a.y = y; a.py = py; a.steps = steps; a.n = (uint32_t)n;
        a.tab = reinterpret_cast<const TabEntry *>(table); a.tab_len = (uint32_t)table_len; a.n_turns = n_turns;
        a.mu = mu; a.b2 = barrier * barrier;
        for (size_t q = 0; q < n_samples; q++) a.samples[q] = turn_samples[q];
        a.n_samples = (uint32_t)n_samples; a.out = out; a.status = status_word();
        a.m_in = matrix_in; a.m_out = matrix_out;
        hipStream_t s = as_stream(stream);
        LaunchTimer timer("k_megno_trace", s);
        const dim3 g(grid_of(n)), b(256);
        if (inverse) {
            if (mu != 0.0) hipLaunchKernelGGL((k_megno_trace<true, true>), g, b, 0, s, a);
            else hipLaunchKernelGGL((k_megno_trace<false, true>), g, b, 0, s, a);
        } else {
            if (mu != 0.0) hipLaunchKernelGGL((k_megno_trace<true, false>), g, b, 0, s, a);
            else hipLaunchKernelGGL((k_megno_trace<false, false>), g, b, 0, s, a);
        }
        launch_check("k_megno_trace");
    });
}

// 1.0 / k for k = 1..n (correctly rounded, computed on the host), kept per
// device in workspace slot 7 and extended when a longer run needs it.
const double *hm::reciprocal_table(uint32_t n, hipStream_t s)
{
    static std::mutex mu;
    static std::map<int, uint32_t> filled;
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lk(mu);
    uint32_t &have = filled[dev];
    const uint32_t want = std::max<uint32_t>(n, 1u << 16);
    if (have >= n) return static_cast<const double *>(workspace((size_t)have * sizeof(double), 7));
    double *d = static_cast<double *>(workspace((size_t)want * sizeof(double), 7));
    std::vector<double> h(want);
    for (uint32_t k = 0; k < want; ++k) h[k] = 1.0 / (double)(k + 1);
    hip_check(hipMemcpyAsync(d, h.data(), want * sizeof(double), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    have = want;
    return d;
}

// Loss compaction of the indicator kernel: the first segment's length and the
// longest segment.
static constexpr uint32_t kIndFirstSegment = 125;
static constexpr uint32_t kIndCompactEvery = 1000;

// Pinned host words receiving the per-segment live counts of k_indicators.
static uint32_t *indicator_counts(size_t n)
{
    static std::mutex mu;
    static uint32_t *buf = nullptr;
    static size_t cap = 0;
    std::lock_guard<std::mutex> lk(mu);
    if (cap < n) {
        if (buf) hip_check(hipHostFree(buf), "hipHostFree");
        buf = nullptr;
        cap = 0;
        hip_check(hipHostMalloc((void **)&buf, n * sizeof(uint32_t), hipHostMallocDefault), "hipHostMalloc");
        cap = n;
    }
    return buf;
}

extern "C" int hm_track_indicators(double *x, double *px, double *y, double *py, uint32_t *steps, size_t n,
                                   const double *table, size_t table_len, uint32_t n_turns, double mu,
                                   double barrier, int mode, const hm_indicator_spec *spec, void *stream)
{
    return guarded([&] {
        if (!spec) throw Error(HM_ERR_ARG, "indicators: spec is NULL");
        if (mode != HM_MODE_EXACT && mode != HM_MODE_FAST && mode != HM_MODE_EXACT_ORBIT)
            throw Error(HM_ERR_ARG, "indicators: unknown mode");
        if (spec->n_samples > (uint32_t)kMaxSamples) throw Error(HM_ERR_ARG, "indicators: at most 16 samples");
        if (spec->n_birkhoff > (uint32_t)kMaxBirk) throw Error(HM_ERR_ARG, "indicators: at most 4 Birkhoff windows");
        for (uint32_t q = 0; q < spec->n_samples; q++) {
            if (spec->samples[q] < 1 || spec->samples[q] > n_turns)
                throw Error(HM_ERR_ARG, "indicators: samples must lie in 1..n_turns");
            if (q && spec->samples[q] <= spec->samples[q - 1])
                throw Error(HM_ERR_ARG, "indicators: samples must increase");
        }
        if (spec->n_samples && !spec->megno && !spec->fli)
            throw Error(HM_ERR_ARG, "indicators: samples given without a megno or fli output");
        for (uint32_t w = 0; w < spec->n_birkhoff; w++)
            if (spec->birk_from[w] >= spec->birk_to[w] || spec->birk_to[w] > n_turns)
                throw Error(HM_ERR_ARG, "indicators: Birkhoff window needs from < to <= n_turns");
        if (spec->n_birkhoff && !spec->birk_out) throw Error(HM_ERR_ARG, "indicators: birk_out is NULL");
        if (spec->hist_len && (!spec->hist || (uint64_t)spec->hist_from + spec->hist_len > (uint64_t)n_turns + 1))
            throw Error(HM_ERR_ARG, "indicators: trajectory window must lie in samples 0..n_turns");
        if (spec->n_megno_birkhoff > (uint32_t)kMaxBirk || spec->n_tune_birkhoff > (uint32_t)kMaxBirk)
            throw Error(HM_ERR_ARG, "indicators: at most 4 megno_birkhoff and 4 tune_birkhoff turn samples");
        if (spec->n_megno_birkhoff && !spec->megno_birkhoff_out)
            throw Error(HM_ERR_ARG, "indicators: megno_birkhoff_out is NULL");
        if (spec->n_tune_birkhoff && !spec->tune_birkhoff_out)
            throw Error(HM_ERR_ARG, "indicators: tune_birkhoff_out is NULL");
        for (uint32_t w = 0; w < spec->n_megno_birkhoff; w++)
            if (spec->megno_birkhoff_samples[w] == 0) throw Error(HM_ERR_ARG, "indicators: turn samples must be >= 1");
        for (uint32_t w = 0; w < spec->n_tune_birkhoff; w++)
            if (spec->tune_birkhoff_samples[w] == 0) throw Error(HM_ERR_ARG, "indicators: turn samples must be >= 1");
        if (n == 0 || n_turns == 0) return;
        if (!x || !px || !y || !py || !steps || !table) throw Error(HM_ERR_ARG, "indicators: NULL pointer");
        hipStream_t s = as_stream(stream);
        IndicatorArgs a{};
        a.x = x; a.px = px; a.y = y; a.py = py; a.steps = steps; a.n = (uint32_t)n;
        a.tab = reinterpret_cast<const TabEntry *>(table); a.tab_len = (uint32_t)table_len; a.n_turns = n_turns;
        a.mu = mu; a.b2 = barrier * barrier; a.renorm_every = spec->renorm_every;
        for (uint32_t q = 0; q < spec->n_samples; q++) a.samples[q] = spec->samples[q];
        a.n_samples = spec->n_samples;
        for (int q = 0; q < 4; q++) a.fli_dir[q] = spec->fli_dir[q];
        if (a.fli_dir[0] == 0 && a.fli_dir[1] == 0 && a.fli_dir[2] == 0 && a.fli_dir[3] == 0) a.fli_dir[0] = 1.0;
        a.megno = spec->megno; a.fli = spec->fli; a.matrix = spec->matrix; a.log_scale = spec->log_scale;
        a.n_birk = spec->n_birkhoff; a.birk_out = spec->birk_out;
        a.n_mb = spec->n_megno_birkhoff; a.mb_out = spec->megno_birkhoff_out;
        a.n_tb = spec->n_tune_birkhoff; a.tb_out = spec->tune_birkhoff_out;
        if (a.n_birk || a.n_mb || a.n_tb) {
            // one table of weights: the history windows', then birkhoff_weights(T_j)
            // of every construct-form turn sample
            size_t total = 0;
            for (uint32_t w = 0; w < a.n_birk; w++) {
                a.birk_from[w] = spec->birk_from[w];
                a.birk_to[w] = spec->birk_to[w];
                a.birk_off[w] = (uint32_t)total;
                total += spec->birk_to[w] - spec->birk_from[w];
            }
            for (uint32_t w = 0; w < a.n_mb; w++) {
                a.mb_T[w] = spec->megno_birkhoff_samples[w];
                a.mb_off[w] = (uint32_t)total;
                total += a.mb_T[w];
            }
            for (uint32_t w = 0; w < a.n_tb; w++) {
                a.tb_T[w] = spec->tune_birkhoff_samples[w];
                a.tb_off[w] = (uint32_t)total;
                total += a.tb_T[w];
            }
            if (total > 0xffffffffull) throw Error(HM_ERR_ARG, "indicators: weight tables exceed 2^32 entries");
            std::vector<double> w(std::max<size_t>(total, 1));
            for (uint32_t q = 0; q < a.n_birk; q++)
                host_birkhoff_weights(a.birk_to[q] - a.birk_from[q], w.data() + a.birk_off[q]);
            for (uint32_t q = 0; q < a.n_mb; q++) host_birkhoff_weights(a.mb_T[q], w.data() + a.mb_off[q]);
            for (uint32_t q = 0; q < a.n_tb; q++) host_birkhoff_weights(a.tb_T[q], w.data() + a.tb_off[q]);
            double *dw = static_cast<double *>(workspace(w.size() * sizeof(double), 1));
            hip_check(hipMemcpyAsync(dw, w.data(), total * sizeof(double), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
            hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
            a.birk_w = dw;
        }
        a.hist_from = spec->hist_from; a.hist_len = spec->hist_len; a.hist = spec->hist;
        a.inv_idx = reciprocal_table(n_turns, s);
        a.status = status_word();
        const bool oct = mu != 0.0;
        LaunchTimer timer("k_indicators", s);
        // Birkhoff windows, and the trajectory and construct windows, are
        // compiled into their own variants, so the MEGNO/FLI-only kernel
        // carries none of their code
        const bool extra = a.hist_len || a.n_mb || a.n_tb;
        const int V = extra ? 2 : (a.n_birk ? 1 : 0);
        using K = void (*)(IndicatorArgs);
        auto pick = [&](auto mode_tag) -> K {
            constexpr int M = decltype(mode_tag)::value;
            if (oct) return V == 2 ? k_indicators<M, true, 2> : V == 1 ? k_indicators<M, true, 1> : k_indicators<M, true, 0>;
            return V == 2 ? k_indicators<M, false, 2> : V == 1 ? k_indicators<M, false, 1> : k_indicators<M, false, 0>;
        };
        K k = mode == HM_MODE_EXACT       ? pick(std::integral_constant<int, kExact>{})
              : mode == HM_MODE_EXACT_ORBIT ? pick(std::integral_constant<int, kExactOrbit>{})
                                            : pick(std::integral_constant<int, kFast>{});
        a.first = a.last = 1;
        a.k0 = 0;
        a.k1 = n_turns;
        if (n_turns <= kIndFirstSegment || n > 0x7fffffffull) {
            hipLaunchKernelGGL(k, dim3(grid_of(n)), dim3(256), 0, s, a);
            launch_check("k_indicators");
            return;
        }
        // Segments of kIndFirstSegment turns doubling up to kIndCompactEvery,
        // each over the particles still alive: the survivors' matrices and sums
        // wait in `state` between launches, and a launch is sized from the live
        // count of the segment two back (read by the host while the GPU runs),
        // so lost particles stop occupying lanes.
        std::vector<uint32_t> seg;
        for (uint32_t done = 0, len = kIndFirstSegment; done < n_turns;) {
            const uint32_t t = std::min(len, n_turns - done);
            seg.push_back(t);
            done += t;
            len = std::min<uint32_t>(len * 2u, kIndCompactEvery);
        }
        const size_t S = seg.size(), N = n;
        double *state = static_cast<double *>(workspace(N * ind_state_len(V) * sizeof(double), 10));
        uint32_t *lists = static_cast<uint32_t *>(workspace(sizeof(uint32_t) * (2 * N + S + 64), 11));
        uint32_t *counts = lists + 2 * N;  // counts[i]: live entries entering segment i (i >= 1)
        uint32_t *host = indicator_counts(S + 1);
        hip_check(hipMemsetAsync(counts, 0, sizeof(uint32_t) * (S + 1), s), "hipMemsetAsync");
        std::vector<hipEvent_t> ev(S, nullptr);
        auto cleanup = [&] {
            for (auto e : ev)
                if (e) (void)hipEventDestroy(e);
        };
        try {
            uint64_t bound = N;
            uint32_t start = 0;
            for (size_t q = 0; q < S; ++q) {
                if (q >= 2) {
                    hip_check(hipEventSynchronize(ev[q - 2]), "hipEventSynchronize");
                    bound = std::min<uint64_t>(bound, host[q - 1]);
                }
                const bool last = q + 1 == S;
                a.k0 = start;
                a.k1 = start + seg[q];
                a.first = q == 0;
                a.last = last;
                a.idx_in = q == 0 ? nullptr : (q % 2 ? lists : lists + N);
                a.count_in = q == 0 ? nullptr : counts + q;
                a.idx_out = last ? nullptr : (q % 2 ? lists + N : lists);
                a.count_out = last ? nullptr : counts + q + 1;
                a.state = state;
                if (bound > 0) {
                    hipLaunchKernelGGL(k, dim3(grid_of(bound)), dim3(256), 0, s, a);
                    launch_check("k_indicators");
                }
                if (!last) {
                    hip_check(hipMemcpyAsync(host + q + 1, counts + q + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s),
                              "hipMemcpyAsync");
                    hip_check(hipEventCreateWithFlags(&ev[q], hipEventDisableTiming), "hipEventCreate");
                    hip_check(hipEventRecord(ev[q], s), "hipEventRecord");
                }
                start += seg[q];
            }
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
    });
}

extern "C" int hm_tangent_apply(double *mats, const double *x, const double *px, const double *y, const double *py,
                                size_t n, const double *table, size_t table_len, uint32_t table_index, double mu,
                                int multiply, void *stream)
{
    return guarded([&] {
        if (table_index >= table_len) throw Error(HM_ERR_RANGE, "tangent: table index out of range");
        if (n == 0) return;
        hipStream_t s = as_stream(stream);
        LaunchTimer timer("k_tangent_apply", s);
        hipLaunchKernelGGL(k_tangent_apply, dim3(grid_of(n)), dim3(256), 0, s, mats, x, px, y, py, (uint32_t)n,
                           reinterpret_cast<const TabEntry *>(table), table_index, mu, multiply);
        launch_check("k_tangent_apply");
    });
}

extern "C" int hm_matrix_left_multiply(double *mats, const double *left, size_t n, void *stream)
{
    return guarded([&] {
        if (n == 0) return;
        hipStream_t s = as_stream(stream);
        hipLaunchKernelGGL(k_left_multiply, dim3(grid_of(n)), dim3(256), 0, s, mats, left, (uint32_t)n);
        launch_check("k_left_multiply");
    });
}

extern "C" int hm_matrix_identity(double *mats, size_t n, void *stream)
{
    return guarded([&] {
        if (n == 0) return;
        hipStream_t s = as_stream(stream);
        hipLaunchKernelGGL(k_identity, dim3(grid_of(n)), dim3(256), 0, s, mats, (uint32_t)n);
        launch_check("k_identity");
    });
}

extern "C" int hm_matvec(const double *mats, const double *vecs, double *out, size_t n, int reference_order,
                         void *stream)
{
    return guarded([&] {
        if (n == 0) return;
        hipStream_t s = as_stream(stream);
        hipLaunchKernelGGL(k_matvec, dim3(grid_of(n)), dim3(256), 0, s, mats, vecs, out, (uint32_t)n, reference_order);
        launch_check("k_matvec");
    });
}

extern "C" int hm_normalize(double *vecs, size_t n, void *stream)
{
    return guarded([&] {
        if (n == 0) return;
        hipStream_t s = as_stream(stream);
        hipLaunchKernelGGL(k_normalize, dim3(grid_of(n)), dim3(256), 0, s, vecs, (uint32_t)n);
        launch_check("k_normalize");
    });
}

extern "C" int hm_lyapunov_add(double *values, double *values_b, const double *vecs, const double *weights, size_t n,
                               uint64_t n_weights, uint64_t idx, void *stream)
{
    return guarded([&] {
        if (idx >= n_weights) throw Error(HM_ERR_RANGE, "lyapunov: index beyond the weights");
        if (n == 0) return;
        hipStream_t s = as_stream(stream);
        hipLaunchKernelGGL(k_lyap_birk, dim3(grid_of(n)), dim3(256), 0, s, values, values_b, vecs, weights,
                           (uint32_t)n, n_weights, idx);
        launch_check("k_lyap_birk");
    });
}

extern "C" int hm_megno_add(double *l1, double *l2, double *l3, const double *ma, const double *mb, size_t n,
                            uint64_t idx, void *stream)
{
    return guarded([&] {
        if (n == 0) return;
        hipStream_t s = as_stream(stream);
        hipLaunchKernelGGL(k_megno_add, dim3(grid_of(n)), dim3(256), 0, s, l1, l2, l3, ma, mb, (uint32_t)n, idx,
                           (const double *)nullptr, (const uint64_t *)nullptr, (const uint64_t *)nullptr, 0u);
        launch_check("k_megno_add");
    });
}

extern "C" int hm_megno_birkhoff_add(double *l1, double *l2, const double *ma, const double *mb, size_t n,
                                     uint64_t idx, const double *w_flat, const uint64_t *w_off,
                                     const uint64_t *turn_samples, uint32_t n_samples, void *stream)
{
    return guarded([&] {
        if (n == 0) return;
        if (!w_flat || !w_off || !turn_samples) throw Error(HM_ERR_ARG, "megno_birkhoff: NULL tables");
        hipStream_t s = as_stream(stream);
        hipLaunchKernelGGL(k_megno_add, dim3(grid_of(n)), dim3(256), 0, s, l1, l2, (double *)nullptr, ma, mb,
                           (uint32_t)n, idx, w_flat, w_off, turn_samples, n_samples);
        launch_check("k_megno_add");
    });
}

extern "C" int hm_tune_birkhoff_add(const double *x, const double *px, const double *y, const double *py,
                                    double *store_x, double *store_y, double *t1x, double *t2x, double *t1y,
                                    double *t2y, const double *w_flat, const uint64_t *w_off,
                                    const uint64_t *turn_samples, uint32_t n_samples, size_t n, uint64_t idx,
                                    int init, void *stream)
{
    return guarded([&] {
        if (n == 0) return;
        hipStream_t s = as_stream(stream);
        hipLaunchKernelGGL(k_tune_birkhoff_add, dim3(grid_of(n)), dim3(256), 0, s, x, px, y, py, store_x, store_y,
                           t1x, t2x, t1y, t2y, w_flat, w_off, turn_samples, n_samples, (uint32_t)n, idx, init);
        launch_check("k_tune_birkhoff_add");
    });
}
