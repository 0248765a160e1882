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

template <int MODE, bool OCT>
__global__ void __launch_bounds__(256) k_adjoint_forward(AdjointFwdArgs a)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const size_t n = a.n;
    double x = a.x0[i], px = a.px0[i], y = a.y0[i], py = a.py0[i];
    bool alive = !any_nan(x, px, y, py);
    double bsum = 0.0;
    const double m3mu = -3.0 * a.mu;
    for (uint32_t k = 0; k < a.n_turns; ++k) {
        if (k % a.every == 0) {
            const size_t o = (size_t)(k / a.every) * 4 * n;
            a.ckpt[o + i] = x; a.ckpt[o + n + i] = px; a.ckpt[o + 2 * n + i] = y; a.ckpt[o + 3 * n + i] = py;
        }
        if (alive) {
            map_step<MODE, false, OCT>(x, px, y, py, a.mu, m3mu, a.tab[k]);
            if (!(radius2<MODE>(x, px, y, py) < a.b2)) {
                alive = false;
                x = px = y = py = __builtin_nan("");
            }
        }
        bsum = fma(a.weights[k], x * x, bsum);
    }
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
    double *sub;              // scratch [K/kChunk + 1][4][n]
    double *grad_x0;          // [4][n]
    double *grad_p;           // [4][n]: d/d eps, d/d mu, d/d omega_x0, d/d omega_y0
};

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
    double lx = a.seed_s[i], lpx = a.seed_s[n + i], ly = a.seed_s[2 * n + i], lpy = a.seed_s[3 * n + i];
    const double sb = a.seed_b ? a.seed_b[i] : 0.0;
    double g_eps = 0.0, g_mu = 0.0, g_wx = 0.0, g_wy = 0.0;
    const double mu = a.mu, m3mu = -3.0 * mu, mu3 = 3.0 * mu, mu6 = 6.0 * mu;
    const uint32_t T = a.n_turns, K = a.every;
    const uint32_t n_seg = (T + K - 1) / K;
    for (int seg = (int)n_seg - 1; seg >= 0; --seg) {
        const uint32_t s0 = (uint32_t)seg * K, s1 = min(s0 + K, T);
        // level 1: recompute the segment, sub-checkpoint every kChunk turns
        {
            const size_t o = (size_t)seg * 4 * n;
            double x = a.ckpt[o + i], px = a.ckpt[o + n + i], y = a.ckpt[o + 2 * n + i], py = a.ckpt[o + 3 * n + i];
            for (uint32_t k = s0; k < s1; ++k) {
                if ((k - s0) % kChunk == 0) {
                    const size_t q = (size_t)((k - s0) / kChunk) * 4 * n;
                    a.sub[q + i] = x; a.sub[q + n + i] = px; a.sub[q + 2 * n + i] = y; a.sub[q + 3 * n + i] = py;
                }
                map_step<MODE, false, OCT>(x, px, y, py, mu, m3mu, a.tab[k]);
            }
        }
        const uint32_t n_chunk = (s1 - s0 + kChunk - 1) / kChunk;
        for (int ch = (int)n_chunk - 1; ch >= 0; --ch) {
            const uint32_t c0 = s0 + (uint32_t)ch * kChunk;
            const uint32_t len = min((uint32_t)kChunk, s1 - c0);
            // level 2: the chunk's states S[0..len] in registers
            double S[kChunk + 1][4];
            const size_t q = (size_t)ch * 4 * n;
            S[0][0] = a.sub[q + i]; S[0][1] = a.sub[q + n + i]; S[0][2] = a.sub[q + 2 * n + i];
            S[0][3] = a.sub[q + 3 * n + i];
#pragma unroll
            for (int m = 0; m < kChunk; ++m) {
                S[m + 1][0] = S[m][0]; S[m + 1][1] = S[m][1]; S[m + 1][2] = S[m][2]; S[m + 1][3] = S[m][3];
                if ((uint32_t)m < len)
                    map_step<MODE, false, OCT>(S[m + 1][0], S[m + 1][1], S[m + 1][2], S[m + 1][3], mu, m3mu,
                                               a.tab[c0 + m]);
            }
#pragma unroll
            for (int m = kChunk - 1; m >= 0; --m) {
                if ((uint32_t)m >= len) continue;
                const uint32_t k = c0 + m;
                const TabEntry t = a.tab[k];
                const double x = S[m][0], y = S[m][2];
                const double xn = S[m + 1][0], pxn = S[m + 1][1], yn = S[m + 1][2], pyn = S[m + 1][3];
                // Birkhoff seed of sample k+1: d/dx_{k+1} of w_k x_{k+1}^2
                lx = fma(2.0 * sb * a.weights[k], xn, lx);
                // rotation parameters (post-step point)
                const double dwx = fma(lx, pxn, -lpx * xn), dwy = fma(ly, pyn, -lpy * yn);
                const double4 dt = *reinterpret_cast<const double4 *>(a.dtab + 4 * (size_t)k);
                g_wx = fma(dwx, dt.x, g_wx);
                g_wy = fma(dwy, dt.z, g_wy);
                g_eps = fma(dwx, dt.y, fma(dwy, dt.w, g_eps));
                // through the rotations
                const double lxt = fma(t.cx, lx, -t.sx * lpx), lpt = fma(t.sx, lx, t.cx * lpx);
                const double lyt = fma(t.cy, ly, -t.sy * lpy), lqt = fma(t.sy, ly, t.cy * lpy);
                // through the kicks (pre-step point)
                const double xx = x * x, yy = y * y;
                g_mu = fma(lpt, x * fma(-3.0, yy, xx), fma(lqt, y * fma(-3.0, xx, yy), g_mu));
                const double ja = fma(mu3, xx - yy, 2.0 * x);
                const double jb = fma(-mu6 * x, y, -2.0 * y);
                lx = fma(ja, lpt, fma(jb, lqt, lxt));
                ly = fma(jb, lpt, fma(-ja, lqt, lyt));
                lpx = lpt;
                lpy = lqt;
            }
        }
    }
    a.grad_x0[i] = lx; a.grad_x0[n + i] = lpx; a.grad_x0[2 * n + i] = ly; a.grad_x0[3 * n + i] = lpy;
    a.grad_p[i] = g_eps; a.grad_p[n + i] = g_mu; a.grad_p[2 * n + i] = g_wx; a.grad_p[3 * n + i] = g_wy;
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
        *bytes = ((size_t)(every + kChunk - 1) / kChunk + 1) * 4 * n * sizeof(double);
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
