// Reverse-mode (adjoint) differentiation of n-turn tracking, with checkpoints.
// No reference counterpart: the reference has no derivatives beyond the
// one-turn tangent matrix (src/henon.cu:175-265).
//
// Forward sweep (k_adjoint_forward): track from the initial points with table
// index k, write the state every K turns (checkpoints), the final state, the
// Birkhoff-weighted sum B = sum_{k=1..T} w_{k-1} x_k^2 (w = birkhoff_weights(T))
// and whether the particle survived all T turns.
//
// Backward sweep (k_adjoint_segments + k_adjoint_sweep): the K-turn segments
// are recomputed from their checkpoints, writing sub-checkpoints every kChunk
// turns to scratch; then per particle, per chunk (in reverse), the chunk's
// kChunk+1 states are recomputed into registers and the adjoint
// lambda <- J^T lambda (+ the Birkhoff seed) propagated, accumulating the
// parameter derivatives
//   d/dmu     : lambda_p~ . (x^3 - 3 x y^2, y^3 - 3 x^2 y)   (octupole kick)
//   d/domega_k: lambda_x' p' - lambda_p' x'                   (rotation, per plane)
// chained through d omega_k / d(omega_0, epsilon) (a per-turn table).
// Recomputation never inverts the map, which is unstable on chaotic orbits.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>

#include "hm_device.h"
#include "hm_internal.h"

namespace hm {

constexpr int kChunk = 8;

__device__ __forceinline__ uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }

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

// LDS ring of the aligned forward sweep, one per wavefront: three 8-turn
// chunks, per turn sx cx sy cy | weight, pad.
constexpr int kFwdTurn = 6;
constexpr int kFwdChunk = 8 * kFwdTurn;
constexpr int kFwdWave = 3 * kFwdChunk;
constexpr int kFwdBlock = 256;

// Every lane runs every turn (a lost particle's lane carries on with values
// that are discarded), so the table record and the weight of a turn are
// wave-uniform. A lost particle is parked at the origin, a fixed point of the
// map, so that its lane never again takes the loss test's rare path.
// With every % 8 == 0 (ALIGNED) the turns run in 8-turn chunks whose table
// records and weights the wavefront stages into its LDS ring two chunks ahead
// (one vector load and one LDS store per lane and chunk); a turn reads its
// entries from LDS one turn ahead, so neither scalar-cache misses nor LDS
// latency sit on the dependency chain. Otherwise the turns run four at a time
// with scalar loads of two alternating prefetched pairs of records.
template <int MODE, bool OCT, bool ALIGNED>
__global__ void __launch_bounds__(kFwdBlock) k_adjoint_forward(AdjointFwdArgs a)
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
    if (!alive) x = px = y = py = 0.0;  // NaN inputs and lanes past n: parked
    const ConstPtr<TabEntry> tab = as_const(a.tab);
    const ConstPtr<double> weights = as_const(a.weights);
    const double mu = a.mu, m3mu = -3.0 * mu, b2 = a.b2, b2s = screen_bound(b2);
    double xx = x * x, yy = y * y;
    double bsum = 0.0;
    // The common turn is the map, the screened radius and one wave-wide test;
    // the exact decision and the parking of a lost point (at the origin, a
    // fixed point of the map, so it never reaches the screen again) run only
    // when some lane of the wavefront is at or past the screen.
    auto turn = [&](const TabEntry &t, double w) {
        bool cand;
        if (MODE == kExact) {
            cand = !(fwd_exact_carry<OCT>(x, px, y, py, xx, yy, mu, t) < b2s);
        } else {
            map_step<kFast, false, OCT>(x, px, y, py, mu, m3mu, t);
            cand = !(radius2<kFast>(x, px, y, py) < b2);
        }
        if (__builtin_expect(__any(cand), 0)) {
            const bool lost = cand && (MODE != kExact || !(exact_radius2(xx, px, yy, py) < b2));
            if (lost) {
                alive = false;
                x = px = y = py = xx = yy = 0.0;
            }
        }
        bsum = fma(w, MODE == kExact ? xx : x * x, bsum);
    };
    auto checkpoint = [&](double *ck) {
        if (in) {
            ck[i] = x; ck[n + i] = px; ck[2 * n + i] = y; ck[3 * n + i] = py;
        }
    };
    const uint32_t T = a.n_turns, K = a.every;
    if constexpr (ALIGNED) {
        __shared__ __align__(16) double ring_all[(kFwdBlock / 64) * kFwdWave];
        double *ring = ring_all + (threadIdx.x >> 6) * kFwdWave;
        const uint32_t lane = threadIdx.x & 63;
        const int G = (int)((T + 7) / 8);
        const double *tabd = reinterpret_cast<const double *>(a.tab);
        // chunk g: lane l < 32 carries table double l (turn l/4), the others the
        // weight of turn l % 8 (identical writes); turns past T read turn T-1
        auto stage_load = [&](int g) -> double {
            const uint32_t t0 = (uint32_t)g * 8;
            return lane < 32 ? tabd[4 * (size_t)umin(t0 + (lane >> 2), T - 1) + (lane & 3)]
                             : a.weights[umin(t0 + (lane & 7), T - 1)];
        };
        auto stage_store = [&](int g, double e) {
            double *slot = ring + ((uint32_t)g % 3) * kFwdChunk;
            slot[lane < 32 ? (lane >> 2) * kFwdTurn + (lane & 3) : (lane & 7) * kFwdTurn + 4] = e;
        };
        auto ring_at = [&](int g, int m, TabEntry &t, double &w) {
            const double *q = ring + ((uint32_t)g % 3) * kFwdChunk + m * kFwdTurn;
            t = TabEntry{q[0], q[1], q[2], q[3]};
            w = q[4];
        };
        auto clampg = [&](int g) { return g < G ? g : G - 1; };
        stage_store(0, stage_load(0));
        stage_store(clampg(1), stage_load(clampg(1)));
        double e = stage_load(clampg(2));
        __builtin_amdgcn_wave_barrier();
        const uint32_t K8 = K / 8;
        double *ck = a.ckpt;
        uint32_t to_ck = 0;  // chunks until the next checkpoint
        TabEntry r;
        double w;
        ring_at(0, 0, r, w);
        for (int g = 0; g < G; ++g) {
            if (to_ck == 0) {
                checkpoint(ck);
                ck += 4 * n;
                to_ck = K8;
            }
            --to_ck;
            stage_store(clampg(g + 2), e);
            e = stage_load(clampg(g + 3));
            __builtin_amdgcn_wave_barrier();  // the stores precede later chunks' reads
            const uint32_t len = umin(8, T - (uint32_t)g * 8);
            if (len == 8) {
#pragma unroll
                for (int m = 0; m < 8; ++m) {
                    TabEntry rn;
                    double wn;
                    if (m + 1 < 8) ring_at(g, m + 1, rn, wn);
                    else ring_at(g + 1, 0, rn, wn);  // the next chunk's (its slot is staged)
                    turn(r, w);
                    r = rn;
                    w = wn;
                }
            } else {
                for (uint32_t m = 0; m < len; ++m) {
                    ring_at(g, (int)m, r, w);
                    turn(r, w);
                }
            }
        }
    } else {
        double *ck = a.ckpt;
        for (uint32_t s0 = 0; s0 < T; s0 += K) {
            checkpoint(ck);
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
    }
    if (!in) return;
    if (!alive) x = px = y = py = bsum = __builtin_nan("");
    a.final_state[i] = x; a.final_state[n + i] = px; a.final_state[2 * n + i] = y; a.final_state[3 * n + i] = py;
    a.birk[i] = bsum;
    a.alive[i] = alive ? 1 : 0;
}

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

// What the adjoint of one turn reads from the tables: the rotation record, the
// angle derivatives and the Birkhoff weight.
struct AdjRec {
    TabEntry t;
    double4 dt;
    double w;
};

// The backward sweep runs over windows of consecutive K-turn segments, the
// last window first. Per window two kernels run:
//  - k_adjoint_segments recomputes every segment of the window from its
//    checkpoint, one thread per (particle, segment), and writes the state at the
//    start of each 8-turn chunk after the first (sub-checkpoints). The segments
//    are independent, so this part runs at full occupancy even for few particles.
//  - the sweep (k_adjoint_sweep_fused for every % 8 == 0, else k_adjoint_sweep),
//    one thread per particle, walks the window's chunks in reverse: the chunk's
//    nine states S0..S8 (S8 = the next chunk's S0) come from its sub-checkpoint
//    (the segment's checkpoint for a first chunk), and the adjoint runs back
//    through its turns. The states of the chunk before it are recomputed in the
//    same straight-line block: the two dependency chains are independent and
//    interleave, which matters at one wavefront per SIMD, where nothing else
//    hides FP64 latency.
// The adjoint and the parameter sums carry from window to window in the output
// arrays.
struct AdjointSegArgs {
    const double *ckpt;       // [n_seg][4][n]
    uint32_t n;
    const TabEntry *tab;
    uint32_t n_turns, every;
    double mu;
    const uint8_t *alive;     // [n]
    double *sub;              // [window segments][seg_chunks - 1][4][n]
    uint32_t seg_first, seg_chunks;
};

template <int MODE, bool OCT>
__global__ void __launch_bounds__(256) k_adjoint_segments(AdjointSegArgs a)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n || !a.alive[i]) return;
    const size_t n = a.n;
    const uint32_t w = blockIdx.y, seg = a.seg_first + w;
    const uint32_t s0 = seg * a.every, s1 = umin(s0 + a.every, a.n_turns);
    const uint32_t chunks = (s1 - s0 + kChunk - 1) / kChunk;
    const double mu = a.mu, m3mu = -3.0 * mu;
    const ConstPtr<TabEntry> tab = as_const(a.tab);
    const double *ck = a.ckpt + (size_t)seg * 4 * n;
    double v[4] = {ck[i], ck[n + i], ck[2 * n + i], ck[3 * n + i]};
    double *out = a.sub + (size_t)w * (a.seg_chunks - 1) * 4 * n;
    for (uint32_t j = 1; j < chunks; ++j) {
        const ConstPtr<TabEntry> p = tab + (s0 + (j - 1) * kChunk);
#pragma unroll
        for (int m = 0; m < kChunk; ++m) step4<MODE, OCT>(v, mu, m3mu, load_record(p + m));
        out[i] = v[0]; out[n + i] = v[1]; out[2 * n + i] = v[2]; out[3 * n + i] = v[3];
        out += 4 * n;
    }
}

struct AdjointSweepArgs {
    const double *ckpt;       // [n_seg][4][n]
    uint32_t n;
    const TabEntry *tab;
    const double *dtab;       // [T][4]: d wx/d wx0, d wx/d eps, d wy/d wy0, d wy/d eps
    uint32_t n_turns, every;
    double mu;
    const double *weights;    // [T]
    const double *seed_s;     // [4][n]: dL/d(final state)
    const double *seed_b;     // [n]:    dL/dB, or NULL
    const uint8_t *alive;     // [n]
    const double *sub;        // k_adjoint_segments' output for this window
    uint32_t seg_first, seg_end, seg_chunks;
    int last_window;          // the window holding turn T-1: start from the seeds
    double *grad_x0;          // [4][n]: the adjoint, carried between windows
    double *grad_p;           // [4][n]: d/d eps, d/d mu, d/d omega_x0, d/d omega_y0
};

// A chunk: turns c0 .. c0+len-1 (len <= kChunk) of segment seg, its j-th.
struct Chunk {
    uint32_t seg, j, c0, len;
};

// Reverse sweep for any checkpoint interval: one thread per particle walks the
// window's chunks in reverse with scalar table loads, recomputing the states of
// the chunk before the current one while it runs the adjoint over the current.
template <int MODE, bool OCT>
__global__ void __launch_bounds__(256) k_adjoint_sweep(AdjointSweepArgs a)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = a.n;
    if (i >= a.n) return;
    if (!a.alive[i]) {
        if (a.last_window)
            for (int c = 0; c < 4; c++) { a.grad_x0[c * n + i] = 0.0; a.grad_p[c * n + i] = 0.0; }
        return;
    }
    AdjAcc acc;
    if (a.last_window) {
        acc = AdjAcc{a.seed_s[i], a.seed_s[n + i], a.seed_s[2 * n + i], a.seed_s[3 * n + i], 0.0, 0.0, 0.0, 0.0};
    } else {
        acc = AdjAcc{a.grad_x0[i], a.grad_x0[n + i], a.grad_x0[2 * n + i], a.grad_x0[3 * n + i],
                     a.grad_p[i],  a.grad_p[n + i],  a.grad_p[2 * n + i],  a.grad_p[3 * n + i]};
    }
    const double bseed = 2.0 * (a.seed_b ? a.seed_b[i] : 0.0);
    const double mu = a.mu, m3mu = -3.0 * mu, mu3 = 3.0 * mu, mu6 = 6.0 * mu;
    const uint32_t T = a.n_turns, K = a.every;
    const uint32_t n_seg = (T + K - 1) / K;
    const ConstPtr<TabEntry> tab = as_const(a.tab);
    const ConstPtr<double> dtab = as_const(a.dtab), weights = as_const(a.weights);
    auto load4 = [&](const double *p, double (&v)[4]) {
        v[0] = p[i]; v[1] = p[n + i]; v[2] = p[2 * n + i]; v[3] = p[3 * n + i];
    };
    auto adjoint_at = [&](uint32_t k, const double (&pre)[4], const double (&post)[4]) {
        const ConstPtr<double> d = dtab + 4 * (size_t)k;
        const double4 dt{d[0], d[1], d[2], d[3]};
        adjoint_turn<OCT>(pre, post, load_record(tab + k), bseed * weights[k], dt, mu3, mu6, acc);
    };
    auto fill = [&](double (&S)[kChunk][4], uint32_t c0, uint32_t len) {
#pragma unroll
        for (int m = 1; m < kChunk; ++m) {
            copy4(S[m], S[m - 1]);
            if ((uint32_t)m < len) step4<MODE, OCT>(S[m], mu, m3mu, load_record(tab + c0 + m - 1));
        }
    };
    auto sweep = [&](const double (&S)[kChunk][4], uint32_t c0, uint32_t len, const double (&after)[4]) {
#pragma unroll
        for (int m = kChunk - 1; m >= 0; --m) {
            if ((uint32_t)m >= len) continue;
            if ((uint32_t)m + 1 == len) adjoint_at(c0 + m, S[m], after);
            else adjoint_at(c0 + m, S[m], S[m + 1 < kChunk ? m + 1 : m]);
        }
    };
    auto seg_chunks = [&](uint32_t seg) -> uint32_t {
        const uint32_t s0 = seg * K;
        return (umin(s0 + K, T) - s0 + kChunk - 1) / kChunk;
    };
    auto chunk = [&](uint32_t seg, uint32_t j) -> Chunk {
        const uint32_t s1 = umin(seg * K + K, T), c0 = seg * K + j * kChunk;
        return Chunk{seg, j, c0, umin(kChunk, s1 - c0)};
    };
    // the chunk before c (in turn order) inside this window
    auto pred = [&](const Chunk &c, Chunk &p) -> bool {
        if (c.j > 0) { p = chunk(c.seg, c.j - 1); return true; }
        if (c.seg > a.seg_first) { p = chunk(c.seg - 1, seg_chunks(c.seg - 1) - 1); return true; }
        return false;
    };
    auto start_of = [&](const Chunk &c) -> const double * {
        if (c.j == 0) return a.ckpt + (size_t)c.seg * 4 * n;
        return a.sub + ((size_t)(c.seg - a.seg_first) * (a.seg_chunks - 1) + c.j - 1) * 4 * n;
    };
    double SA[kChunk][4], SB[kChunk][4], nxt[4], P[4];
    Chunk cur = chunk(a.seg_end - 1, seg_chunks(a.seg_end - 1) - 1), prv{};
    load4(start_of(cur), SA[0]);
    fill(SA, cur.c0, cur.len);
    if (a.seg_end < n_seg) {
        load4(a.ckpt + (size_t)a.seg_end * 4 * n, nxt);  // the state after the window
    } else {
        double v[4];
#pragma unroll
        for (int m = 0; m < kChunk; ++m)
            if ((uint32_t)m + 1 == cur.len) copy4(v, SA[m]);
        step4<MODE, OCT>(v, mu, m3mu, load_record(tab + T - 1));
        copy4(nxt, v);
    }
    bool has = pred(cur, prv);
    load4(has ? start_of(prv) : a.ckpt, P);
    // One chunk: S holds cur's states, R receives prv's. Returns whether a
    // chunk is left.
    auto body = [&](double (&S)[kChunk][4], double (&R)[kChunk][4]) __attribute__((always_inline)) -> bool {
        Chunk pp{};
        const bool has2 = has && pred(prv, pp);
        double Q[4];
        load4(has2 ? start_of(pp) : a.ckpt, Q);  // consumed by the next chunk
        if (has) {
            copy4(R[0], P);
            fill(R, prv.c0, prv.len);
        }
        sweep(S, cur.c0, cur.len, nxt);
        if (!has) return false;
        copy4(nxt, S[0]);
        cur = prv;
        prv = pp;
        has = has2;
        copy4(P, Q);
        return true;
    };
    while (body(SA, SB) && body(SB, SA)) {
    }
    a.grad_x0[i] = acc.lx; a.grad_x0[n + i] = acc.lpx; a.grad_x0[2 * n + i] = acc.ly; a.grad_x0[3 * n + i] = acc.lpy;
    a.grad_p[i] = acc.g_eps; a.grad_p[n + i] = acc.g_mu; a.grad_p[2 * n + i] = acc.g_wx; a.grad_p[3 * n + i] = acc.g_wy;
}

constexpr int kRingSlots = 3;                 // chunks per ring
constexpr int kRingTurn = 10;                 // ring record: sx cx sy cy | 4 angle derivatives | weight, pad

// Reverse sweep for every % 8 == 0, one wavefront per 64 particles holding
// both dependency chains in registers: per step it runs the adjoint back
// through the current chunk (states X) while it recomputes the states of the
// chunk before it (Y) from that chunk's sub-checkpoint, one turn of each per
// scheduling region, so the two chains interleave inside the wavefront. The
// wavefront stages table records, angle derivatives and weights into its own
// LDS ring of 3 chunks (one vector load and one LDS store per lane and chunk,
// loaded a step before the step that first reads them) and reads a turn's
// entries one region ahead. Nothing crosses wavefronts: no workgroup barrier,
// no state traffic through LDS. The window's top chunk (possibly short) runs
// first with scalar table loads.
constexpr int kFusedBlock = 256;

template <int MODE, bool OCT>
__global__ void __launch_bounds__(kFusedBlock) k_adjoint_sweep_fused(AdjointSweepArgs a)
{
    __shared__ __align__(16) double ring_all[(kFusedBlock / 64) * kRingSlots * kChunk * kRingTurn];
    const uint32_t lane = threadIdx.x & 63;
    double *const ring = ring_all + (threadIdx.x >> 6) * (kRingSlots * kChunk * kRingTurn);
    const uint32_t i = blockIdx.x * kFusedBlock + threadIdx.x;
    const size_t n = a.n;
    // lanes past n read particle n-1; lost particles compute on parked or stale
    // states; only live lanes store
    const uint32_t ic = i < a.n ? i : a.n - 1;
    const bool live = i < a.n && a.alive[ic];
    if (!live && i < a.n && a.last_window)
        for (int c = 0; c < 4; c++) { a.grad_x0[c * n + i] = 0.0; a.grad_p[c * n + i] = 0.0; }
    if (!__any(live)) return;  // the whole wavefront

    const double mu = a.mu, m3mu = -3.0 * mu, mu3 = 3.0 * mu, mu6 = 6.0 * mu;
    const uint32_t T = a.n_turns, K = a.every;
    const uint32_t n_seg = (T + K - 1) / K;
    const uint32_t K8 = K / kChunk;
    const uint32_t G = (T + kChunk - 1) / kChunk;
    const int g_lo = (int)(a.seg_first * K8);
    const int gt = (int)umin(a.seg_end * K8, G) - 1;
    const int steps = gt - g_lo;  // full chunks below the top one
    const double *tabd = reinterpret_cast<const double *>(a.tab);
    const double *sub_base = a.sub ? a.sub : a.ckpt;
    auto clampg = [&](int g) { return g < g_lo ? g_lo : g; };
    auto slot = [&](int g) { return ring + ((uint32_t)g % kRingSlots) * (kChunk * kRingTurn); };
    auto load4 = [&](const double *p, double (&v)[4]) {
        v[0] = p[ic]; v[1] = p[n + ic]; v[2] = p[2 * n + ic]; v[3] = p[3 * n + ic];
    };
    auto start_at = [&](uint32_t seg, uint32_t j) -> const double * {
        return j == 0 ? a.ckpt + (size_t)seg * 4 * n
                      : sub_base + ((size_t)(seg - a.seg_first) * (K8 - 1) + j - 1) * 4 * n;
    };
    // ring record of a turn: sx cx sy cy | 4 angle derivatives | weight, pad.
    // Lane l < 32 stages table double l (turn l/4), lane l >= 32 derivative
    // double l-32, every lane the weight of turn l % 8 (eight identical writes).
    auto stage_load = [&](int g, double &e0, double &e1) {
        const size_t b = (size_t)g * kChunk;
        const double *src = lane < 32 ? tabd + 4 * b + lane : a.dtab + 4 * b + (lane - 32);
        e0 = *src;
        e1 = a.weights[b + (lane & 7)];
    };
    auto stage_store = [&](int g, double e0, double e1) {
        double *s = slot(g);
        s[((lane & 31) >> 2) * kRingTurn + (lane & 3) + (lane < 32 ? 0 : 4)] = e0;
        s[(lane & 7) * kRingTurn + 8] = e1;
    };
    auto rec_p = [&](int g, int m) -> TabEntry {
        const double *q = slot(g) + m * kRingTurn;
        return TabEntry{q[0], q[1], q[2], q[3]};
    };
    auto rec_c = [&](int g, int m) -> AdjRec {
        const double *q = slot(g) + m * kRingTurn;
        return AdjRec{TabEntry{q[0], q[1], q[2], q[3]}, double4{q[4], q[5], q[6], q[7]}, q[8]};
    };

    AdjAcc acc;
    if (a.last_window) {
        acc = AdjAcc{a.seed_s[ic], a.seed_s[n + ic], a.seed_s[2 * n + ic], a.seed_s[3 * n + ic], 0.0, 0.0, 0.0, 0.0};
    } else {
        acc = AdjAcc{a.grad_x0[ic], a.grad_x0[n + ic], a.grad_x0[2 * n + ic], a.grad_x0[3 * n + ic],
                     a.grad_p[ic],  a.grad_p[n + ic],  a.grad_p[2 * n + ic],  a.grad_p[3 * n + ic]};
    }
    const double bseed = 2.0 * (a.seed_b ? a.seed_b[ic] : 0.0);
    const ConstPtr<TabEntry> tab = as_const(a.tab);
    const ConstPtr<double> dtab = as_const(a.dtab), weights = as_const(a.weights);

    // Ring schedule: the step sweeping chunk g first stores chunk g-1 (loaded
    // one step earlier), which its producer half reads, into the slot chunk g+2
    // held; its consumer half reads chunk g.
    double e0 = 0.0, e1 = 0.0;
    if (steps > 0) {
        stage_load(gt - 1, e0, e1);
        stage_store(gt - 1, e0, e1);
        stage_load(clampg(gt - 2), e0, e1);
        __builtin_amdgcn_wave_barrier();
    }
    // Sub-checkpoint walk of the producer (chunk h, segment hs, index hj):
    // P is the start of the chunk produced next.
    double P[4];
    int h = g_lo;
    uint32_t hs = 0, hj = 0;
    {
        const int g1 = clampg(gt - 1);
        const uint32_t s1 = (uint32_t)g1 / K8;
        load4(start_at(s1, (uint32_t)g1 - s1 * K8), P);
        h = clampg(gt - 2);
        hs = (uint32_t)h / K8;
        hj = (uint32_t)h - hs * K8;
    }
    auto next_start = [&](double (&Q)[4]) {
        load4(start_at(hs, hj), Q);
        const bool down = h > g_lo, wrap = hj == 0;
        h = down ? h - 1 : h;
        hs = down && wrap ? hs - 1 : hs;
        hj = down ? (wrap ? K8 - 1 : hj - 1) : hj;
    };

    double SA[kChunk][4], SB[kChunk][4], nxt[4];
    {
        // the top chunk (scalar table loads) while chunk gt-1 is produced into SA
        const uint32_t ct = (uint32_t)gt * kChunk, lent = umin(kChunk, T - ct);
        const uint32_t st_seg = (uint32_t)gt / K8;
        double S[kChunk][4];
        load4(start_at(st_seg, (uint32_t)gt - st_seg * K8), S[0]);
#pragma unroll
        for (int m = 1; m < kChunk; ++m) {
            copy4(S[m], S[m - 1]);
            if ((uint32_t)m < lent) step4<MODE, OCT>(S[m], mu, m3mu, load_record(tab + ct + m - 1));
        }
        if (a.seg_end < n_seg) {
            load4(a.ckpt + (size_t)a.seg_end * 4 * n, nxt);  // the state after the window
        } else {
            double v[4];
#pragma unroll
            for (int m = 0; m < kChunk; ++m)
                if ((uint32_t)m + 1 == lent) copy4(v, S[m]);
            step4<MODE, OCT>(v, mu, m3mu, load_record(tab + T - 1));
            copy4(nxt, v);
        }
        double Q[4];
        next_start(Q);
        copy4(SA[0], P);
        const int gp = clampg(gt - 1);
#pragma unroll
        for (int m = 1; m < kChunk; ++m) {
            copy4(SA[m], SA[m - 1]);
            step4<MODE, OCT>(SA[m], mu, m3mu, rec_p(gp, m - 1));
        }
        copy4(P, Q);
#pragma unroll
        for (int m = kChunk - 1; m >= 0; --m) {
            if ((uint32_t)m >= lent) continue;
            const uint32_t k = ct + m;
            const ConstPtr<double> d = dtab + 4 * (size_t)k;
            const double4 dt{d[0], d[1], d[2], d[3]};
            if ((uint32_t)m + 1 == lent)
                adjoint_turn<OCT>(S[m], nxt, load_record(tab + k), bseed * weights[k], dt, mu3, mu6, acc);
            else
                adjoint_turn<OCT>(S[m], S[m + 1 < kChunk ? m + 1 : m], load_record(tab + k), bseed * weights[k], dt,
                                  mu3, mu6, acc);
        }
        copy4(nxt, S[0]);
    }

    // One step: the adjoint through chunk g (states X) and the states of chunk
    // g-1 into Y. In the last step the producer runs on clamped inputs and its
    // states are not used.
    auto step = [&](double (&X)[kChunk][4], double (&Y)[kChunk][4], int c) __attribute__((always_inline)) {
        const int g = gt - c, gp = clampg(g - 1);
        stage_store(gp, e0, e1);
        stage_load(clampg(g - 2), e0, e1);
        __builtin_amdgcn_wave_barrier();  // the ring stores precede later reads
        double Q[4];
        next_start(Q);
        copy4(Y[0], P);
        double post[4];
        copy4(post, nxt);
        AdjRec rc = rec_c(g, kChunk - 1);
        TabEntry rp = rec_p(gp, 0);
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
            const int m = kChunk - 1 - j;
            AdjRec rcn{};
            TabEntry rpn{};
            if (m > 0) rcn = rec_c(g, m - 1);
            if (j + 2 < kChunk) rpn = rec_p(gp, j + 1);
            adjoint_turn<OCT>(X[m], post, rc.t, bseed * rc.w, rc.dt, mu3, mu6, acc);
            copy4(post, X[m]);
            if (j + 1 < kChunk) {
                copy4(Y[j + 1], Y[j]);
                step4<MODE, OCT>(Y[j + 1], mu, m3mu, rp);
            }
            rc = rcn;
            rp = rpn;
            // one turn of each chain per scheduling region; the next region's
            // ring entries are requested one region ahead
            __builtin_amdgcn_sched_barrier(0);
        }
        copy4(nxt, X[0]);
        copy4(P, Q);
    };
    for (int c = 1; c <= steps; c += 2) {
        step(SA, SB, c);
        if (c + 1 > steps) break;
        step(SB, SA, c + 1);
    }
    if (!live) return;
    a.grad_x0[i] = acc.lx; a.grad_x0[n + i] = acc.lpx; a.grad_x0[2 * n + i] = acc.ly; a.grad_x0[3 * n + i] = acc.lpy;
    a.grad_p[i] = acc.g_eps; a.grad_p[n + i] = acc.g_mu; a.grad_p[2 * n + i] = acc.g_wx; a.grad_p[3 * n + i] = acc.g_wy;
}

template <int MODE, bool OCT>
static void launch_forward(bool aligned, uint32_t grid, hipStream_t s, const AdjointFwdArgs &a)
{
    if (aligned) hipLaunchKernelGGL((k_adjoint_forward<MODE, OCT, true>), dim3(grid), dim3(kFwdBlock), 0, s, a);
    else hipLaunchKernelGGL((k_adjoint_forward<MODE, OCT, false>), dim3(grid), dim3(kFwdBlock), 0, s, a);
}

template <int MODE, bool OCT>
static void launch_sweep(bool aligned, size_t n, hipStream_t s, const AdjointSweepArgs &w)
{
    if (aligned)
        hipLaunchKernelGGL((k_adjoint_sweep_fused<MODE, OCT>), dim3((uint32_t)((n + kFusedBlock - 1) / kFusedBlock)),
                           dim3(kFusedBlock), 0, s, w);
    else hipLaunchKernelGGL((k_adjoint_sweep<MODE, OCT>), dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, w);
}

// Scratch layout: sub-checkpoints of every 8-turn chunk but the first, per
// segment. A window is as many segments as the scratch holds.
static size_t segment_bytes(size_t n, uint32_t every)
{
    return (size_t)((every + kChunk - 1) / kChunk - 1) * 4 * n * sizeof(double);
}

// Default scratch: every segment at once, up to 4 GiB (HM_ADJOINT_SCRATCH_BYTES
// overrides the cap; at least one segment).
static constexpr uint64_t kAdjointScratchCap = 4ull << 30;
// Turn indices, segment ends (s0 + every <= 2 n_turns) stay within 32 bits.
static constexpr uint32_t kAdjointMaxTurns = 1u << 31;

static uint64_t adjoint_scratch_cap()
{
    const char *e = std::getenv("HM_ADJOINT_SCRATCH_BYTES");
    if (e && *e) {
        const unsigned long long v = std::strtoull(e, nullptr, 10);
        if (v > 0) return v;
    }
    return kAdjointScratchCap;
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
        if (n_turns > kAdjointMaxTurns) throw Error(HM_ERR_RANGE, "adjoint: more than 2^31 turns");
        if (mode != HM_MODE_EXACT && mode != HM_MODE_FAST) throw Error(HM_ERR_ARG, "adjoint: unknown mode");
        if (n == 0) return;
        if (n > UINT32_MAX) throw Error(HM_ERR_RANGE, "adjoint: too many particles");
        every = std::min(every, n_turns);  // one segment either way
        AdjointFwdArgs a{x0, px0, y0, py0, (uint32_t)n, reinterpret_cast<const TabEntry *>(table), n_turns, every,
                         mu, barrier * barrier, weights, checkpoints, final_state, birk, alive};
        hipStream_t s = as_stream(stream);
        const uint32_t grid = (uint32_t)((n + kFwdBlock - 1) / kFwdBlock);
        const bool oct = mu != 0.0;
        const bool aligned = every % 8 == 0;
        LaunchTimer timer("k_adjoint_forward", s);
        if (mode == HM_MODE_EXACT) {
            if (oct) launch_forward<kExact, true>(aligned, grid, s, a);
            else launch_forward<kExact, false>(aligned, grid, s, a);
        } else {
            if (oct) launch_forward<kFast, true>(aligned, grid, s, a);
            else launch_forward<kFast, false>(aligned, grid, s, a);
        }
        launch_check("k_adjoint_forward");
    });
}

extern "C" int hm_adjoint_scratch_bytes(size_t n, uint32_t n_turns, uint32_t every, size_t *bytes)
{
    return guarded([&] {
        if (!bytes || every == 0) throw Error(HM_ERR_ARG, "adjoint: bad arguments");
        if (n_turns) every = std::min(every, n_turns);
        const uint64_t n_seg = n_turns ? (n_turns + (uint64_t)every - 1) / every : 0;
        const uint64_t seg = segment_bytes(n, every);
        const uint64_t fit = seg ? std::max<uint64_t>(1, adjoint_scratch_cap() / seg) : 0;
        *bytes = (size_t)(std::min<uint64_t>(n_seg, fit) * seg);
    });
}

extern "C" int hm_adjoint_backward(const double *checkpoints, size_t n, const double *table, const double *dtable,
                                   size_t table_len, uint32_t n_turns, double mu, uint32_t every,
                                   const double *weights, const double *seed_state, const double *seed_birk,
                                   const uint8_t *alive, double *scratch, size_t scratch_bytes, double *grad_x0,
                                   double *grad_params, int mode, void *stream)
{
    return guarded([&] {
        if (every == 0) throw Error(HM_ERR_ARG, "adjoint: checkpoint interval must be positive");
        if (n_turns == 0 || n_turns > table_len) throw Error(HM_ERR_RANGE, "adjoint: n_turns must lie in 1..table_len");
        if (n_turns > kAdjointMaxTurns) throw Error(HM_ERR_RANGE, "adjoint: more than 2^31 turns");
        if (mode != HM_MODE_EXACT && mode != HM_MODE_FAST) throw Error(HM_ERR_ARG, "adjoint: unknown mode");
        if (n == 0) return;
        if (n > UINT32_MAX) throw Error(HM_ERR_RANGE, "adjoint: too many particles");
        every = std::min(every, n_turns);  // one segment either way
        if (!checkpoints || !table || !dtable || !weights || !seed_state || !alive || !grad_x0 || !grad_params)
            throw Error(HM_ERR_ARG, "adjoint: NULL pointer");
        const uint32_t n_seg = (n_turns + every - 1) / every;
        const uint32_t seg_chunks = (every + kChunk - 1) / kChunk;
        const size_t seg = segment_bytes(n, every);
        // segments per window: all of them when no sub-checkpoints are needed
        // (every <= 8); else what the scratch holds (grid.y bounds one launch)
        uint32_t win = n_seg;
        if (seg) {
            if (!scratch) throw Error(HM_ERR_ARG, "adjoint: NULL scratch");
            if (scratch_bytes < seg)
                throw Error(HM_ERR_ARG, "adjoint: scratch smaller than one segment (see hm_adjoint_scratch_bytes)");
            win = (uint32_t)std::min<uint64_t>({(uint64_t)n_seg, scratch_bytes / seg, 65535});
        }
        hipStream_t s = as_stream(stream);
        const uint32_t grid = (uint32_t)((n + 255) / 256);
        const bool oct = mu != 0.0;
        const bool exact = mode == HM_MODE_EXACT;
        const bool aligned = every % kChunk == 0;
        // segments [first, end): sub-checkpoints at sub
        auto segments = [&](uint32_t first, uint32_t end, double *sub) {
            AdjointSegArgs g{checkpoints, (uint32_t)n, reinterpret_cast<const TabEntry *>(table), n_turns, every, mu,
                             alive, sub, first, seg_chunks};
            const dim3 gd(grid, end - first);
            if (exact) {
                if (oct) hipLaunchKernelGGL((k_adjoint_segments<kExact, true>), gd, dim3(256), 0, s, g);
                else hipLaunchKernelGGL((k_adjoint_segments<kExact, false>), gd, dim3(256), 0, s, g);
            } else {
                if (oct) hipLaunchKernelGGL((k_adjoint_segments<kFast, true>), gd, dim3(256), 0, s, g);
                else hipLaunchKernelGGL((k_adjoint_segments<kFast, false>), gd, dim3(256), 0, s, g);
            }
            launch_check("k_adjoint_segments");
        };
        auto sweep = [&](uint32_t first, uint32_t end, const double *sub) {
            AdjointSweepArgs w{checkpoints, (uint32_t)n, reinterpret_cast<const TabEntry *>(table), dtable, n_turns,
                               every, mu, weights, seed_state, seed_birk, alive, sub, first, end, seg_chunks,
                               end == n_seg ? 1 : 0, grad_x0, grad_params};
            if (exact) {
                if (oct) launch_sweep<kExact, true>(aligned, n, s, w);
                else launch_sweep<kExact, false>(aligned, n, s, w);
            } else {
                if (oct) launch_sweep<kFast, true>(aligned, n, s, w);
                else launch_sweep<kFast, false>(aligned, n, s, w);
            }
            launch_check("k_adjoint_sweep");
        };
        LaunchTimer timer("k_adjoint_backward", s);
        for (uint32_t end = n_seg; end > 0;) {
            const uint32_t first = end > win ? end - win : 0;
            if (seg) segments(first, end, scratch);
            sweep(first, end, scratch);
            end = first;
        }
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
