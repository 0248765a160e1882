// Tracking kernels: the persistent turn loop (replaces one launch per turn of
// src/henon.cu:3155-3185 / gpu_henon_track :636-691), readback observables
// (:67-111) and the per-particle tangent matrix (:3187-3219, :175-249).
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <vector>

#include "hm_device.h"
#include "hm_internal.h"

namespace hm {

constexpr uint32_t kFlagRange = 1u;

// Per-lane fallback used when the live lanes of a wavefront do not share one
// turn index (particles reset while others were not) or when that index range
// leaves the table: each lane reads its own table record, and a lane whose
// index would leave [0, tab_len) stops unchanged and raises kFlagRange.
template <int MODE, bool INV, bool OCT>
__device__ __attribute__((noinline)) void track_lanes(double &x, double &px, double &y, double &py, bool &alive,
                                                      uint32_t &done, uint32_t s, const TabEntry *__restrict__ tab,
                                                      uint32_t tab_len, uint32_t n_turns, double mu, double b2,
                                                      uint32_t *status)
{
    const double m3mu = -3.0 * mu;
    for (uint32_t k = 0; k < n_turns; ++k) {
        if (alive) {
            const int64_t ti = INV ? (int64_t)s - 1 - (int64_t)k : (int64_t)s + (int64_t)k;
            if (ti < 0 || ti >= (int64_t)tab_len) {
                atomicOr(status, kFlagRange);
                alive = false;
                done = k;
            } else {
                const TabEntry t = tab[ti];
                map_step<MODE, INV, OCT>(x, px, y, py, mu, m3mu, t);
                if (!(radius2<MODE>(x, px, y, py) < b2)) {
                    x = px = y = py = __builtin_nan("");
                    alive = false;
                    done = k;
                }
            }
        }
        if (!__any(alive)) break;
    }
}

// One thread per particle (optionally through a live-index list), n_turns
// turns in registers. The first live lane's turn index s0 is read into a
// scalar register; when every live lane of the wavefront shares it (the normal
// case: all particles of a tracker advance together), the table record of turn
// s0+k is the same for the whole wavefront and is fetched with a scalar load,
// one turn ahead, straight into SGPR operands of the VALU ops. Losses are kept
// as a lane mask, so a turn costs the map's VALU ops plus one compare.
template <int MODE, bool INV, bool OCT>
__global__ void __launch_bounds__(256)
k_track(double *__restrict__ gx, double *__restrict__ gpx, double *__restrict__ gy, double *__restrict__ gpy,
        uint32_t *__restrict__ gsteps, const uint32_t *__restrict__ idx_in, uint32_t n,
        const TabEntry *__restrict__ tab, uint32_t tab_len, uint32_t n_turns, double mu, double b2,
        uint32_t *__restrict__ idx_out, uint32_t *__restrict__ count_out, uint32_t *__restrict__ status)
{
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in_range = tid < n;
    uint32_t j = 0;
    uint32_t s = 0;
    double x = __builtin_nan(""), px = x, y = x, py = x;
    if (in_range) {
        j = idx_in ? idx_in[tid] : tid;
        x = gx[j];
        px = gpx[j];
        y = gy[j];
        py = gpy[j];
        s = gsteps[j];
    }
    // henon.cu:431-434: a particle with a NaN coordinate is not stepped
    const bool alive0 = in_range && !any_nan(x, px, y, py);
    bool alive = alive0;
    const double m3mu = -3.0 * mu;
    const uint64_t live = __ballot(alive);
    if (live == 0) return;

    const int first = __builtin_ctzll(live);
    const uint32_t s0 = __builtin_amdgcn_readlane(s, first);
    const bool same = __all(!alive || s == s0);
    const bool in_table = INV ? (s0 >= n_turns) : ((uint64_t)s0 + n_turns <= (uint64_t)tab_len);
    uint32_t done = n_turns;  // turns survived

    if (same && in_table && n_turns > 0) {
        // `live` is a wave-uniform lane mask: the loss test is one v_cmp into an
        // SGPR pair and the rest of the bookkeeping is scalar. A lane keeps
        // computing after its loss (on values that no longer matter) and is
        // set to NaN once, after the loop.
        uint64_t live_mask = live;
        const uint64_t my_bit = 1ull << __lane_id();
        const TabEntry *tp = INV ? tab + (s0 - 1) : tab + s0;
        // Two turns per iteration: the allocator alternates the state registers
        // instead of copying them back, and the records of the next two turns
        // are requested one iteration ahead.
        auto rec = [&](uint32_t k) -> TabEntry { return INV ? *(tp - k) : tp[k]; };
        const uint32_t last = n_turns - 1;
        TabEntry n0 = rec(0), n1 = rec(last < 1 ? last : 1);
        uint32_t k = 0;
        for (; k + 1 < n_turns; k += 2) {
            const TabEntry t0 = n0, t1 = n1;
            n0 = rec(k + 2 <= last ? k + 2 : last);
            n1 = rec(k + 3 <= last ? k + 3 : last);
            map_step<MODE, INV, OCT>(x, px, y, py, mu, m3mu, t0);
            uint64_t lost = live_mask & ~__ballot(radius2<MODE>(x, px, y, py) < b2);
            if (lost) {
                if (lost & my_bit) done = k;
                live_mask &= ~lost;
                if (!live_mask) break;
            }
            map_step<MODE, INV, OCT>(x, px, y, py, mu, m3mu, t1);
            lost = live_mask & ~__ballot(radius2<MODE>(x, px, y, py) < b2);
            if (lost) {
                if (lost & my_bit) done = k + 1;
                live_mask &= ~lost;
                if (!live_mask) break;
            }
        }
        if (k == last && live_mask) {
            map_step<MODE, INV, OCT>(x, px, y, py, mu, m3mu, n0);
            const uint64_t lost = live_mask & ~__ballot(radius2<MODE>(x, px, y, py) < b2);
            if (lost & my_bit) done = k;
            live_mask &= ~lost;
        }
        alive = (live_mask & my_bit) != 0;
        if (alive0 && !alive) x = px = y = py = __builtin_nan("");
    } else if (n_turns > 0) {
        track_lanes<MODE, INV, OCT>(x, px, y, py, alive, done, s, tab, tab_len, n_turns, mu, b2, status);
    }

    if (alive0) {
        gx[j] = x;
        gpx[j] = px;
        gy[j] = y;
        gpy[j] = py;
        gsteps[j] = INV ? s - done : s + done;
    }
    if (idx_out) {
        // wavefront-aggregated append of the survivors to the next live list
        const uint64_t m = __ballot(alive);
        if (m) {
            const int leader = __builtin_ctzll(m);
            uint32_t base = 0;
            if ((int)__lane_id() == leader) base = atomicAdd(count_out, (uint32_t)__popcll(m));
            base = __builtin_amdgcn_readlane(base, leader);
            if (alive) {
                const uint64_t below = m & ((1ull << __lane_id()) - 1ull);
                idx_out[base + __popcll(below)] = j;
            }
        }
    }
}

template <int MODE, bool INV, bool OCT>
static void launch_track(double *x, double *px, double *y, double *py, uint32_t *steps, const uint32_t *idx_in,
                         uint32_t n, const TabEntry *tab, uint32_t tab_len, uint32_t n_turns, double mu, double b2,
                         uint32_t *idx_out, uint32_t *count_out, int block, hipStream_t s)
{
    const uint32_t grid = (n + block - 1) / block;
    if (grid == 0) return;
    LaunchTimer timer("k_track", s);
    hipLaunchKernelGGL((k_track<MODE, INV, OCT>), dim3(grid), dim3(block), 0, s, x, px, y, py, steps, idx_in, n,
                       tab, tab_len, n_turns, mu, b2, idx_out, count_out, status_word());
    launch_check("k_track");
}

using TrackFn = void (*)(double *, double *, double *, double *, uint32_t *, const uint32_t *, uint32_t,
                         const TabEntry *, uint32_t, uint32_t, double, double, uint32_t *, uint32_t *, int,
                         hipStream_t);

static TrackFn pick_track(int mode, bool inv, bool oct)
{
    static const TrackFn table[2][2][2] = {
        {{launch_track<kExact, false, false>, launch_track<kExact, false, true>},
         {launch_track<kExact, true, false>, launch_track<kExact, true, true>}},
        {{launch_track<kFast, false, false>, launch_track<kFast, false, true>},
         {launch_track<kFast, true, false>, launch_track<kFast, true, true>}},
    };
    return table[mode][inv ? 1 : 0][oct ? 1 : 0];
}

// ---------------------------------------------------------------- observables
// src/henon.cu:67-111, same expression order.
__global__ void k_observable(int kind, const double *__restrict__ x, const double *__restrict__ px,
                             const double *__restrict__ y, const double *__restrict__ py, uint32_t n,
                             double *__restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double a = x[i], b = px[i], c = y[i], d = py[i];
    double r;
    switch (kind) {
    case HM_OBS_RADIUS: r = sqrt((((a * a) + (b * b)) + (c * c)) + (d * d)); break;
    case HM_OBS_ACTION: r = (((a * a) + (b * b)) + (c * c)) + (d * d); break;
    case HM_OBS_ACTION_X: r = (a * a) + (b * b); break;
    case HM_OBS_ACTION_Y: r = (c * c) + (d * d); break;
    case HM_OBS_ANGLE_X: r = atan2(b, a); break;
    default: r = atan2(d, c); break;
    }
    out[i] = r;
}

// ---------------------------------------------------------------- tangent matrix
// J at table[ti] (henon.cu:175-196) or J^-1 (henon.cu:224-249; pow(u,2) = u*u).
__global__ void k_tangent_matrix(const double *__restrict__ x, const double *__restrict__ px,
                                 const double *__restrict__ y, const double *__restrict__ py,
                                 const uint8_t *__restrict__ valid, uint32_t n, const TabEntry *__restrict__ tab,
                                 uint32_t ti, double mu, int reverse, double *__restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double m[16];
    if (valid && !valid[i]) {
        for (int k = 0; k < 16; k++) m[k] = __builtin_nan("");
    } else {
        const TabEntry t = tab[ti];
        if (!reverse) {
            const Jac jc = jac_coeffs<kExact>(x[i], y[i], mu * 3, mu * 6);
            jac_dense_exact(jc, t, m);
        } else {
            const double sx = t.sx, cx = t.cx, sy = t.sy, cy = t.cy;
            const double u = (cx * x[i]) - (px[i] * sx);
            const double v = (cy * y[i]) - (py[i] * sy);
            const double d = (u * u) - (v * v);
            m[0] = cx; m[1] = -sx; m[2] = 0; m[3] = 0;
            m[4] = ((((-3 * cx) * mu) * d) - ((2 * cx) * u)) + sx;
            m[5] = (cx + (((3 * mu) * sx) * d)) + ((2 * sx) * u);
            m[6] = ((2 * cy) * v) * (((3 * mu) * u) + 1);
            m[7] = ((2 * sy) * v) * (((-3 * mu) * u) - 1);
            m[8] = 0; m[9] = 0; m[10] = cy; m[11] = -sy;
            m[12] = ((2 * cx) * v) * (((3 * mu) * u) + 1);
            m[13] = ((2 * sx) * v) * (((-3 * mu) * u) - 1);
            m[14] = ((((3 * cy) * mu) * d) + ((2 * cy) * u)) + sy;
            m[15] = (cy - (((3 * mu) * sy) * d)) - ((2 * sy) * u);
        }
    }
    for (int k = 0; k < 16; k++) out[(size_t)i * 16 + k] = m[k];
}

}  // namespace hm

using namespace hm;

static constexpr int kDefaultBlock = 256;
static constexpr int kDefaultCompactEvery = 0;  // compaction off until tuned on hardware

extern "C" int hm_track(double *x, double *px, double *y, double *py, uint32_t *steps, size_t n,
                        const double *table, size_t table_len, uint32_t n_turns, double mu, double barrier,
                        double kick_module, int inverse, int mode, const hm_launch_cfg *cfg, void *stream)
{
    return guarded([&] {
        if (!std::isnan(kick_module))
            throw Error(HM_ERR_UNSUPPORTED, "track: the random kick (kick_module) is not available in this build");
        if (mode != HM_MODE_EXACT && mode != HM_MODE_FAST) throw Error(HM_ERR_ARG, "track: unknown mode");
        if (n == 0 || n_turns == 0) return;
        if (!x || !px || !y || !py || !steps || !table) throw Error(HM_ERR_ARG, "track: NULL pointer");
        if (n > 0xffffffffull || table_len > 0xffffffffull) throw Error(HM_ERR_ARG, "track: size exceeds 2^32");
        int block = cfg && cfg->block ? cfg->block : kDefaultBlock;
        if (block < 64 || block > 1024 || block % 64) throw Error(HM_ERR_ARG, "track: block must be 64..1024, /64");
        if (block > 256) block = 256;  // kernel is compiled for at most 256 threads
        int every = cfg && cfg->compact_every ? cfg->compact_every : kDefaultCompactEvery;
        hipStream_t s = as_stream(stream);
        const double b2 = barrier * barrier;  // henon.cu:2717, :3173
        const bool oct = mu != 0.0;           // henon.cu:445-448
        TrackFn fn = pick_track(mode, inverse != 0, oct);
        const TabEntry *tab = reinterpret_cast<const TabEntry *>(table);
        const uint32_t N = (uint32_t)n;

        if (every <= 0 || (uint32_t)every >= n_turns) {
            fn(x, px, y, py, steps, nullptr, N, tab, (uint32_t)table_len, n_turns, mu, b2, nullptr, nullptr, block, s);
            return;
        }
        // Segments of `every` turns; survivors of each segment are compacted into
        // a dense index list so later segments run on full wavefronts.
        uint32_t *lists = static_cast<uint32_t *>(workspace(sizeof(uint32_t) * (2 * (size_t)N + 64), 0));
        uint32_t *count = lists + 2 * (size_t)N;
        uint32_t *cur = nullptr, *nxt = lists;
        uint32_t n_cur = N, remaining = n_turns;
        while (remaining > 0 && n_cur > 0) {
            const uint32_t turns = remaining < (uint32_t)every ? remaining : (uint32_t)every;
            const bool last = turns == remaining;
            if (!last) hip_check(hipMemsetAsync(count, 0, sizeof(uint32_t), s), "hipMemsetAsync");
            fn(x, px, y, py, steps, cur, n_cur, tab, (uint32_t)table_len, turns, mu, b2, last ? nullptr : nxt,
               last ? nullptr : count, block, s);
            remaining -= turns;
            if (!last) {
                uint32_t c = 0;
                hip_check(hipMemcpyAsync(&c, count, sizeof(uint32_t), hipMemcpyDeviceToHost, s), "hipMemcpyAsync");
                hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
                n_cur = c;
                cur = nxt;
                nxt = (nxt == lists) ? lists + N : lists;
            }
        }
    });
}

extern "C" int hm_observable(int kind, const double *x, const double *px, const double *y, const double *py,
                             size_t n, double *out, void *stream)
{
    return guarded([&] {
        if (kind < HM_OBS_RADIUS || kind > HM_OBS_ANGLE_Y) throw Error(HM_ERR_ARG, "observable: unknown kind");
        if (n == 0) return;
        if (!x || !px || !y || !py || !out) throw Error(HM_ERR_ARG, "observable: NULL pointer");
        hipStream_t s = as_stream(stream);
        const uint32_t grid = (uint32_t)((n + 255) / 256);
        LaunchTimer timer("k_observable", s);
        hipLaunchKernelGGL(k_observable, dim3(grid), dim3(256), 0, s, kind, x, px, y, py, (uint32_t)n, out);
        launch_check("k_observable");
    });
}

extern "C" int hm_tangent_matrix(const double *x, const double *px, const double *y, const double *py,
                                 const uint8_t *valid, size_t n, const double *table, size_t table_len,
                                 uint32_t table_index, double mu, int reverse, double *out, void *stream)
{
    return guarded([&] {
        if (n == 0) return;
        if (!x || !px || !y || !py || !table || !out) throw Error(HM_ERR_ARG, "tangent_matrix: NULL pointer");
        if (table_index >= table_len) throw Error(HM_ERR_RANGE, "tangent_matrix: table index out of range");
        hipStream_t s = as_stream(stream);
        const uint32_t grid = (uint32_t)((n + 255) / 256);
        LaunchTimer timer("k_tangent_matrix", s);
        hipLaunchKernelGGL(k_tangent_matrix, dim3(grid), dim3(256), 0, s, x, px, y, py, valid, (uint32_t)n,
                           reinterpret_cast<const TabEntry *>(table), table_index, mu, reverse, out);
        launch_check("k_tangent_matrix");
    });
}
