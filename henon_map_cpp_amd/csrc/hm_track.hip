// Tracking kernels: the persistent turn loop (replaces one launch per turn of
// src/henon.cu:3155-3185 / gpu_henon_track :636-691), readback observables
// (:67-111) and the per-particle tangent matrix (:3187-3219, :175-249).
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <algorithm>
#include <mutex>
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
__device__ __forceinline__ void track_lanes(double &x, double &px, double &y, double &py, bool &alive,
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

// P particles per lane, n_turns turns in registers. Wavefront w owns the 64*P
// consecutive entries [64*P*w, 64*P*(w+1)) of the particle list (or of the
// live-index list `idx_in`), lane l the entries 64*q + l. The particle count is
// read from device memory (`count_in`) when given, so a launch can be sized
// from an upper bound without a host round trip.
//
// The first live lane's turn index s0 is read into a scalar register; when every
// live particle of the wavefront shares it (the normal case: all particles of a
// tracker advance together), the table record of turn s0+k is the same for the
// whole wavefront and is fetched with a scalar load, an iteration ahead, straight
// into SGPR operands of the VALU ops. Losses are kept as wave-uniform lane masks,
// so a turn costs the map's VALU ops plus one compare per particle.
struct TrackArgs {
    double *gx, *gpx, *gy, *gpy;
    uint32_t *gsteps;
    const uint32_t *idx_in, *count_in;
    uint32_t n_arg;
    uint32_t slots;  // wavefronts launched in mixed mode (0 = one particle per lane everywhere)
    const TabEntry *tab;
    uint32_t tab_len, n_turns;
    double mu, b2;
    uint32_t *idx_out, *count_out, *status;
};

template <int MODE, bool INV, bool OCT, int P>
__device__ __forceinline__ void track_wave(const TrackArgs &A, uint32_t base, uint32_t n)
{
    double *__restrict__ gx = A.gx, *__restrict__ gpx = A.gpx, *__restrict__ gy = A.gy, *__restrict__ gpy = A.gpy;
    uint32_t *__restrict__ gsteps = A.gsteps;
    const uint32_t *__restrict__ idx_in = A.idx_in;
    const TabEntry *__restrict__ tab = A.tab;
    const uint32_t tab_len = A.tab_len, n_turns = A.n_turns;
    const double mu = A.mu, b2 = A.b2;
    uint32_t *__restrict__ idx_out = A.idx_out;
    uint32_t *__restrict__ count_out = A.count_out;
    uint32_t *__restrict__ status = A.status;
    const uint32_t lane = __lane_id();
    double x[P], px[P], y[P], py[P];
    uint32_t j[P], s[P], done[P];
    bool alive0[P];
    uint64_t live[P];
    const uint64_t my_bit = 1ull << lane;
    uint64_t any_live = 0;
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const uint32_t i = base + 64u * q + lane;
        const bool in_range = i < n;
        j[q] = 0;
        s[q] = 0;
        x[q] = px[q] = y[q] = py[q] = __builtin_nan("");
        if (in_range) {
            j[q] = idx_in ? idx_in[i] : i;
            x[q] = gx[j[q]];
            px[q] = gpx[j[q]];
            y[q] = gy[j[q]];
            py[q] = gpy[j[q]];
            s[q] = gsteps[j[q]];
        }
        // henon.cu:431-434: a particle with a NaN coordinate is not stepped
        alive0[q] = in_range && !any_nan(x[q], px[q], y[q], py[q]);
        live[q] = __ballot(alive0[q]);
        any_live |= live[q];
        done[q] = n_turns;
    }
    if (!any_live) return;

    const double m3mu = -3.0 * mu;
    // turn index of the first live particle (q-major), and whether all share it
    int qf = 0;
#pragma unroll
    for (int q = P - 1; q >= 0; --q)
        if (live[q]) qf = q;
    uint32_t sf = 0;
#pragma unroll
    for (int q = 0; q < P; ++q)
        if (q == qf) sf = s[q];
    uint64_t lf = 0;
#pragma unroll
    for (int q = 0; q < P; ++q)
        if (q == qf) lf = live[q];
    const uint32_t s0 = __builtin_amdgcn_readlane(sf, __builtin_ctzll(lf));
    bool same = true;
#pragma unroll
    for (int q = 0; q < P; ++q) same = same && __all(!alive0[q] || s[q] == s0);
    const bool in_table = INV ? (s0 >= n_turns) : ((uint64_t)s0 + n_turns <= (uint64_t)tab_len);

    if (same && in_table && n_turns > 0) {
        const TabEntry *tp = INV ? tab + (s0 - 1) : tab + s0;
        // Two turns per iteration: the allocator alternates the state registers
        // instead of copying them back, and the records of the next two turns
        // are requested one iteration ahead.
        auto rec = [&](uint32_t k) -> TabEntry { return INV ? *(tp - k) : tp[k]; };
        const uint32_t last = n_turns - 1;
        TabEntry n0 = rec(0), n1 = rec(last < 1 ? last : 1);
        uint32_t k = 0;
        bool more = true;
        // exact forward mode carries x*x, y*y from one turn to the next
        constexpr bool CARRY = MODE == kExact && !INV;
        double xx[P], yy[P];
#pragma unroll
        for (int q = 0; q < P; ++q) {
            xx[q] = x[q] * x[q];
            yy[q] = y[q] * y[q];
        }
        // exact mode screens losses with a fused radius (screen_radius2) and
        // confirms candidates with the reference's expression
        constexpr bool SCREEN = MODE == kExact;
        const double b2s = SCREEN ? screen_bound(b2) : b2;
        auto turn = [&](const TabEntry &t, uint32_t kk) {
            double r2[P];
#pragma unroll
            for (int q = 0; q < P; ++q) {
                if (CARRY) {
                    r2[q] = fwd_exact_carry<OCT>(x[q], px[q], y[q], py[q], xx[q], yy[q], mu, t);
                } else {
                    map_step<MODE, INV, OCT>(x[q], px[q], y[q], py[q], mu, m3mu, t);
                    r2[q] = SCREEN ? screen_radius2(x[q], px[q], y[q], py[q]) : radius2<MODE>(x[q], px[q], y[q], py[q]);
                }
            }
            uint64_t all = 0;
#pragma unroll
            for (int q = 0; q < P; ++q) {
                uint64_t lost = live[q] & ~__ballot(r2[q] < b2s);
                if (SCREEN && lost) {
                    const double e = CARRY ? exact_radius2(xx[q], px[q], yy[q], py[q])
                                           : exact_radius2(x[q] * x[q], px[q], y[q] * y[q], py[q]);
                    lost &= ~__ballot(e < b2);
                }
                if (lost) {
                    if (lost & my_bit) done[q] = kk;
                    live[q] &= ~lost;
                }
                all |= live[q];
            }
            more = all != 0;
        };
        for (; k + 1 < n_turns; k += 2) {
            const TabEntry t0 = n0, t1 = n1;
            n0 = rec(k + 2 <= last ? k + 2 : last);
            n1 = rec(k + 3 <= last ? k + 3 : last);
            turn(t0, k);
            if (!more) break;
            turn(t1, k + 1);
            if (!more) break;
        }
        if (more && k == last) turn(n0, k);
#pragma unroll
        for (int q = 0; q < P; ++q)
            if (alive0[q] && !(live[q] & my_bit)) x[q] = px[q] = y[q] = py[q] = __builtin_nan("");
    } else if (n_turns > 0) {
#pragma unroll
        for (int q = 0; q < P; ++q) {
            bool a = alive0[q];
            track_lanes<MODE, INV, OCT>(x[q], px[q], y[q], py[q], a, done[q], s[q], tab, tab_len, n_turns, mu, b2,
                                        status);
            live[q] = __ballot(a);
        }
    }

#pragma unroll
    for (int q = 0; q < P; ++q) {
        if (alive0[q]) {
            gx[j[q]] = x[q];
            gpx[j[q]] = px[q];
            gy[j[q]] = y[q];
            gpy[j[q]] = py[q];
            gsteps[j[q]] = INV ? s[q] - done[q] : s[q] + done[q];
        }
        if (idx_out) {
            // wavefront-aggregated append of the survivors to the next live list
            const uint64_t m = live[q];
            if (m) {
                const int leader = __builtin_ctzll(m);
                uint32_t slot = 0;
                if ((int)lane == leader) slot = atomicAdd(count_out, (uint32_t)__popcll(m));
                slot = __builtin_amdgcn_readlane(slot, leader);
                if (m & my_bit) idx_out[slot + __popcll(m & (my_bit - 1ull))] = j[q];
            }
        }
    }
}

// Wavefront w owns 64*P consecutive list entries. In mixed mode (slots > 0,
// exactly `slots` wavefronts launched) the first w2 wavefronts take two particles
// per lane so that n particles fit in one resident round of wavefronts:
// w2 = ceil((n - 64 slots) / 64) when n exceeds 64 slots.
template <int MODE, bool INV, bool OCT>
__global__ void __launch_bounds__(256, 8) k_track(TrackArgs A)
{
    const uint32_t n = A.count_in ? *A.count_in : A.n_arg;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    uint32_t w2 = 0;
    if (A.slots && n > 64u * A.slots) w2 = min((n - 64u * A.slots + 63u) / 64u, A.slots);
    if (wave < w2) {
        track_wave<MODE, INV, OCT, 2>(A, wave * 128u, n);
    } else {
        const uint32_t base = w2 * 128u + (wave - w2) * 64u;
        if (base < n) track_wave<MODE, INV, OCT, 1>(A, base, n);
    }
}

struct TrackLaunch {
    double *x, *px, *y, *py;
    uint32_t *steps;
    const uint32_t *idx_in, *count_in;
    uint32_t n_arg, grid_bound;  // exact count when count_in is NULL; upper bound of it otherwise
    const TabEntry *tab;
    uint32_t tab_len, n_turns;
    double mu, b2;
    uint32_t *idx_out, *count_out;
    int block;
    uint32_t slots;  // resident wavefront slots of the device (0 = plain one-per-lane launch)
};

template <int MODE, bool INV, bool OCT>
static void launch_track(const TrackLaunch &a, hipStream_t s)
{
    TrackArgs A{a.x, a.px, a.y, a.py, a.steps, a.idx_in, a.count_in, a.n_arg, 0u, a.tab, a.tab_len, a.n_turns,
                a.mu, a.b2, a.idx_out, a.count_out, status_word()};
    uint64_t waves = ((uint64_t)a.grid_bound + 63) / 64;
    // mixed mode when one particle per lane would need more than one round of
    // resident wavefronts and two per lane suffice
    if (a.slots && waves > a.slots && (uint64_t)a.grid_bound <= 128ull * a.slots) {
        A.slots = a.slots;
        waves = a.slots;
    }
    const uint64_t grid = (waves * 64 + a.block - 1) / a.block;
    if (grid == 0) return;
    LaunchTimer timer("k_track", s);
    hipLaunchKernelGGL((k_track<MODE, INV, OCT>), dim3((uint32_t)grid), dim3(a.block), 0, s, A);
    launch_check("k_track");
}

using TrackFn = void (*)(const TrackLaunch &, hipStream_t);

static TrackFn pick_track(int mode, bool inv, bool oct)
{
    if (mode == kExact) {
        if (inv) return oct ? launch_track<kExact, true, true> : launch_track<kExact, true, false>;
        return oct ? launch_track<kExact, false, true> : launch_track<kExact, false, false>;
    }
    if (inv) return oct ? launch_track<kFast, true, true> : launch_track<kFast, true, false>;
    return oct ? launch_track<kFast, false, true> : launch_track<kFast, false, false>;
}

// Resident wavefront slots: CUs x 4 SIMDs x 8 wavefronts (the kernel keeps
// its VGPRs within 64 so that 8 fit per SIMD).
static uint32_t resident_slots()
{
    int dev = 0, cus = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    return (uint32_t)cus * 32u;
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
static constexpr int kDefaultCompactEvery = 1000;  // turns between compactions
static constexpr uint32_t kFirstSegment = 125;     // first segments are short: fast losses go early

// Pinned host words receiving the per-segment live counts.
static uint32_t *pinned_counts(size_t n)
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
        if (n > 0x7fffffffull || table_len > 0xffffffffull) throw Error(HM_ERR_ARG, "track: size exceeds 2^31");
        const int block = cfg && cfg->block ? cfg->block : kDefaultBlock;
        if (block < 64 || block > 256 || block % 64) throw Error(HM_ERR_ARG, "track: block must be 64, 128, 192 or 256");
        const int every = cfg && cfg->compact_every ? cfg->compact_every : kDefaultCompactEvery;
        // 1: one particle per lane; 2: one or two (mixed launch). The default is
        // one in exact mode and mixed in fast mode, the faster of the two for each
        // on the 1000x1000-grid, 1e5-turn scan.
        int pair_mode = cfg ? cfg->particles_per_lane : 0;
        if (pair_mode < 0 || pair_mode > 2) throw Error(HM_ERR_ARG, "track: particles_per_lane must be 0, 1 or 2");
        if (pair_mode == 0) pair_mode = mode == HM_MODE_EXACT ? 1 : 2;
        hipStream_t s = as_stream(stream);
        const double b2 = barrier * barrier;  // henon.cu:2717, :3173
        const bool oct = mu != 0.0;           // henon.cu:445-448
        const TabEntry *tab = reinterpret_cast<const TabEntry *>(table);
        const uint32_t N = (uint32_t)n;

        TrackLaunch a{};
        a.x = x; a.px = px; a.y = y; a.py = py; a.steps = steps;
        a.tab = tab; a.tab_len = (uint32_t)table_len; a.mu = mu; a.b2 = b2; a.block = block;
        a.slots = pair_mode == 1 ? 0u : resident_slots();

        if (every < 0 || (uint32_t)every >= n_turns) {
            a.n_arg = a.grid_bound = N;
            a.n_turns = n_turns;
            pick_track(mode, inverse != 0, oct)(a, s);
            return;
        }
        // Segments: kFirstSegment turns, doubling up to `every`. Survivors of a
        // segment are appended to a dense live-index list that the next segment
        // reads, with its length read on the device. The host sizes segment i's
        // grid from the count of segment i-2 (an upper bound), so it never waits
        // for the segment that feeds the launch. (Letting late segments grow past
        // `every` saves launches but carries more dead lanes: 10% slower on the
        // 1e5-turn scan.)
        std::vector<uint32_t> seg;
        for (uint32_t done = 0, len = std::min<uint32_t>(kFirstSegment, (uint32_t)every); done < n_turns;) {
            const uint32_t t = std::min(len, n_turns - done);
            seg.push_back(t);
            done += t;
            len = std::min<uint32_t>(len * 2u, (uint32_t)every);
        }
        const size_t S = seg.size();
        uint32_t *lists = static_cast<uint32_t *>(workspace(sizeof(uint32_t) * (2 * (size_t)N + S + 64), 0));
        uint32_t *counts = lists + 2 * (size_t)N;  // counts[i]: length of segment i's input list (i >= 1)
        uint32_t *host = pinned_counts(S + 1);
        hip_check(hipMemsetAsync(counts, 0, sizeof(uint32_t) * (S + 1), s), "hipMemsetAsync");
        std::vector<hipEvent_t> ev(S, nullptr);
        auto cleanup = [&] {
            for (auto e : ev)
                if (e) (void)hipEventDestroy(e);
        };
        try {
            uint64_t bound = N;  // upper bound of counts[i]
            for (size_t i = 0; i < S; ++i) {
                if (i >= 2) {
                    hip_check(hipEventSynchronize(ev[i - 2]), "hipEventSynchronize");
                    bound = std::min<uint64_t>(bound, host[i - 1]);
                    if (host[i - 1] == 0) break;  // every particle is lost or NaN
                }
                const bool last = i + 1 == S;
                a.n_turns = seg[i];
                a.idx_in = i == 0 ? nullptr : (i % 2 ? lists : lists + N);
                a.count_in = i == 0 ? nullptr : counts + i;
                a.n_arg = N;
                a.grid_bound = (uint32_t)bound;
                a.idx_out = last ? nullptr : (i % 2 ? lists + N : lists);
                a.count_out = last ? nullptr : counts + i + 1;
                pick_track(mode, inverse != 0, oct)(a, s);
                if (!last) {
                    hip_check(hipMemcpyAsync(host + i + 1, counts + i + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s),
                              "hipMemcpyAsync");
                    hip_check(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming), "hipEventCreate");
                    hip_check(hipEventRecord(ev[i], s), "hipEventRecord");
                }
            }
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
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
