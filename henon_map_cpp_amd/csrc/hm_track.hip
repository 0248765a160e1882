// Tracking kernels: the persistent turn loop (replaces one launch per turn of
// src/henon.cu:3155-3185 / gpu_henon_track :636-691), readback observables
// (:67-111) and the per-particle tangent matrix (:3187-3219, :175-249).
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "hm_device.h"
#include "hm_internal.h"
#include "hm_wave.h"

namespace hm {

// Wavefront w owns 64*P consecutive list entries. In mixed mode (slots > 0,
// exactly `slots` wavefronts launched) the first w2 wavefronts take two particles
// per lane so that n particles fit in one resident round of wavefronts:
// w2 = ceil((n - 64 slots) / 64) when n exceeds 64 slots.
// RING: the table records come through a per-wavefront LDS ring instead of
// scalar loads (the launch picks it when there are few wavefronts per SIMD).
template <int MODE, bool INV, bool OCT, bool RING>
__global__ void __launch_bounds__(256, RING ? 2 : 8) k_track(TrackArgs A)
{
    const uint32_t n = A.count_in ? *A.count_in : A.n_arg;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    uint32_t w2 = 0;
    if (A.slots && n > 64u * A.slots) w2 = min((n - 64u * A.slots + 63u) / 64u, A.slots);
    if (wave < w2) {
        track_wave<MODE, INV, OCT, 2, RING>(A, wave * 128u, n);
    } else {
        const uint32_t base = w2 * 128u + (wave - w2) * 64u;
        if (base < n) track_wave<MODE, INV, OCT, 1, RING>(A, base, n);
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
    int staging;     // HM_STAGING_*: table records by scalar loads or through the LDS ring
    uint32_t simds;  // SIMDs of the device (CUs x 4)
};

// Automatic table staging: the LDS ring up to this many wavefronts per SIMD.
static constexpr uint64_t kRingMaxWavesPerSimd = 2;

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
    const bool ring = a.staging == HM_STAGING_RING ||
                      (a.staging == HM_STAGING_AUTO && waves <= (uint64_t)kRingMaxWavesPerSimd * a.simds);
    LaunchTimer timer("k_track", s);
    if (ring) hipLaunchKernelGGL((k_track<MODE, INV, OCT, true>), dim3((uint32_t)grid), dim3(a.block), 0, s, A);
    else hipLaunchKernelGGL((k_track<MODE, INV, OCT, false>), dim3((uint32_t)grid), dim3(a.block), 0, s, A);
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

// Launch configurations found by hm_autotune_track, per device, arithmetic
// mode, octupole on/off and particle-count class (floor(log2 n)).
struct TuneKey {
    int dev, mode, oct, nbits;
    bool operator<(const TuneKey &o) const
    {
        return std::tie(dev, mode, oct, nbits) < std::tie(o.dev, o.mode, o.oct, o.nbits);
    }
};
static std::mutex g_tune_mu;
static std::map<TuneKey, hm_launch_cfg> g_tuned;

static TuneKey tune_key(int mode, bool oct, size_t n)
{
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    int nbits = 0;
    while ((size_t(2) << nbits) <= n) ++nbits;
    return TuneKey{dev, mode, oct ? 1 : 0, nbits};
}

static bool tuned_config(int mode, bool oct, size_t n, hm_launch_cfg *out)
{
    const TuneKey k = tune_key(mode, oct, n);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(k);
    if (it == g_tuned.end()) return false;
    *out = it->second;
    return true;
}

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

// Argument checks and the launch description shared by hm_track and
// hm_track_noalloc. Returns false when there is nothing to do.
static bool prepare_track(double *x, double *px, double *y, double *py, uint32_t *steps, size_t n,
                          const double *table, size_t table_len, uint32_t n_turns, double mu, double barrier,
                          double kick_module, int mode, const hm_launch_cfg *cfg, const char *who, TrackLaunch &a,
                          int &every)
{
    const std::string w(who);
    if (!std::isnan(kick_module))
        throw Error(HM_ERR_ARG, w + ": a finite kick_module needs hm_track_kicked (seed and draw index)");
    if (mode != HM_MODE_EXACT && mode != HM_MODE_FAST) throw Error(HM_ERR_ARG, w + ": unknown mode");
    if (n == 0 || n_turns == 0) return false;
    if (!x || !px || !y || !py || !steps || !table) throw Error(HM_ERR_ARG, w + ": NULL pointer");
    if (n > 0x7fffffffull || table_len > 0xffffffffull) throw Error(HM_ERR_ARG, w + ": size exceeds 2^31");
    hm_launch_cfg tuned{};
    if (!cfg && tuned_config(mode, mu != 0.0, n, &tuned)) cfg = &tuned;
    const int block = cfg && cfg->block ? cfg->block : kDefaultBlock;
    if (block < 64 || block > 256 || block % 64) throw Error(HM_ERR_ARG, w + ": block must be 64, 128, 192 or 256");
    every = cfg && cfg->compact_every ? cfg->compact_every : kDefaultCompactEvery;
    // 1: one particle per lane; 2: one or two (mixed launch). The default is
    // one in exact mode and mixed in fast mode, the faster of the two for each
    // on the 1000x1000-grid, 1e5-turn scan.
    int pair_mode = cfg ? cfg->particles_per_lane : 0;
    if (pair_mode < 0 || pair_mode > 2) throw Error(HM_ERR_ARG, w + ": particles_per_lane must be 0, 1 or 2");
    if (pair_mode == 0) pair_mode = mode == HM_MODE_EXACT ? 1 : 2;
    a = TrackLaunch{};
    a.x = x; a.px = px; a.y = y; a.py = py; a.steps = steps;
    a.tab = reinterpret_cast<const TabEntry *>(table);
    a.tab_len = (uint32_t)table_len;
    a.mu = mu;
    a.b2 = barrier * barrier;  // henon.cu:2717, :3173
    a.block = block;
    a.slots = pair_mode == 1 ? 0u : resident_slots();
    a.staging = cfg ? cfg->table_staging : HM_STAGING_AUTO;
    if (a.staging < HM_STAGING_AUTO || a.staging > HM_STAGING_RING)
        throw Error(HM_ERR_ARG, w + ": table_staging must be 0 (auto), 1 (scalar) or 2 (ring)");
    a.simds = device_simds();
    return true;
}

// Segment lengths: kFirstSegment turns, doubling up to `every`.
static std::vector<uint32_t> segment_plan(uint32_t n_turns, int every)
{
    std::vector<uint32_t> seg;
    for (uint32_t done = 0, len = std::min<uint32_t>(kFirstSegment, (uint32_t)every); done < n_turns;) {
        const uint32_t t = std::min(len, n_turns - done);
        seg.push_back(t);
        done += t;
        len = std::min<uint32_t>(len * 2u, (uint32_t)every);
    }
    return seg;
}

extern "C" int hm_track(double *x, double *px, double *y, double *py, uint32_t *steps, size_t n,
                        const double *table, size_t table_len, uint32_t n_turns, double mu, double barrier,
                        double kick_module, int inverse, int mode, const hm_launch_cfg *cfg, void *stream)
{
    return guarded([&] {
        TrackLaunch a{};
        int every = 0;
        if (!prepare_track(x, px, y, py, steps, n, table, table_len, n_turns, mu, barrier, kick_module, mode, cfg,
                           "track", a, every))
            return;
        hipStream_t s = as_stream(stream);
        const bool oct = mu != 0.0;  // henon.cu:445-448
        const uint32_t N = (uint32_t)n;

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
        const std::vector<uint32_t> seg = segment_plan(n_turns, every);
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

// ---------------------------------------------------------------- allocation-free tracking
// hm_track_noalloc keeps its live-index lists and per-segment counts in this
// file-scope array instead of the library's growable workspace. Each device
// holds its own copy (8.3 MiB) from module load until the process ends. Host
// code never passes the array's name to a kernel: the name is only a host-side
// stand-in; the device address comes from hipGetSymbolAddress, per device.
constexpr uint32_t kNoAllocMaxParticles = 1u << 20;
constexpr uint32_t kNoAllocMaxSegments = 1u << 16;
__device__ __align__(256) uint32_t scratch[2 * (size_t)kNoAllocMaxParticles + kNoAllocMaxSegments + 64];

static constexpr int kMaxDevices = 64;
static std::mutex g_noalloc_mu;
static uint32_t *g_noalloc_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the first
// call there.
static uint32_t *noalloc_scratch()
{
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) throw Error(HM_ERR_ARG, "track_noalloc: device index beyond 64");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        uint32_t *ptr = nullptr;
        hip_check(hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_noalloc_ptr[dev] = ptr;
    }
    return g_noalloc_ptr[dev];
}

__global__ void k_zero_words(uint32_t *__restrict__ p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

extern "C" int hm_track_noalloc(double *x, double *px, double *y, double *py, uint32_t *steps, size_t n,
                                const double *table, size_t table_len, uint32_t n_turns, double mu, double barrier,
                                double kick_module, int inverse, int mode, const hm_launch_cfg *cfg, void *stream)
{
    return guarded([&] {
        if (n > kNoAllocMaxParticles)
            throw Error(HM_ERR_ARG, "track_noalloc: at most 1048576 particles per call (hm_track takes more)");
        TrackLaunch a{};
        int every = 0;
        if (!prepare_track(x, px, y, py, steps, n, table, table_len, n_turns, mu, barrier, kick_module, mode, cfg,
                           "track_noalloc", a, every))
            return;
        hipStream_t s = as_stream(stream);
        const bool oct = mu != 0.0;
        const uint32_t N = (uint32_t)n;
        const TrackFn launch = pick_track(mode, inverse != 0, oct);
        if (every < 0 || (uint32_t)every >= n_turns) {
            a.n_arg = a.grid_bound = N;
            a.n_turns = n_turns;
            launch(a, s);
            return;
        }
        // hm_track's segments, each launched over all n particles: a segment
        // reads its live count on the device and wavefronts past it return at
        // once, so the host never waits for a count
        uint32_t S = 0;
        for (uint32_t done = 0, len = std::min<uint32_t>(kFirstSegment, (uint32_t)every); done < n_turns; ++S) {
            done += std::min(len, n_turns - done);
            len = std::min<uint32_t>(len * 2u, (uint32_t)every);
        }
        if (S > kNoAllocMaxSegments)
            throw Error(HM_ERR_ARG, "track_noalloc: more than 65536 compaction segments (raise compact_every)");
        uint32_t *lists = noalloc_scratch();
        uint32_t *counts = lists + 2 * (size_t)kNoAllocMaxParticles;  // counts[i]: segment i's input length
        // the previous call's counts are still there: clear them with a kernel
        hipLaunchKernelGGL(k_zero_words, dim3((S + 1 + 255) / 256), dim3(256), 0, s, counts, S + 1);
        launch_check("k_zero_words");
        uint32_t len = std::min<uint32_t>(kFirstSegment, (uint32_t)every), done = 0;
        for (uint32_t i = 0; i < S; ++i) {
            const bool last = i + 1 == S;
            a.n_turns = std::min(len, n_turns - done);
            a.idx_in = i == 0 ? nullptr : (i % 2 ? lists : lists + N);
            a.count_in = i == 0 ? nullptr : counts + i;
            a.n_arg = a.grid_bound = N;
            a.idx_out = last ? nullptr : (i % 2 ? lists + N : lists);
            a.count_out = last ? nullptr : counts + i + 1;
            launch(a, s);
            done += a.n_turns;
            len = std::min<uint32_t>(len * 2u, (uint32_t)every);
        }
    });
}

extern "C" int hm_autotune_track(const double *x, const double *px, const double *y, const double *py,
                                 const uint32_t *steps, size_t n, const double *table, size_t table_len,
                                 uint32_t warm_turns, uint32_t trial_turns, double mu, double barrier, int mode,
                                 hm_launch_cfg *best, double *best_ms, void *stream)
{
    return guarded([&] {
        if (!best) throw Error(HM_ERR_ARG, "autotune: NULL result");
        if (mode != HM_MODE_EXACT && mode != HM_MODE_FAST) throw Error(HM_ERR_ARG, "autotune: unknown mode");
        if (!x || !px || !y || !py || !steps || !table) throw Error(HM_ERR_ARG, "autotune: NULL pointer");
        if (n == 0 || n > 0x7fffffffull || trial_turns == 0) throw Error(HM_ERR_ARG, "autotune: empty trial");
        hipStream_t s = as_stream(stream);
        // the trial runs on copies: the caller's particles are not touched
        std::vector<uint32_t> host_steps(n);
        hip_check(hipMemcpyAsync(host_steps.data(), steps, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s),
                  "hipMemcpyAsync(steps)");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        const uint64_t max_steps = *std::max_element(host_steps.begin(), host_steps.end());
        if (max_steps + 2 > table_len) throw Error(HM_ERR_RANGE, "autotune: no turns left in the table");
        const uint64_t room = table_len - max_steps;
        warm_turns = (uint32_t)std::min<uint64_t>(warm_turns, room / 2);
        trial_turns = (uint32_t)std::min<uint64_t>(trial_turns, room - warm_turns);
        const size_t state = n * (4 * sizeof(double) + sizeof(uint32_t));
        char *work = static_cast<char *>(workspace(state, 5));
        char *warm = static_cast<char *>(workspace(state, 6));
        double *w[4];
        for (int c = 0; c < 4; ++c) w[c] = reinterpret_cast<double *>(work + c * n * sizeof(double));
        uint32_t *ws = reinterpret_cast<uint32_t *>(work + 4 * n * sizeof(double));
        const double *src[4] = {x, px, y, py};
        for (int c = 0; c < 4; ++c)
            hip_check(hipMemcpyAsync(w[c], src[c], n * sizeof(double), hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
        hip_check(hipMemcpyAsync(ws, steps, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
        const double nan = std::nan("");
        hm_launch_cfg base{};
        base.particles_per_lane = mode == HM_MODE_EXACT ? 1 : 2;
        if (warm_turns) {
            const int rc = hm_track(w[0], w[1], w[2], w[3], ws, n, table, table_len, warm_turns, mu, barrier, nan, 0,
                                    mode, &base, stream);
            if (rc != HM_OK) throw Error(rc, hm_last_error());
        }
        hip_check(hipMemcpyAsync(warm, work, state, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync(snapshot)");
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hip_check(hipEventCreate(&e0), "hipEventCreate");
        hip_check(hipEventCreate(&e1), "hipEventCreate");
        float best_t = -1.0f;
        hm_launch_cfg best_c = base;
        try {
            for (int block : {64, 256})
                for (int every : {500, 1000, 2000})
                    for (int pair : {1, 2}) {
                        hm_launch_cfg c{block, every, pair, 0};
                        hip_check(hipMemcpyAsync(work, warm, state, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
                        hip_check(hipEventRecord(e0, s), "hipEventRecord");
                        const int rc = hm_track(w[0], w[1], w[2], w[3], ws, n, table, table_len, trial_turns, mu,
                                                barrier, nan, 0, mode, &c, stream);
                        if (rc != HM_OK) throw Error(rc, hm_last_error());
                        hip_check(hipEventRecord(e1, s), "hipEventRecord");
                        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
                        float ms = 0.0f;
                        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
                        if (best_t < 0.0f || ms < best_t) {
                            best_t = ms;
                            best_c = c;
                        }
                    }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *best = best_c;
        if (best_ms) *best_ms = best_t;
        const TuneKey k = tune_key(mode, mu != 0.0, n);
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[k] = best_c;
    });
}

extern "C" int hm_autotune_clear(void)
{
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned.clear();
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
