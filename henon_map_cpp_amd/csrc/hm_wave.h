// Wavefront-level tracking shared by the tracking kernels (single tracker and
// batched jobs): a wavefront's particles advanced for a number of turns with
// the table record in scalar registers, loss kept as lane masks, and the
// survivors appended to the next live list. Replaces the per-turn launches of
// src/henon.cu:636-691 / :3155-3185 and the host loop :2701-2752.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "hm_device.h"

namespace hm {

constexpr uint32_t kFlagRange = 1u;


// Per-lane fallback used when the live lanes of a wavefront do not share one
// turn index (particles reset while others were not) or when that index range
// leaves the table: each lane reads its own table record, and a lane whose
// index would leave [0, tab_len) stops unchanged and raises kFlagRange.
template <int MODE, bool INV, bool OCT>
__device__ __forceinline__ void track_lanes(double &x, double &px, double &y, double &py, bool &alive,
                                            uint32_t &done, uint32_t s, const TabEntry *__restrict__ tab,
                                            uint32_t tab_len, uint32_t n_turns, double mu, double b2, uint32_t *status)
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

// LDS ring of the staged turn loop, per wavefront: 3 chunks of 8 records.
constexpr int kTrackRingWave = 3 * 8 * 4;
constexpr int kTrackRingWaves = 4;  // wavefronts per workgroup (at most 256 threads)

template <int MODE, bool INV, bool OCT, int P, bool RING = false>
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
    // forward reads records s0 .. s0+n_turns-1, inverse s0-1 down to s0-n_turns:
    // both ends must lie inside [0, tab_len), else the per-lane path flags it
    const bool in_table = INV ? (s0 >= n_turns && s0 <= tab_len) : ((uint64_t)s0 + n_turns <= (uint64_t)tab_len);

    if (same && in_table && n_turns > 0) {
        // the table is read-only for the whole launch: reading it through the
        // constant address space lets wave-uniform records come in by scalar
        // loads even when the table pointer itself was loaded from memory
        const ConstPtr<TabEntry> tp = as_const(INV ? tab + (s0 - 1) : tab + s0);
        auto rec = [&](uint32_t k) -> TabEntry {
            const ConstPtr<TabEntry> r = INV ? tp - k : tp + k;
            return load_record(r);
        };
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
        double r2[P];
        auto turn = [&](const TabEntry &t) {
#pragma unroll
            for (int q = 0; q < P; ++q) {
                if (CARRY) {
                    r2[q] = fwd_exact_carry<OCT>(x[q], px[q], y[q], py[q], xx[q], yy[q], mu, t);
                } else {
                    map_step<MODE, INV, OCT>(x[q], px[q], y[q], py[q], mu, m3mu, t);
                    r2[q] = SCREEN ? screen_radius2(x[q], px[q], y[q], py[q]) : radius2<MODE>(x[q], px[q], y[q], py[q]);
                }
            }
        };
        // The common turn costs one compare and one scalar test per particle
        // slot; everything about a loss (the exact confirmation, the loss turn,
        // the live masks, the early exit) sits on the rare path `settle`, which
        // returns false when no particle of the wavefront is left.
        auto lost_mask = [&]() -> uint64_t {
            uint64_t any = 0;
#pragma unroll
            for (int q = 0; q < P; ++q) any |= live[q] & ~__ballot(r2[q] < b2s);
            return any;
        };
        auto settle = [&](uint32_t kk) -> bool {
            uint64_t all = 0;
#pragma unroll
            for (int q = 0; q < P; ++q) {
                uint64_t lost = live[q] & ~__ballot(r2[q] < b2s);
                if (SCREEN && lost) {
                    const double e = CARRY ? exact_radius2(xx[q], px[q], yy[q], py[q])
                                           : exact_radius2(x[q] * x[q], px[q], y[q] * y[q], py[q]);
                    lost &= ~__ballot(e < b2);
                }
                if (lost & my_bit) done[q] = kk;
                live[q] &= ~lost;
                all |= live[q];
            }
            return all != 0;
        };
        // Four turns per iteration in two halves: each half's two records are
        // requested (one 64-byte scalar load, fixed offsets from a pointer that
        // advances by four records) one half ahead, into the register pair the
        // half before last used, so no register copies are needed. The last
        // turns (fewer than four) go one by one. When the last particle is lost,
        // the rare path ends the loop by shortening its count: the common path
        // has no exit test of its own.
        uint32_t k = 0;
        bool any_live = true;
        if constexpr (RING) {
            // With few wavefronts per SIMD nothing hides a scalar load's latency
            // (an L2 miss the first time a table is read): the wavefront stages
            // the records of 8-turn chunks into its own LDS ring two chunks
            // ahead (one vector load and one LDS store per lane and chunk) and
            // reads each turn's record one turn ahead. Lanes 32..63 repeat
            // lanes 0..31 (identical stores).
            __shared__ __align__(16) double ring_all[kTrackRingWaves * kTrackRingWave];
            double *const ring = ring_all + (threadIdx.x >> 6) * kTrackRingWave;
            const double *tabd = reinterpret_cast<const double *>(tab);
            const uint32_t G = (n_turns + 7) / 8;
            auto clampg = [&](uint32_t g) { return g < G ? g : G - 1; };
            auto stage_load = [&](uint32_t g) -> double {
                const uint32_t kk = min(g * 8 + ((lane & 31) >> 2), n_turns - 1);
                const uint32_t ti = INV ? s0 - 1 - kk : s0 + kk;
                return tabd[4 * (size_t)ti + (lane & 3)];
            };
            auto stage_store = [&](uint32_t g, double e) { ring[(g % 3) * 32 + (lane & 31)] = e; };
            auto ring_at = [&](uint32_t g, uint32_t m) -> TabEntry {
                const double *q = ring + (g % 3) * 32 + m * 4;
                return TabEntry{q[0], q[1], q[2], q[3]};
            };
            stage_store(0, stage_load(0));
            stage_store(clampg(1), stage_load(clampg(1)));
            double e = stage_load(clampg(2));
            __builtin_amdgcn_wave_barrier();
            TabEntry r = ring_at(0, 0);
            for (uint32_t g = 0; g < G; ++g) {
                stage_store(clampg(g + 2), e);
                e = stage_load(clampg(g + 3));
                __builtin_amdgcn_wave_barrier();  // the stores precede later chunks' reads
                const uint32_t k0 = g * 8, len = min(8u, n_turns - k0);
                if (len == 8) {
#pragma unroll
                    for (uint32_t m = 0; m < 8; ++m) {
                        const TabEntry rn = m + 1 < 8 ? ring_at(g, m + 1) : ring_at(g + 1, 0);
                        turn(r);
                        if (lost_mask() && !settle(k0 + m)) any_live = false;
                        r = rn;
                    }
                } else {
                    for (uint32_t m = 0; m < len; ++m) {
                        turn(ring_at(g, m));
                        if (lost_mask() && !settle(k0 + m)) any_live = false;
                    }
                }
                if (!any_live) break;
            }
            k = n_turns;
        } else if (n_turns >= 6) {
            ConstPtr<TabEntry> p = tp;  // record of turn k (forward) / k-th back (inverse)
            auto at = [&](int off) -> TabEntry {
                const ConstPtr<TabEntry> r = INV ? p - off : p + off;
                return load_record(r);
            };
            auto one = [&](const TabEntry &t, uint32_t kk, uint32_t &quads) {
                turn(t);
                if (lost_mask() && !settle(kk)) {
                    any_live = false;
                    quads = 1;
                }
            };
            TabEntry a0 = at(0), a1 = at(1);
            // turns k..k+3 use a0, a1, b0, b1; the next quad's a0, a1 (turns k+4,
            // k+5) must exist, hence (n_turns - 2) / 4 quads
            for (uint32_t quads = (n_turns - 2) / 4; quads; --quads) {
                const TabEntry b0 = at(2), b1 = at(3);
                one(a0, k, quads);
                one(a1, k + 1, quads);
                a0 = at(4);
                a1 = at(5);
                one(b0, k + 2, quads);
                one(b1, k + 3, quads);
                k += 4;
                p = INV ? p - 4 : p + 4;
            }
        }
        for (; any_live && k < n_turns; ++k) {
            turn(rec(k));
            if (lost_mask()) any_live = settle(k);
        }
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

}  // namespace hm
