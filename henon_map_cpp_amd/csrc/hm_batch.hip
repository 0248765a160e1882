// Batched tracking of independent jobs (each with its own particles, table,
// turn count, mu and barrier) in one stream of persistent launches fed by a
// device work queue. No reference counterpart: the reference tracks one
// particle set per call (src/henon.cu:2701-2752, :3155-3185).
//
// The turns are cut into segments (125, 250, 500, then `compact_every` turns).
// Per segment, two launches:
//   k_batch_plan   one workgroup: per job, the live count (the job's particle
//                  count in the first segment, the survivors appended by the
//                  previous segment otherwise) -> 64-particle chunks; an
//                  exclusive scan over jobs gives each job's first chunk id; the
//                  queue head and the jobs' next-segment counters are cleared.
//   k_batch_track  a persistent grid of resident wavefronts; each wavefront
//                  pops chunk ids from the queue head (one atomic per chunk),
//                  finds the chunk's job by binary search over the scan, and
//                  tracks those 64 particles for the job's turns of the segment
//                  (the same wave-level code as hm_track), appending survivors
//                  to the job's next live list. A wavefront leaves when the queue
//                  is drained, so every wavefront exits.
// Nothing is read back by the host between segments, so the whole launch
// sequence can be recorded once in a HIP graph and replayed.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "hm_device.h"
#include "hm_internal.h"
#include "hm_tangent.h"
#include "hm_wave.h"

namespace hm {

struct JobDesc {
    double *x, *px, *y, *py;
    uint32_t *steps;
    const TabEntry *tab;
    uint32_t n, n_turns, tab_len, list_off;
    double mu, b2;
    double *megno;    // MEGNO jobs: output [n]; NULL for tracking-only jobs
    double *state;    // MEGNO jobs: [n][kMegnoState] tangent matrix and sums
    uint32_t renorm;  // MEGNO jobs: renormalisation interval (0 = never)
    uint32_t pad;
};

// Per-particle state a MEGNO job carries from one segment to the next: the
// tangent matrix (16), L1, L2 and |M|_F^2 of the last turn.
constexpr int kMegnoState = 19;

struct BatchQueue {
    uint32_t total;   // chunks of the current segment
    uint32_t split;   // chunks of tracking-only jobs come first: [0, split)
    uint32_t head;    // next tracking chunk id to hand out
    uint32_t head_m;  // next MEGNO chunk id to hand out (from split)
};

struct BatchSeg {
    const JobDesc *jobs;
    uint32_t n_jobs;
    uint32_t n_plain;                    // jobs [0, n_plain) track only, the rest accumulate MEGNO
    const double *inv_idx;               // 1.0 / k, k = 1..max turns
    uint32_t seg_start, seg_len;
    uint32_t first;                      // 1 in the first segment (identity lists)
    const uint32_t *lists_cur;           // live lists read by this segment
    uint32_t *lists_next;                // survivors appended by this segment
    const uint32_t *count_cur;           // [n_jobs]
    uint32_t *count_next;                // [n_jobs]
    uint32_t *chunk_off;                 // [n_jobs + 1] exclusive scan of chunk counts
    BatchQueue *queue;
    uint32_t *status;
};

constexpr int kPlanThreads = 1024;

__global__ void __launch_bounds__(kPlanThreads) k_batch_plan(BatchSeg S)
{
    __shared__ uint32_t wave_sum[kPlanThreads / 64];
    __shared__ uint32_t carry;
    const uint32_t t = threadIdx.x, lane = t & 63u, wv = t >> 6;
    if (t == 0) carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < S.n_jobs; base += kPlanThreads) {
        const uint32_t j = base + t;
        uint32_t nch = 0;
        if (j < S.n_jobs) {
            const JobDesc &J = S.jobs[j];
            if (S.seg_start < J.n_turns) {
                const uint32_t live = S.first ? J.n : S.count_cur[j];
                nch = (live + 63u) / 64u;
            }
            S.count_next[j] = 0;
        }
        // inclusive scan within the wavefront, then across wavefronts
        uint32_t v = nch;
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t u = __shfl_up(v, d, 64);
            if ((int)lane >= d) v += u;
        }
        if (lane == 63) wave_sum[wv] = v;
        __syncthreads();
        uint32_t before = carry;
        for (uint32_t w = 0; w < wv; ++w) before += wave_sum[w];
        if (j < S.n_jobs) S.chunk_off[j] = before + v - nch;
        __syncthreads();
        if (t == kPlanThreads - 1) carry = before + v;
        __syncthreads();
    }
    __syncthreads();
    if (t == 0) {
        S.chunk_off[S.n_jobs] = carry;
        const uint32_t split = S.n_plain < S.n_jobs ? S.chunk_off[S.n_plain] : carry;
        S.queue->total = carry;
        S.queue->split = split;
        S.queue->head = 0;
        S.queue->head_m = split;
    }
}

// Identity tangent matrices and zero sums for every particle of the MEGNO jobs.
__global__ void k_megno_init(double *__restrict__ state, uint64_t n_particles)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_particles) return;
    double *st = state + i * kMegnoState;
    for (int q = 0; q < 16; q++) st[q] = (q % 5 == 0) ? 1.0 : 0.0;
    st[16] = 0.0;
    st[17] = 0.0;
    st[18] = 4.0;  // |I|_F^2
}

template <int MODE>
__global__ void __launch_bounds__(256, 8) k_batch_track(BatchSeg S)
{
    const uint32_t lane = __lane_id();
    // written by the plan launch, read-only here: scalar loads
    const ConstPtr<uint32_t> chunk_off = as_const(S.chunk_off);
    const ConstPtr<uint32_t> count_cur = as_const(S.count_cur);
    const ConstPtr<JobDesc> jobs = as_const(S.jobs);
    const uint32_t total = *as_const(&S.queue->split);
    for (;;) {
        uint32_t c = 0;
        if (lane == 0) c = atomicAdd(&S.queue->head, 1u);
        c = __builtin_amdgcn_readfirstlane(c);
        if (c >= total) break;
        // job = the last j with chunk_off[j] <= c (jobs with no chunk are skipped)
        uint32_t lo = 0, hi = S.n_jobs;  // chunk_off[lo] <= c < chunk_off[hi]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (chunk_off[mid] <= c) lo = mid;
            else hi = mid;
        }
        const uint32_t job = lo;
        const ConstPtr<JobDesc> J = jobs + job;
        const uint32_t n = S.first ? J->n : count_cur[job];
        const uint32_t base = (c - chunk_off[job]) * 64u;
        const uint32_t n_turns = J->n_turns, list_off = J->list_off;
        const uint32_t turns = min(S.seg_len, n_turns - S.seg_start);
        const bool last = S.seg_start + S.seg_len >= n_turns;
        const double mu = J->mu;
        TrackArgs A{J->x, J->px, J->y, J->py, J->steps,
                    S.first ? nullptr : S.lists_cur + list_off, nullptr, n, 0u,
                    J->tab, J->tab_len, turns, mu, J->b2,
                    last ? nullptr : S.lists_next + list_off, last ? nullptr : S.count_next + job, S.status};
        if (mu != 0.0) track_wave<MODE, false, true, 1>(A, base, n);
        else track_wave<MODE, false, false, 1>(A, base, n);
    }
}

// One chunk of a MEGNO job for the segment's turns: the fused tangent map and
// MEGNO recursion of k_indicators (same operations in the same order, so the
// same results), with the matrix and sums kept per particle between segments.
template <int MODE, bool OCT>
__device__ __forceinline__ void megno_chunk(const BatchSeg &S, ConstPtr<JobDesc> J, uint32_t job, uint32_t base,
                                            uint32_t n)
{
    const uint32_t lane = __lane_id(), i = base + lane;
    const bool in = i < n;
    const uint32_t list_off = J->list_off, n_turns = J->n_turns;
    const uint32_t turns = min(S.seg_len, n_turns - S.seg_start);
    const bool last = S.seg_start + S.seg_len >= n_turns;
    double *gx = J->x, *gpx = J->px, *gy = J->y, *gpy = J->py;
    uint32_t *gsteps = J->steps;
    uint32_t j = 0, s = 0;
    double x = __builtin_nan(""), px = x, y = x, py = x;
    double m[16], l1 = 0.0, l2 = 0.0, fb2 = 0.0;
    double *st = nullptr;
    if (in) {
        j = S.first ? i : S.lists_cur[list_off + i];
        x = gx[j]; px = gpx[j]; y = gy[j]; py = gpy[j];
        s = gsteps[j];
        st = J->state + (size_t)j * kMegnoState;
#pragma unroll
        for (int q = 0; q < 16; q++) m[q] = st[q];
        l1 = st[16]; l2 = st[17]; fb2 = st[18];
    } else {
#pragma unroll
        for (int q = 0; q < 16; q++) m[q] = 0.0;
    }
    const bool alive0 = in && !any_nan(x, px, y, py);
    // held (some but not all coordinates NaN): not tracked, but its matrix and
    // sums go on at the held point with table index steps + job turn, as in
    // k_indicators; it stays in the live list and off the straight runs
    const bool held = in && !alive0 && !(__builtin_isnan(x) && __builtin_isnan(px) && __builtin_isnan(y) &&
                                         __builtin_isnan(py));
    if (held) s += S.seg_start;
    bool alive = alive0;
    uint32_t done = turns;
    const double mu = J->mu, b2 = J->b2, mu3 = mu * 3, mu6 = mu * 6, m3mu = -3.0 * mu;
    const TabEntry *tab = J->tab;
    const uint32_t tab_len = J->tab_len;
    const uint64_t live0 = __ballot(alive0);
    uint32_t s0 = 0;
    if (live0) s0 = __builtin_amdgcn_readlane(s, __builtin_ctzll(live0));
    const bool uniform = live0 && __all(!alive0 || s == s0) && !__any(held) && (uint64_t)s0 + turns <= (uint64_t)tab_len;
    const bool any_live = __any(alive0 || held);
    const ConstPtr<TabEntry> tp = as_const(tab + s0);
    const ConstPtr<double> inv_idx = as_const(S.inv_idx);
    const uint32_t renorm = J->renorm;
    uint32_t renorm_in = renorm ? renorm - S.seg_start % renorm : 0;
    // as in k_indicators (same operations, so the same values): the MEGNO sums
    // of turn k are formed at the top of turn k+1, or at once on a
    // renormalisation turn
    bool pending = false;
    double w_prev = 0.0, inv_prev = 0.0;
    auto megno_update = [&]() {
        const double fa2 = frob2_fused(m);
        l1 = l1 + (0.5 * log10_fdl(fa2 / fb2)) * w_prev;
        l2 = l2 + inv_prev * l1;
        fb2 = fa2;
        pending = false;
    };
    double xx = x * x, yy = y * y;  // carried between turns in exact mode
    const double b2s = screen_bound(b2);
    for (uint32_t k = 0; k < turns && any_live; ++k) {
        if (uniform && pending) {
            // Plain turns up to the next renormalisation turn, as k_indicators
            // runs them: one straight block per turn (the previous turn's MEGNO
            // sums, the matrix product and the map interleave), the next turn's
            // table record and reciprocal requested a turn ahead. A particle
            // lost in the run keeps its loss turn and becomes NaN when the run
            // ends; its MEGNO is NaN whatever the sums hold.
            uint32_t stop = turns;
            if (renorm_in) stop = min(stop, k + renorm_in - 1);
            if (stop > k) {
                const bool alive_start = alive;
                TabEntry t = load_record(tp + k);
                double inv_k = inv_idx[S.seg_start + k];
                for (uint32_t kk = k; kk < stop; ++kk) {
                    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): this turn's entries
                    __builtin_amdgcn_sched_barrier(0);
                    const uint32_t kn = min(kk + 1, turns - 1);
                    const TabEntry tn = load_record(tp + kn);
                    const double inv_n = inv_idx[S.seg_start + kn];
                    {
                        const double fa2 = frob2_fused(m);
                        l1 = l1 + (0.5 * log10_fdl(fa2 / fb2)) * w_prev;
                        l2 = l2 + inv_prev * l1;
                        fb2 = fa2;
                    }
                    const Jac jc = jac_coeffs<MODE>(x, y, mu3, mu6);
                    mat_mul_structured<MODE>(jc, t, m);
                    const bool out = forward_turn_out<MODE, OCT>(x, px, y, py, xx, yy, mu, m3mu, t, b2, b2s);
                    done = (alive & out) ? kk : done;
                    alive = alive & !out;
                    w_prev = (double)(S.seg_start + kk + 1);
                    inv_prev = inv_k;
                    t = tn;
                    inv_k = inv_n;
                }
                if (alive_start && !alive) {
                    x = px = y = py = xx = yy = __builtin_nan("");
#pragma unroll
                    for (int q = 0; q < 16; q++) m[q] = __builtin_nan("");
                }
                if (renorm_in) renorm_in -= stop - k;
                k = stop;
                if (k == turns || !__any(alive)) break;
            }
        }
        if (pending) megno_update();
        TabEntry t;
        if (uniform) {
            t = load_record(tp + k);
        } else {
            const uint64_t ti = (uint64_t)s + k;
            if (alive && ti >= tab_len) {
                atomicOr(S.status, 1u);
                alive = false;
                done = k;
            }
            t = tab[ti < tab_len ? ti : tab_len - 1];
        }
        const Jac jc = jac_coeffs<MODE>(x, y, mu3, mu6);
        mat_mul_structured<MODE>(jc, t, m);
        if (alive && forward_turn_lost<MODE, OCT>(x, px, y, py, xx, yy, mu, m3mu, t, b2, b2s)) {
            x = px = y = py = xx = yy = __builtin_nan("");
#pragma unroll
            for (int q = 0; q < 16; q++) m[q] = __builtin_nan("");
            alive = false;
            done = k;
        }
        const uint32_t idx = S.seg_start + k + 1;
        pending = true;
        w_prev = (double)idx;
        inv_prev = inv_idx[idx - 1];
        if (renorm_in && --renorm_in == 0) {
            renorm_in = renorm;
            megno_update();
            if (alive) {
                const double fa = sqrt(frob2<MODE>(m));
                const double inv = 1.0 / fa;
#pragma unroll
                for (int q = 0; q < 16; q++) m[q] = m[q] * inv;
                fb2 = frob2_fused(m);
            }
        }
        if (!__any(alive || held)) break;
    }
    if (pending) megno_update();
    if (alive0) {
        gx[j] = x; gpx[j] = px; gy[j] = y; gpy[j] = py;
        gsteps[j] = s + done;
    }
    if (alive0 || held) {
#pragma unroll
        for (int q = 0; q < 16; q++) st[q] = m[q];
        st[16] = l1; st[17] = l2; st[18] = fb2;
    }
    // a particle leaves the live list when lost (or all NaN from the start):
    // its MEGNO is final then; survivors and held particles get theirs after
    // the job's last turn
    const bool keep = alive || held;
    if (in && (!keep || last)) J->megno[j] = keep ? inv_idx[n_turns - 1] * l2 : __builtin_nan("");
    if (!last) {
        const uint64_t mlive = __ballot(keep);
        if (mlive) {
            const int leader = __builtin_ctzll(mlive);
            uint32_t slot = 0;
            if ((int)lane == leader) slot = atomicAdd(S.count_next + job, (uint32_t)__popcll(mlive));
            slot = __builtin_amdgcn_readlane(slot, leader);
            const uint64_t bit = 1ull << lane;
            if (mlive & bit) S.lists_next[list_off + slot + __popcll(mlive & (bit - 1ull))] = j;
        }
    }
}

template <int MODE>
__global__ void __launch_bounds__(256) k_batch_megno(BatchSeg S)
{
    const uint32_t lane = __lane_id();
    const ConstPtr<uint32_t> chunk_off = as_const(S.chunk_off);
    const ConstPtr<uint32_t> count_cur = as_const(S.count_cur);
    const ConstPtr<JobDesc> jobs = as_const(S.jobs);
    const uint32_t total = *as_const(&S.queue->total);
    for (;;) {
        uint32_t c = 0;
        if (lane == 0) c = atomicAdd(&S.queue->head_m, 1u);
        c = __builtin_amdgcn_readfirstlane(c);
        if (c >= total) break;
        uint32_t lo = 0, hi = S.n_jobs;  // chunk_off[lo] <= c < chunk_off[hi]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (chunk_off[mid] <= c) lo = mid;
            else hi = mid;
        }
        const uint32_t job = lo;
        const ConstPtr<JobDesc> J = jobs + job;
        const uint32_t n = S.first ? J->n : count_cur[job];
        const uint32_t base = (c - chunk_off[job]) * 64u;
        if (J->mu != 0.0) megno_chunk<MODE, true>(S, J, job, base, n);
        else megno_chunk<MODE, false>(S, J, job, base, n);
    }
}

static uint32_t batch_resident_waves()
{
    int dev = 0, cus = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    return (uint32_t)cus * 32u;  // 4 SIMDs x 8 wavefronts (the kernel stays within 64 VGPRs)
}

}  // namespace hm

using namespace hm;

static constexpr uint32_t kBatchFirstSegment = 125;

namespace {

// Launch configurations found by hm_autotune_batch, per device, mode and
// job-mix class: floor(log2) of the job count, of the particles in total and of
// the longest job's turns, and the share of MEGNO particles in quarters.
struct BatchKey {
    int dev, mode, jobs_bits, n_bits, t_bits, megno_q;
    bool operator<(const BatchKey &o) const
    {
        return std::tie(dev, mode, jobs_bits, n_bits, t_bits, megno_q) <
               std::tie(o.dev, o.mode, o.jobs_bits, o.n_bits, o.t_bits, o.megno_q);
    }
};
std::mutex g_batch_tune_mu;
std::map<BatchKey, hm_batch_cfg> g_batch_tuned;

int floor_log2(uint64_t v)
{
    int b = 0;
    while ((uint64_t(2) << b) <= v) ++b;
    return b;
}

BatchKey batch_key(const hm_job *jobs, size_t n_jobs, int mode)
{
    uint64_t total = 0, megno = 0;
    uint32_t max_t = 0;
    for (size_t k = 0; k < n_jobs; ++k) {
        total += jobs[k].n;
        if (jobs[k].megno) megno += jobs[k].n;
        if (jobs[k].n) max_t = std::max(max_t, jobs[k].n_turns);
    }
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    const int q = total ? (int)std::min<uint64_t>(3, (4 * megno) / total) : 0;
    return BatchKey{dev, mode, floor_log2(n_jobs), floor_log2(std::max<uint64_t>(total, 1)),
                    floor_log2(std::max<uint32_t>(max_t, 1)), q};
}

// Instantiated graphs of earlier calls, keyed by everything the recorded launch
// sequence depends on (the job descriptors are read from device memory, so
// particle and table pointers, mu and barriers may change between replays).
struct GraphKey {
    int dev, mode, block;
    uint32_t grid, every, max_turns, n_plain;
    size_t n_jobs;
    uint64_t total_n, megno_n;
    const void *ws, *mstate, *inv_idx;
    bool operator<(const GraphKey &o) const
    {
        return std::tie(dev, mode, block, grid, every, max_turns, n_plain, n_jobs, total_n, megno_n, ws, mstate,
                        inv_idx) < std::tie(o.dev, o.mode, o.block, o.grid, o.every, o.max_turns, o.n_plain,
                                            o.n_jobs, o.total_n, o.megno_n, o.ws, o.mstate, o.inv_idx);
    }
};
constexpr size_t kGraphCache = 8;
std::mutex g_graph_mu;
std::vector<std::pair<GraphKey, hipGraphExec_t>> g_graphs;  // oldest first

// The batched launch sequence for a resolved configuration.
void batch_run(const hm_job *jobs, size_t n_jobs, int mode, int block, int every, uint32_t waves, bool graph,
               hipStream_t s)
{
    // tracking-only jobs first, then the MEGNO jobs (each kernel takes one range)
    std::vector<size_t> order;
    order.reserve(n_jobs);
    for (size_t k = 0; k < n_jobs; ++k)
        if (!jobs[k].megno) order.push_back(k);
    const uint32_t n_plain = (uint32_t)order.size();
    for (size_t k = 0; k < n_jobs; ++k)
        if (jobs[k].megno) order.push_back(k);
    std::vector<JobDesc> desc(n_jobs);
    uint64_t total_n = 0, megno_n = 0;
    uint32_t max_turns = 0;
    for (size_t r = 0; r < n_jobs; ++r) {
        const size_t k = order[r];
        const hm_job &j = jobs[k];
        if (j.n > 0x7fffffffull) throw Error(HM_ERR_ARG, "batch: a job exceeds 2^31 particles");
        if (j.n && (!j.x || !j.px || !j.y || !j.py || !j.steps || !j.table))
            throw Error(HM_ERR_ARG, "batch: NULL pointer in job " + std::to_string(k));
        if (j.table_len > 0xffffffffull) throw Error(HM_ERR_ARG, "batch: table exceeds 2^32 turns");
        if (!std::isnan(j.kick_module)) throw Error(HM_ERR_UNSUPPORTED, "batch: the random kick is not available");
        JobDesc &d = desc[r];
        d.x = j.x; d.px = j.px; d.y = j.y; d.py = j.py; d.steps = j.steps;
        d.megno = j.megno;
        d.renorm = j.renorm_every;
        d.state = nullptr;
        if (j.megno) {
            d.state = reinterpret_cast<double *>((uintptr_t)(megno_n * kMegnoState * sizeof(double)));  // offset
            megno_n += j.n;
        }
        d.tab = reinterpret_cast<const TabEntry *>(j.table);
        d.n = (uint32_t)j.n;
        d.n_turns = j.n ? j.n_turns : 0;
        d.tab_len = (uint32_t)j.table_len;
        d.list_off = (uint32_t)total_n;
        d.mu = j.mu;
        d.b2 = j.barrier * j.barrier;  // henon.cu:2717
        total_n += j.n;
        if (total_n > 0xffffffffull) throw Error(HM_ERR_ARG, "batch: more than 2^32 particles in total");
        max_turns = std::max(max_turns, d.n_turns);
    }
    if (max_turns == 0) return;

    // segment schedule
    std::vector<std::pair<uint32_t, uint32_t>> seg;  // (start, length)
    for (uint64_t done = 0, len = std::min<uint32_t>(kBatchFirstSegment, (uint32_t)every); done < max_turns;) {
        seg.emplace_back((uint32_t)done, (uint32_t)len);
        done += len;
        len = std::min<uint64_t>(len * 2, (uint64_t)every);
    }

    // device buffers (workspace slot 4): descriptors | lists x2 | counts x2 | scan | queue
    auto align = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t b_desc = align(sizeof(JobDesc) * n_jobs);
    const size_t b_list = align(sizeof(uint32_t) * std::max<uint64_t>(total_n, 1));
    const size_t b_cnt = align(sizeof(uint32_t) * n_jobs);
    const size_t b_off = align(sizeof(uint32_t) * (n_jobs + 1));
    const size_t b_q = align(sizeof(BatchQueue));
    char *ws = static_cast<char *>(workspace(b_desc + 2 * b_list + 2 * b_cnt + b_off + b_q, 4));
    double *mstate = megno_n ? static_cast<double *>(workspace(megno_n * kMegnoState * sizeof(double), 9))
                             : nullptr;
    for (auto &d : desc)
        if (d.megno) d.state = reinterpret_cast<double *>(reinterpret_cast<char *>(mstate) +
                                                          reinterpret_cast<uintptr_t>(d.state));
    const double *inv_idx = megno_n ? reciprocal_table(max_turns, s) : nullptr;
    JobDesc *d_desc = reinterpret_cast<JobDesc *>(ws);
    uint32_t *lists[2] = {reinterpret_cast<uint32_t *>(ws + b_desc),
                          reinterpret_cast<uint32_t *>(ws + b_desc + b_list)};
    uint32_t *counts[2] = {reinterpret_cast<uint32_t *>(ws + b_desc + 2 * b_list),
                           reinterpret_cast<uint32_t *>(ws + b_desc + 2 * b_list + b_cnt)};
    uint32_t *d_off = reinterpret_cast<uint32_t *>(ws + b_desc + 2 * b_list + 2 * b_cnt);
    BatchQueue *d_q = reinterpret_cast<BatchQueue *>(ws + b_desc + 2 * b_list + 2 * b_cnt + b_off);

    hip_check(hipMemcpyAsync(d_desc, desc.data(), sizeof(JobDesc) * n_jobs, hipMemcpyHostToDevice, s),
              "hipMemcpyAsync(jobs)");
    // the descriptors are read from pageable memory: make sure the copy has
    // consumed them before `desc` goes out of scope
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");

    waves = std::max<uint32_t>(waves, (uint32_t)block / 64u);
    const uint32_t grid = (waves * 64u + (uint32_t)block - 1) / (uint32_t)block;
    uint32_t *status = status_word();
    auto enqueue = [&](hipStream_t q) {
        if (megno_n) {
            hipLaunchKernelGGL(k_megno_init, dim3((uint32_t)((megno_n + 255) / 256)), dim3(256), 0, q, mstate,
                               (uint64_t)megno_n);
            launch_check("k_megno_init");
        }
        for (size_t i = 0; i < seg.size(); ++i) {
            BatchSeg S{d_desc, (uint32_t)n_jobs, n_plain, inv_idx, seg[i].first, seg[i].second,
                       i == 0 ? 1u : 0u, lists[i % 2], lists[(i + 1) % 2], counts[i % 2], counts[(i + 1) % 2],
                       d_off, d_q, status};
            hipLaunchKernelGGL(k_batch_plan, dim3(1), dim3(kPlanThreads), 0, q, S);
            launch_check("k_batch_plan");
            if (megno_n) {
                if (mode == HM_MODE_EXACT)
                    hipLaunchKernelGGL(k_batch_megno<kExact>, dim3(grid), dim3(block), 0, q, S);
                else
                    hipLaunchKernelGGL(k_batch_megno<kFast>, dim3(grid), dim3(block), 0, q, S);
                launch_check("k_batch_megno");
            }
            if (mode == HM_MODE_EXACT)
                hipLaunchKernelGGL(k_batch_track<kExact>, dim3(grid), dim3(block), 0, q, S);
            else
                hipLaunchKernelGGL(k_batch_track<kFast>, dim3(grid), dim3(block), 0, q, S);
            launch_check("k_batch_track");
        }
    };
    if (!graph) {
        LaunchTimer timer("k_batch_track", s);
        enqueue(s);
        return;
    }
    // record the whole segment sequence once per job shape and replay it
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    const GraphKey key{dev, mode, block, grid, (uint32_t)every, max_turns, n_plain, n_jobs, total_n, megno_n,
                       ws, mstate, inv_idx};
    std::lock_guard<std::mutex> lk(g_graph_mu);
    hipGraphExec_t ge = nullptr;
    for (auto &e : g_graphs)
        if (!(e.first < key) && !(key < e.first)) ge = e.second;
    if (!ge) {
        hipStream_t cs = nullptr;
        hipGraph_t g = nullptr;
        hip_check(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreate");
        try {
            hip_check(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            enqueue(cs);
            hip_check(hipStreamEndCapture(cs, &g), "hipStreamEndCapture");
            hip_check(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0), "hipGraphInstantiate");
        } catch (...) {
            if (ge) (void)hipGraphExecDestroy(ge);
            if (g) (void)hipGraphDestroy(g);
            (void)hipStreamDestroy(cs);
            throw;
        }
        (void)hipGraphDestroy(g);
        (void)hipStreamDestroy(cs);
        if (g_graphs.size() == kGraphCache) {
            (void)hipGraphExecDestroy(g_graphs.front().second);
            g_graphs.erase(g_graphs.begin());
        }
        g_graphs.emplace_back(key, ge);
    }
    {
        LaunchTimer timer("k_batch_track", s);
        hip_check(hipGraphLaunch(ge, s), "hipGraphLaunch");
    }
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

struct BatchCfgResolved {
    int block, every;
    uint32_t waves;
    bool graph;
};

BatchCfgResolved resolve(const hm_batch_cfg *cfg)
{
    BatchCfgResolved r;
    r.every = cfg && cfg->compact_every ? cfg->compact_every : 1000;
    if (r.every <= 0) throw Error(HM_ERR_ARG, "batch: compact_every must be positive");
    r.block = cfg && cfg->block ? cfg->block : 256;
    if (r.block < 64 || r.block > 256 || r.block % 64) throw Error(HM_ERR_ARG, "batch: block must be 64, 128, 192 or 256");
    r.waves = cfg && cfg->waves > 0 ? (uint32_t)cfg->waves : batch_resident_waves();
    r.graph = cfg && cfg->use_graph;
    return r;
}

}  // namespace

extern "C" int hm_track_batch(const hm_job *jobs, size_t n_jobs, int mode, const hm_batch_cfg *cfg, void *stream)
{
    return guarded([&] {
        if (mode != HM_MODE_EXACT && mode != HM_MODE_FAST) throw Error(HM_ERR_ARG, "batch: unknown mode");
        if (n_jobs == 0) return;
        if (!jobs) throw Error(HM_ERR_ARG, "batch: NULL job array");
        if (n_jobs > (1u << 24)) throw Error(HM_ERR_ARG, "batch: too many jobs");
        hm_batch_cfg tuned{};
        if (!cfg) {
            const BatchKey k = batch_key(jobs, n_jobs, mode);
            std::lock_guard<std::mutex> lk(g_batch_tune_mu);
            auto it = g_batch_tuned.find(k);
            if (it != g_batch_tuned.end()) {
                tuned = it->second;
                cfg = &tuned;
            }
        }
        const BatchCfgResolved r = resolve(cfg);
        batch_run(jobs, n_jobs, mode, r.block, r.every, r.waves, r.graph, as_stream(stream));
    });
}

extern "C" int hm_autotune_batch(const hm_job *jobs, size_t n_jobs, int mode, uint32_t trial_turns,
                                 hm_batch_cfg *best, double *best_ms, void *stream)
{
    return guarded([&] {
        if (!best) throw Error(HM_ERR_ARG, "autotune_batch: NULL result");
        if (mode != HM_MODE_EXACT && mode != HM_MODE_FAST) throw Error(HM_ERR_ARG, "autotune_batch: unknown mode");
        if (n_jobs == 0 || !jobs || trial_turns == 0) throw Error(HM_ERR_ARG, "autotune_batch: empty trial");
        if (n_jobs > (1u << 24)) throw Error(HM_ERR_ARG, "autotune_batch: too many jobs");
        hipStream_t s = as_stream(stream);
        // trial copies of every job's particles (and MEGNO outputs), each job
        // cut to at most trial_turns turns; the caller's arrays are only read
        std::vector<hm_job> trial(jobs, jobs + n_jobs);
        std::vector<void *> bufs;
        auto release = [&] {
            for (void *b : bufs) (void)hipFree(b);
        };
        try {
            for (size_t k = 0; k < n_jobs; ++k) {
                hm_job &t = trial[k];
                const size_t n = jobs[k].n;
                t.n_turns = std::min(jobs[k].n_turns, trial_turns);
                if (!n) continue;
                void *b = nullptr;
                hip_check(hipMalloc(&b, n * (4 * sizeof(double) + sizeof(uint32_t)) + (jobs[k].megno ? 8 * n + 8 : 0)),
                          "hipMalloc(autotune_batch)");
                bufs.push_back(b);
                double *d = static_cast<double *>(b);
                t.x = d; t.px = d + n; t.y = d + 2 * n; t.py = d + 3 * n;
                t.steps = reinterpret_cast<uint32_t *>(d + 4 * n);
                if (jobs[k].megno) t.megno = reinterpret_cast<double *>(t.steps + n + (n & 1));
            }
            auto restore = [&] {
                for (size_t k = 0; k < n_jobs; ++k) {
                    const size_t n = jobs[k].n;
                    if (!n) continue;
                    hip_check(hipMemcpyAsync(trial[k].x, jobs[k].x, n * 8, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
                    hip_check(hipMemcpyAsync(trial[k].px, jobs[k].px, n * 8, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
                    hip_check(hipMemcpyAsync(trial[k].y, jobs[k].y, n * 8, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
                    hip_check(hipMemcpyAsync(trial[k].py, jobs[k].py, n * 8, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
                    hip_check(hipMemcpyAsync(trial[k].steps, jobs[k].steps, n * 4, hipMemcpyDeviceToDevice, s),
                              "hipMemcpyAsync");
                }
            };
            hipEvent_t e0 = nullptr, e1 = nullptr;
            hip_check(hipEventCreate(&e0), "hipEventCreate");
            hip_check(hipEventCreate(&e1), "hipEventCreate");
            const uint32_t resident = batch_resident_waves();
            double best_t = 0.0;
            hm_batch_cfg best_c{};
            bool have = false;
            // first the defaults untimed (tables into L2, code objects loaded)
            restore();
            batch_run(trial.data(), n_jobs, mode, 256, 1000, resident, false, s);
            for (int block : {64, 256})
                for (int every : {250, 1000, 4000})
                    for (uint32_t waves : {resident, resident / 2}) {
                        restore();
                        hip_check(hipEventRecord(e0, s), "hipEventRecord");
                        batch_run(trial.data(), n_jobs, mode, block, every, waves, false, s);
                        hip_check(hipEventRecord(e1, s), "hipEventRecord");
                        hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
                        float ms = 0.0f;
                        hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
                        if (!have || ms < best_t) {
                            have = true;
                            best_t = ms;
                            best_c = hm_batch_cfg{block, every, (int32_t)waves, 0};
                        }
                    }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            *best = best_c;
            if (best_ms) *best_ms = best_t;
            const BatchKey key = batch_key(jobs, n_jobs, mode);
            std::lock_guard<std::mutex> lk(g_batch_tune_mu);
            g_batch_tuned[key] = best_c;
        } catch (...) {
            (void)hipStreamSynchronize(s);
            release();
            throw;
        }
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        release();
    });
}

extern "C" int hm_autotune_batch_clear(void)
{
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_batch_tune_mu);
        g_batch_tuned.clear();
    });
}
