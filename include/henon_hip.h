/*
 * henon_hip.h -- C ABI of the MI355X (gfx950) 4D modulated Henon-map tracker.
 *
 * Plain C: pointers, sizes and status codes; no C++ or torch types. Every entry
 * point returns HM_OK (0) or an HM_ERR_* code; hm_last_error() gives the
 * message of the calling thread's last failure. Device pointers are HIP device
 * allocations (hm_malloc, or any hipMalloc'd / torch CUDA tensor memory); a
 * `stream` argument is a hipStream_t passed as void* (NULL = the null stream).
 * Kernel launches are asynchronous with respect to the host unless stated.
 * The library keeps per-device scratch buffers (live lists, staging, tables);
 * calls that overlap in time on one device, from several host threads or
 * streams, share them and must be serialised by the caller.
 *
 * Layouts in HBM
 *   particles : structure of arrays, x, px, y, py (double[n]) and steps (uint32[n])
 *   table     : one 32-byte record per turn, {sin wx(n), cos wx(n), sin wy(n), cos wy(n)}
 *               (double[turns][4]); built on the host by hm_build_table
 *   matrices  : double[n][16], row-major 4x4 per particle (the reference's layout)
 *
 * Each declaration names the reference interface it replaces (paths relative to
 * the reference checkout).
 */
#ifndef HENON_HIP_H_
#define HENON_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status ------------------------------------------------------------ */
enum {
    HM_OK = 0,
    HM_ERR_ARG = 1,          /* invalid argument (ValueError in Python) */
    HM_ERR_RANGE = 2,        /* turn count / table range violated (RuntimeError) */
    HM_ERR_HIP = 3,          /* HIP runtime error */
    HM_ERR_UNSUPPORTED = 4,  /* feature not available in this build */
    HM_ERR_ALLOC = 5         /* device or host allocation failed */
};

/* HM_MODE_EXACT: bit-exact with the reference CPU path (no fused multiply-adds).
 * HM_MODE_FAST: fused multiply-adds everywhere.
 * HM_MODE_EXACT_ORBIT (hm_track_indicators only; no reference counterpart):
 *   the map and the loss test of HM_MODE_EXACT (trajectory, losses and Birkhoff
 *   tunes bit-exact), the tangent products and norms of HM_MODE_FAST (MEGNO and
 *   FLI to rounding). */
enum { HM_MODE_EXACT = 0, HM_MODE_FAST = 1, HM_MODE_EXACT_ORBIT = 2 };

/* Last error message of the calling thread (never NULL).
 * Replaces check_cuda_errors' print-and-exit(1) (src/henon.cu:17-25). */
const char *hm_last_error(void);

/* Library version string. */
const char *hm_version(void);

/* ---- device and memory plumbing ---------------------------------------- */
/* Number of visible HIP devices (0 when none). Replaces
 * is_cuda_device_available (src/bind.cpp:28-39). */
int hm_device_count(int *count);
/* Free and total device memory of the current device, in bytes (no reference
 * counterpart: the engine sizes its trajectory chunks from it). */
int hm_mem_info(size_t *free_bytes, size_t *total_bytes);
int hm_malloc(void **ptr, size_t bytes);
int hm_free(void *ptr);
/* kind: 0 host->device, 1 device->host, 2 device->device. Asynchronous on
 * `stream`; host buffers must stay valid until the stream is synchronised. */
int hm_memcpy(void *dst, const void *src, size_t bytes, int kind, void *stream);
int hm_memset(void *dst, int value, size_t bytes, void *stream);
int hm_stream_synchronize(void *stream);

/* Read and clear the device-side status word: bit 0 set when a kernel met a
 * table index outside [0, table_len) and froze that particle instead of
 * reading out of bounds. Synchronises `stream`. */
int hm_take_status(uint32_t *flags, void *stream);

/* ---- modulation tables (host, glibc libm) ------------------------------- */
/* src/modulation.cpp:3-12 (basic_modulation) */
int hm_basic_modulation(double tune, double omega, double epsilon, int start, int end, double *out);
/* src/modulation.cpp:30-60 (sps_modulation) */
int hm_sps_modulation(double tune, double epsilon, int start, int end, double *out);
/* The sps loop with caller-supplied line coefficients and angular frequencies:
 * out[i] = 2 pi tune (1 + epsilon sum_j coef[j] cos(freq[j] (i+start))).
 * With the SPS lines it reproduces hm_sps_modulation bit for bit. New kind
 * ("custom"), no reference counterpart. */
int hm_lines_modulation(double tune, double epsilon, const double *coef, const double *freq, int n_lines,
                        int start, int end, double *out);
/* The SPS line set (src/modulation.cpp:14-28): 7 coefficients and frequencies. */
int hm_sps_lines(double *coef, double *freq);
/* src/modulation.cpp:62-96 (gaussian_modulation, uniform_modulation). The
 * reference seeds from std::random_device; here the seed is explicit. */
int hm_gaussian_modulation(double tune, double sigma, int start, int end, uint64_t seed, double *out);
int hm_uniform_modulation(double from, double to, int start, int end, uint64_t seed, double *out);
/* src/henon.cu:2664-2692 (sin/cos tables of compute_a_modulation): writes the
 * interleaved table[n][4] from omega_x[n], omega_y[n] (host memory). */
int hm_build_table(const double *omega_x, const double *omega_y, size_t n, double *table);
/* src/dynamic_indicator.cpp:63-79 (birkhoff_weights), host. */
int hm_birkhoff_weights(uint32_t n_weights, double *out);

/* ---- tracking ------------------------------------------------------------ */
/* Launch configuration; zero fields take the tuned defaults. */
typedef struct hm_launch_cfg {
    int32_t block;              /* threads per workgroup: 64, 128, 192 or 256 (0 = 256) */
    int32_t compact_every;      /* turns between loss compactions (0 = 1000, <0 = never); segments
                                   start at 125 turns and double up to this */
    int32_t particles_per_lane; /* 1 = one per lane; 2 = one, or two for the leading wavefronts
                                   when the live particles exceed one round of resident
                                   wavefronts; 0 = 1 in exact mode, 2 in fast mode */
    int32_t table_staging;      /* how a wavefront reads the turn's table record: HM_STAGING_SCALAR
                                   (scalar loads one half-quad ahead), HM_STAGING_RING (8-turn
                                   chunks staged through a per-wavefront LDS ring), or
                                   HM_STAGING_AUTO (0): the ring when a launch has at most 2
                                   wavefronts per SIMD, where nothing else hides a scalar load's
                                   latency */
} hm_launch_cfg;
enum { HM_STAGING_AUTO = 0, HM_STAGING_SCALAR = 1, HM_STAGING_RING = 2 };

/* Track n particles for n_turns turns of the modulated map.
 * Replaces henon_tracker_gpu::track (src/henon.cu:3155-3185) + gpu_henon_track
 * (:636-691) and the CPU henon_tracker::track (:2701-2752): one persistent launch
 * per compaction segment instead of one launch per turn.
 *   x, px, y, py, steps : device SoA, updated in place (steps = survival turns)
 *   table, table_len    : device table[table_len][4]
 *   mu                  : octupole strength (the octupole kick is skipped when 0)
 *   barrier             : loss radius; lost when x^2+px^2+y^2+py^2 >= barrier^2,
 *                         then all four coordinates become NaN
 *   kick_module         : must be NaN (for the random kick, hm_track_kicked)
 *   inverse             : 0 forward (table index steps), 1 inverse (steps-1)
 *   mode                : HM_MODE_EXACT (bit-exact with the reference CPU path)
 *                         or HM_MODE_FAST (fused multiply-adds)
 *   cfg                 : NULL for the configuration hm_autotune_track chose for this
 *                         device, mode and size class, else the defaults
 * The caller owns the global turn counter and its range check (the Python/C++
 * layer raises "The number of turns is too large."). */
int hm_track(double *x, double *px, double *y, double *py, uint32_t *steps, size_t n,
             const double *table, size_t table_len, uint32_t n_turns, double mu, double barrier,
             double kick_module, int inverse, int mode, const hm_launch_cfg *cfg, void *stream);

/* hm_track that allocates nothing. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc). Same arguments, results
 * (bit for bit) and status bit as hm_track; it calls no hipMalloc, hipFree or
 * other allocating routine and never waits on the host. Its live-index lists
 * and segment counts sit in a static device array of 8.3 MiB that each device
 * holds from library load until the process ends, which bounds n to
 * 1048576 particles and the compaction segments of one call to 65536
 * (n_turns / compact_every); larger calls return HM_ERR_ARG. Every call on a
 * device shares that device's array: calls that overlap in time on one device
 * (several streams or host threads) race and must be serialised by the
 * caller. With hm_profile_enable(1) launches are bracketed by HIP events. */
int hm_track_noalloc(double *x, double *px, double *y, double *py, uint32_t *steps, size_t n,
                     const double *table, size_t table_len, uint32_t n_turns, double mu, double barrier,
                     double kick_module, int inverse, int mode, const hm_launch_cfg *cfg, void *stream);

/* hm_track with the random 4D kick of src/henon.cu:362-405, 420-542: every
 * turn, before the forward map (after the inverse map), each particle moves by
 * kick_module (finite) along a direction drawn from four standard normals. The
 * reference draws them from clock-seeded cuRAND states (not reproducible); here
 * particle i's k-th turn of the call uses Philox-4x32-10 keyed by `seed` at draw
 * index draw + k (Box-Muller normals), so a seed and draw index fix the kicks,
 * and a caller continues the stream by advancing `draw` by n_turns per call.
 * Same loss rule, NaN semantics and status bit 0 (turn index outside the table)
 * as hm_track. Replaces the kick branch of henon_step / gpu_henon_track
 * (src/henon.cu:420-542, 636-691). */
int hm_track_kicked(double *x, double *px, double *y, double *py, uint32_t *steps, size_t n, const double *table,
                    size_t table_len, uint32_t n_turns, double mu, double barrier, double kick_module, int inverse,
                    int mode, uint64_t seed, uint64_t draw, void *stream);

/* Autotuner (no reference counterpart). Times hm_track launch configurations
 * (block 64 / 256, compact_every 500 / 1000 / 2000, one or two particles per
 * lane) on a copy of the given particles: first warm_turns turns with the
 * defaults (so that the fast early losses are behind), then trial_turns turns
 * from that state per candidate. Returns the fastest configuration in *best
 * (its trial time in *best_ms) and remembers it for this device, mode, mu != 0
 * and particle-count class (floor(log2 n)): hm_track with cfg == NULL uses the
 * remembered configuration. The caller's particles are not modified; every
 * configuration gives bit-identical results. Synchronises `stream`. */
int hm_autotune_track(const double *x, const double *px, const double *y, const double *py,
                      const uint32_t *steps, size_t n, const double *table, size_t table_len, uint32_t warm_turns,
                      uint32_t trial_turns, double mu, double barrier, int mode, hm_launch_cfg *best,
                      double *best_ms, void *stream);
/* Forget every remembered configuration (hm_track with cfg == NULL returns to
 * the defaults). */
int hm_autotune_clear(void);

/* ---- batched jobs ---------------------------------------------------------- */
/* One independent tracking job: its own device particles, table, turn count,
 * mu and barrier (the arguments of one hm_track call, forward direction). */
typedef struct hm_job {
    double *x, *px, *y, *py;    /* device SoA, updated in place */
    uint32_t *steps;            /* device; table index and survival turns */
    size_t n;                   /* particles */
    const double *table;        /* device table[table_len][4] */
    size_t table_len;
    uint32_t n_turns;
    uint32_t renorm_every;      /* MEGNO jobs: renormalise the tangent matrix every this many
                                   turns (0 = never) */
    double mu, barrier;
    double kick_module;         /* must be NaN */
    double *megno;              /* NULL: tracking only. Else device double[n] receiving the
                                   MEGNO (construct form, as hm_track_indicators) after
                                   n_turns turns, NaN for lost particles; the tangent matrix
                                   starts from the identity */
} hm_job;

typedef struct hm_batch_cfg {
    int32_t block;         /* threads per workgroup: 64..256 (0 = 256) */
    int32_t compact_every; /* turns per segment (0 = 1000); segments start at 125 turns */
    int32_t waves;         /* persistent wavefronts per launch (0 = CUs x 32, one resident round) */
    int32_t use_graph;     /* 1: record the launch sequence in a HIP graph, replay it and wait */
} hm_batch_cfg;

/* Track every job for its own n_turns, all jobs in one sequence of persistent
 * launches fed by a device work queue of 64-particle chunks, with loss
 * compaction per job between segments. Results are identical to one hm_track
 * call per job with the same mode (MEGNO jobs: to hm_track_indicators with the
 * same renormalisation and one sample at n_turns). No counterpart in the reference, which tracks
 * one particle set per call (henon_tracker::track, src/henon.cu:2701-2752,
 * :3155-3185). jobs: host array. Asynchronous on `stream` unless use_graph. */
int hm_track_batch(const hm_job *jobs, size_t n_jobs, int mode, const hm_batch_cfg *cfg, void *stream);
/* With cfg == NULL, hm_track_batch uses the configuration hm_autotune_batch
 * chose for this device, mode and job-mix class (floor(log2) of the job count,
 * of the particles in total and of the longest job's turns, and the share of
 * MEGNO particles in quarters), else the defaults. With use_graph, the recorded
 * launch sequence is kept (up to 8 job shapes per process) and replayed by later
 * calls with the same job shapes; the descriptors (particles, tables, mu,
 * barriers) are re-read from device memory on every call. */

/* Batch autotuner (no reference counterpart). Times hm_track_batch launch
 * configurations (block 64 / 256, compact_every 250 / 1000 / 4000, one or half
 * a round of resident wavefronts) on copies of the jobs' particles with every
 * job cut to at most trial_turns turns, returns the fastest in *best (its time
 * in *best_ms) and remembers it for the job-mix class. The caller's arrays are
 * not modified; every configuration gives bit-identical results. Allocates the
 * copies (device memory) for the duration of the call. Synchronises `stream`. */
int hm_autotune_batch(const hm_job *jobs, size_t n_jobs, int mode, uint32_t trial_turns, hm_batch_cfg *best,
                      double *best_ms, void *stream);
/* Forget every remembered batch configuration. */
int hm_autotune_batch_clear(void);

/* ---- readback observables ------------------------------------------------ */
enum {
    HM_OBS_RADIUS = 0,   /* src/henon.cu:67-74  sqrt(x^2+px^2+y^2+py^2) */
    HM_OBS_ACTION = 1,   /* :76-83   x^2+px^2+y^2+py^2 */
    HM_OBS_ACTION_X = 2, /* :86-93   x^2+px^2 */
    HM_OBS_ACTION_Y = 3, /* :95-102  y^2+py^2 */
    HM_OBS_ANGLE_X = 4,  /* :104-111 atan2(px, x) */
    HM_OBS_ANGLE_Y = 5   /* :104-111 atan2(py, y) */
};
/* Replaces gpu_evaluate_radius/_action/_action_x/_action_y/gpu_evaluate_angle
 * (src/henon.cu:67-111). out: device double[n]. */
int hm_observable(int kind, const double *x, const double *px, const double *y, const double *py, size_t n,
                  double *out, void *stream);

/* ---- shadow (ghost) particles --------------------------------------------
 * Layout of the particle arrays with ghosts (src/henon.cu:737-817): n_base base
 * particles, then `sets` ghost sets of n_base entries each; entry
 * n_base * (1 + g) + b is ghost g of base particle b. The ghosts themselves are
 * built on the host (henon_map_engine's particles_4d.add_ghost). */

/* Rescales every ghost's displacement from its base to length target:
 * g <- b + (g - b) * (target / |g - b|). Bases are untouched.
 * Replaces gpu_particle_renormalize / particles_4d_gpu::renormalize
 * (src/henon.cu:1257-1282; CPU: 826-858). */
int hm_ghost_renormalize(double *x, double *px, double *y, double *py, size_t n_base, size_t sets, double target,
                         void *stream);
/* Displacement of every ghost from its base: module_out (device double
 * [n_base][sets], or NULL) = |g - b|; direction_out (device double
 * [4][n_base][sets], or NULL) = (b + (g - b) * (1 / |g - b|)) - b per
 * coordinate (x, px, y, py). NaN where the ghost or its base holds a NaN.
 * Replaces gpu_particle_displacements / gpu_particle_directions and the
 * get_displacement_module / get_displacement_direction accessors
 * (src/henon.cu:1284-1442; CPU: 860-952). */
int hm_ghost_displacement(const double *x, const double *px, const double *y, const double *py, size_t n_base,
                          size_t sets, double *module_out, double *direction_out, void *stream);

/* Alignment indices of the deviation vectors d_k = (g_k - b) / |g_k - b| of
 * the first ghost sets (henon_map_cpp/dynamic_indicators.py:154-366, which
 * computes them in numpy from shadow particles it tracks alongside):
 * HM_ALIGN_SALI (needs sets >= 2): min(|d_0 + d_1|, |d_0 - d_1|);
 * HM_ALIGN_GALI (needs sets >= 4): |det [d_0 d_1 d_2 d_3]| (the product of the
 * singular values). out: device double[n_base], the index at the current
 * state (the running minimum over time is the caller's); NaN where a
 * deviation holds a NaN. */
enum { HM_ALIGN_SALI = 0, HM_ALIGN_GALI = 1 };
int hm_ghost_alignment(const double *x, const double *px, const double *y, const double *py, size_t n_base,
                       size_t sets, int kind, double *out, void *stream);

/* Tangent matrix of one turn at the current points, table index table_index:
 * J (reverse=0) or the analytic inverse J^-1 (reverse=1); particles with
 * valid[i]==0 get a NaN matrix. out: device double[n][16].
 * Replaces henon_tracker_gpu::get_tangent_matrix (src/henon.cu:3187-3219),
 * which computed it on the host. */
int hm_tangent_matrix(const double *x, const double *px, const double *y, const double *py,
                      const uint8_t *valid, size_t n, const double *table, size_t table_len,
                      uint32_t table_index, double mu, int reverse, double *out, void *stream);

/* ---- tangent map, MEGNO, FLI, Birkhoff ----------------------------------- */
/* henon_tracker::megno (src/henon.cu:2755-2887): per turn k, J at the pre-step
 * point with table index k, M <- J M, a_k = (k+1) log10(tr M_{k+1}/tr M_k), then
 * the map step (table index steps, or steps - 1 with inverse = 1: the inverse
 * map, as the reference does); out[i][q] = S2[turn_samples[q]] / turn_samples[q]
 * with S2 the running sum of (running sum of a_k)/(k+1); a sample equal to
 * n_turns reads 0 (the reference's never-written slot). M starts from
 * matrix_in[i] (device double[n][16]) or, when NULL, the identity; the two
 * traces start at 4 either way, as in the reference. matrix_out (device
 * double[n][16], or NULL) receives the final M. At most 16 samples.
 * out: device double[n][n_samples]. */
int hm_megno_trace(double *x, double *px, double *y, double *py, uint32_t *steps, size_t n, const double *table,
                   size_t table_len, uint32_t n_turns, double mu, double barrier, int inverse,
                   const uint32_t *turn_samples, size_t n_samples, const double *matrix_in, double *matrix_out,
                   double *out, void *stream);

#define HM_MAX_SAMPLES 16
#define HM_MAX_BIRKHOFF 4
/* What hm_track_indicators records; NULL outputs are skipped. */
typedef struct hm_indicator_spec {
    uint32_t renorm_every;             /* M <- M/|M|_F every k turns (0 = never) */
    uint32_t n_samples;                /* turn counts at which MEGNO / FLI are recorded */
    uint32_t samples[HM_MAX_SAMPLES];  /* 1..n_turns, increasing */
    double fli_dir[4];                 /* FLI tangent vector (normalised; zero = (1,0,0,0)) */
    double *megno;                     /* device [n_samples][n] */
    double *fli;                       /* device [n_samples][n] */
    double *matrix;                    /* device [n][16]: initial M in (identity if NULL), final M out */
    double *log_scale;                 /* device [n]: log10 of the scale factored out of M, in/out */
    uint32_t n_birkhoff;               /* Birkhoff tune windows over samples (from, to] */
    uint32_t birk_from[HM_MAX_BIRKHOFF];
    uint32_t birk_to[HM_MAX_BIRKHOFF];
    double *birk_out;                  /* device [n_birkhoff][2][n]: tune x, tune y */
    uint32_t hist_from, hist_len;      /* trajectory window: samples [hist_from, hist_from+hist_len) */
    double *hist;                      /* device [hist_len][4][n] (x, px, y, py) */
    /* megno_birkhoff_construct (src/henon.cu:2366-2402) fed every turn of this
     * call, idx = 1, 2, ...: for turn sample T_j, L2_j += L1 w_j[idx-1] / idx
     * while idx <= T_j, w_j = birkhoff_weights(T_j), L1 the MEGNO recursion's
     * first layer; out [n_megno_birkhoff][n] holds L2_j (get_values()). */
    uint32_t n_megno_birkhoff;
    uint32_t megno_birkhoff_samples[HM_MAX_BIRKHOFF];
    double *megno_birkhoff_out;
    /* tune_birkhoff_construct (src/henon.cu:2533-2584), first_add at the
     * initial state and add after every turn: with the wrapped phase advance
     * d (d <= 0 gets + 2 pi), tune1_j += d w_j[idx] while idx < T_j, tune2_j +=
     * d w_j[idx - T_j] while T_j <= idx < 2 T_j; out [n_tune_birkhoff][4][n]:
     * tune1 x, tune2 x, tune1 y, tune2 y (the getters' sums). */
    uint32_t n_tune_birkhoff;
    uint32_t tune_birkhoff_samples[HM_MAX_BIRKHOFF];
    double *tune_birkhoff_out;
} hm_indicator_spec;

/* Fused tracking + tangent map + indicators, one thread per particle with the
 * 4x4 tangent matrix in registers (replaces the per-turn construct launches of
 * src/henon.cu:267-359, 2223-2299, 2428-2655 driven from Python, and the CPU
 * loops of :2755-2967). Per turn: J at the pre-step point (table index
 * steps+k), M <- J M, map step, loss test; MEGNO in the construct form
 * (src/henon.cu:2259-2283: L1 += log10(|M_k+1|_F/|M_k|_F) idx, L2 += L1/idx,
 * value L2/idx); FLI = log10|M_t v0| / t; Birkhoff tunes as
 * dynamic_indicator.cpp:81-131 over the (n_turns+1)-sample trajectory; the
 * trajectory window feeds hm_fft_tunes. mode: HM_MODE_EXACT, HM_MODE_FAST or
 * HM_MODE_EXACT_ORBIT. */
int hm_track_indicators(double *x, double *px, double *y, double *py, uint32_t *steps, size_t n,
                        const double *table, size_t table_len, uint32_t n_turns, double mu, double barrier, int mode,
                        const hm_indicator_spec *spec, void *stream);

/* src/henon.cu:267-313 (get_matrices_and_multiply / _set): M <- J M (multiply=1)
 * or M <- J (multiply=0), J at the current points with table index table_index. */
int hm_tangent_apply(double *mats, const double *x, const double *px, const double *y, const double *py, size_t n,
                     const double *table, size_t table_len, uint32_t table_index, double mu, int multiply,
                     void *stream);
/* src/henon.cu:1721-1741 (matrix_4d_vector::multiply): M <- L M per particle. */
int hm_matrix_left_multiply(double *mats, const double *left, size_t n, void *stream);
/* src/henon.cu:1708-1719, 1808-1822 (reset to the identity). */
int hm_matrix_identity(double *mats, size_t n, void *stream);
/* src/henon.cu:316-338 (matrix_vector_multiply, reference_order=1, out may equal
 * vecs) and :1758-1787 (get_vector, reference_order=0). */
int hm_matvec(const double *mats, const double *vecs, double *out, size_t n, int reference_order, void *stream);
/* src/henon.cu:341-359 (normalize_vectors). */
int hm_normalize(double *vecs, size_t n, void *stream);
/* src/henon.cu:2083-2098 (kernel_lyap_birk). */
int hm_lyapunov_add(double *values, double *values_b, const double *vecs, const double *weights, size_t n,
                    uint64_t n_weights, uint64_t idx, void *stream);
/* src/henon.cu:2253-2284 (kernel_megno). */
int hm_megno_add(double *l1, double *l2, double *l3, const double *ma, const double *mb, size_t n, uint64_t idx,
                 void *stream);
/* src/henon.cu:2366-2402 (kernel_megno_birkhoff); l2: device [n_samples][n],
 * weights of sample j at w_flat + w_off[j] (all device). */
int hm_megno_birkhoff_add(double *l1, double *l2, const double *ma, const double *mb, size_t n, uint64_t idx,
                          const double *w_flat, const uint64_t *w_off, const uint64_t *turn_samples,
                          uint32_t n_samples, void *stream);
/* src/henon.cu:2533-2584 (kernel_tune_birkhoff_init when init=1, else
 * kernel_tune_birkhoff); tune arrays: device [n_samples][n]. */
int hm_tune_birkhoff_add(const double *x, const double *px, const double *y, const double *py, double *store_x,
                         double *store_y, double *t1x, double *t2x, double *t1y, double *t2y, const double *w_flat,
                         const uint64_t *w_off, const uint64_t *turn_samples, uint32_t n_samples, size_t n,
                         uint64_t idx, int init, void *stream);

/* ---- stored trajectories and tunes --------------------------------------- */
/* History layout: hist[sample][4][n] (x, px, y, py), sample 0 = the state before
 * the first turn. */
/* Track n_turns (exact arithmetic) writing samples 0..n_turns into hist
 * (device double[n_turns+1][4][n]); the trajectory store of
 * henon_tracker::birkhoff_tunes / all_tunes (src/henon.cu:2913-2943, 3004-3037). */
int hm_track_history(double *x, double *px, double *y, double *py, uint32_t *steps, size_t n, const double *table,
                     size_t table_len, uint32_t n_turns, double mu, double barrier, int inverse, double *hist,
                     void *stream);
/* Per-coordinate means over n_samples samples, summed in sample order from 0.0
 * (std::accumulate, src/henon.cu:3041-3051). means: device double[4][n]. */
int hm_history_means(const double *hist, size_t n, uint32_t n_samples, double *means, void *stream);
/* Birkhoff tunes (src/dynamic_indicator.cpp:81-131, 198-253) of the windows
 * [from[w], to[w]] (inclusive sample indices) of hist, after subtracting means
 * (device double[4][n], or NULL); NaN when a sample of the window is NaN.
 * from, to: host arrays. out: device double[n_windows][2][n] (tune x, tune y). */
int hm_birkhoff_tunes_history(const double *hist, size_t n, uint32_t n_samples, const double *means,
                              const uint32_t *from, const uint32_t *to, size_t n_windows, double *out, void *stream);
/* FFT tunes (src/dynamic_indicator.cpp:32-61, 134-196, 255-331): Hanning window
 * 2 sin^2(pi i/N), forward DFT of (x-mean) + i (px-mean), magnitude, first
 * maximum, three-point interpolation, |1 - (k/N + asin(A sin(2pi/N))/2pi)|; 1.0 /
 * 0.0 at the spectrum's edges; NaN when a sample of the window is NaN. Windows
 * start at sample from[w] and have len[w] samples (host arrays); any length of
 * at least 2. Power-of-two lengths 8..2048 use a Stockham radix-8 FFT in LDS
 * (from a sample-major history the windows are first corner-turned into
 * per-particle signal rows: library workspace of at most 4 GiB, particles taken
 * in batches that fit; the environment variable HM_FFT_ROW_BYTES overrides the
 * budget); other lengths up to about 10^4 samples one LDS-resident transform per
 * signal (mixed radix with prime passes up to 257, or chirp-z over a smooth
 * length that fits); longer ones four-step mixed-radix or chirp-z transforms
 * through HBM scratch (also sized by HM_FFT_ROW_BYTES, 8 GiB by default).
 * hm_fft_tunes_rows reads particle rows in place. out: device
 * double[n_windows][2][n]. */
int hm_fft_tunes(const double *hist, size_t n, uint32_t n_samples, const double *means, const uint32_t *from,
                 const uint32_t *len, size_t n_windows, double *out, void *stream);

/* The same trajectory store as particle rows: rows[particle][4][row_len]
 * (x, px, y, py), samples 0..n_turns at row positions 0..n_turns; row_len a
 * multiple of 16 of at least n_turns+1 (positions past n_turns hold the final
 * state), rows 16-byte aligned. Each particle's samples are contiguous, so the
 * tune kernels below read them in place (no corner turn); every store
 * completes 128-byte lines. means: NULL, or device double[4][n] that receives
 * each coordinate's mean over samples 0..n_turns, equal bit for bit to
 * hm_rows_means(rows, n, row_len, n_turns + 1) (the mean removal of all_tunes,
 * src/henon.cu:3041-3051) without a second pass over the rows. */
int hm_track_rows(double *x, double *px, double *y, double *py, uint32_t *steps, size_t n, const double *table,
                  size_t table_len, uint32_t n_turns, double mu, double barrier, int inverse, double *rows,
                  size_t row_len, double *means, void *stream);
/* hm_history_means, hm_birkhoff_tunes_history and hm_fft_tunes on particle rows;
 * n_samples <= row_len. Means and FFT tunes equal the sample-major results bit
 * for bit; Birkhoff tunes sum the same terms in a wavefront-parallel order
 * (equal to rounding). `means` is NULL or hm_rows_means of the same rows and
 * n_samples: four finite means of a particle then prove that none of its samples
 * is NaN, and hm_fft_tunes_rows reads only the rows of the plane it transforms. */
int hm_rows_means(const double *rows, size_t n, size_t row_len, uint32_t n_samples, double *means, void *stream);
int hm_birkhoff_tunes_rows(const double *rows, size_t n, size_t row_len, uint32_t n_samples, const double *means,
                           const uint32_t *from, const uint32_t *to, size_t n_windows, double *out, void *stream);
int hm_fft_tunes_rows(const double *rows, size_t n, size_t row_len, uint32_t n_samples, const double *means,
                      const uint32_t *from, const uint32_t *len, size_t n_windows, double *out, void *stream);
/* Birkhoff and FFT tunes of the same windows [from, to] of particle rows, as
 * all_tunes computes both (src/henon.cu:2969-3078: birkhoff_tunes' windows and
 * the FFT of samples from..to): equal to hm_birkhoff_tunes_rows (within
 * rounding: another summation order) and hm_fft_tunes_rows (bit for bit) with
 * len = to - from + 1. A window of 8, 16, ..., 2048 samples is read once for
 * both (one workgroup per particle and plane); other lengths run the two
 * separately. birkhoff_out, fft_out: device double[n_windows][2][n]. */
int hm_tunes_rows(const double *rows, size_t n, size_t row_len, uint32_t n_samples, const double *means,
                  const uint32_t *from, const uint32_t *to, size_t n_windows, double *birkhoff_out, double *fft_out,
                  void *stream);

/* ---- reverse-mode derivatives ---------------------------------------------- */
/* No counterpart in the reference, whose only derivative is the one-turn tangent
 * matrix (src/henon.cu:175-265); these extend it to whole-orbit gradients.
 * Loss model: L = L(s_T, B) with s_T the state after n_turns forward turns from
 * x0 (table rows 0..n_turns-1) and B = sum_{k<n_turns} w[k] x_{k+1}^2 the
 * weighted time average of x^2 (w = hm_birkhoff_weights(n_turns) or any
 * weights). */
/* Forward sweep: writes the state before turns 0, every, 2 every, ... into
 * checkpoints (device double[ceil(n_turns/every)][4][n]), the final state
 * (double[4][n], NaN when lost), B (double[n]) and alive (uint8[n], 1 when the
 * particle stayed inside the barrier for all n_turns). Same arithmetic as
 * hm_track for the same mode: in exact mode final_state equals hm_track's. */
int hm_adjoint_forward(const double *x0, const double *px0, const double *y0, const double *py0, size_t n,
                       const double *table, size_t table_len, uint32_t n_turns, double mu, double barrier,
                       uint32_t every, const double *weights, double *checkpoints, double *final_state, double *birk,
                       uint8_t *alive, int mode, void *stream);
/* Bytes of scratch for hm_adjoint_backward with n particles, n_turns turns and
 * checkpoint interval every: the 8-turn sub-checkpoints of every K-turn segment
 * (4 doubles per particle per chunk after a segment's first), capped at 4 GiB
 * (environment variable HM_ADJOINT_SCRATCH_BYTES sets another cap) but never
 * below one segment. 0 when every <= 8. */
int hm_adjoint_scratch_bytes(size_t n, uint32_t n_turns, uint32_t every, size_t *bytes);
/* Backward sweep: given seed_state = dL/ds_T (double[4][n]) and seed_birk =
 * dL/dB (double[n], or NULL for 0), writes grad_x0 = dL/dx0 (double[4][n]) and
 * grad_params = (dL/d epsilon, dL/d mu, dL/d omega_x0, dL/d omega_y0)
 * (double[4][n]), with dtable[k] = (d omega_x(k)/d omega_x0, d omega_x(k)/d eps,
 * d omega_y(k)/d omega_y0, d omega_y(k)/d eps) (device double[n_turns][4], from
 * hm_lines_modulation_derivatives). Lost particles get zero gradients. The
 * K-turn segments are recomputed from their checkpoints, all at once, into
 * sub-checkpoints every 8 turns in scratch (scratch_bytes of device memory,
 * at least one segment's worth: fewer bytes than all segments need make the
 * sweep run in windows of segments); then each particle's 8-turn chunks are
 * recomputed and swept in reverse. The map is never inverted. */
int hm_adjoint_backward(const double *checkpoints, size_t n, const double *table, const double *dtable,
                        size_t table_len, uint32_t n_turns, double mu, uint32_t every, const double *weights,
                        const double *seed_state, const double *seed_birk, const uint8_t *alive, double *scratch,
                        size_t scratch_bytes, double *grad_x0, double *grad_params, int mode, void *stream);
/* Host: for omega(k) = 2 pi tune (1 + eps sum_j coef_j cos(freq_j (k+start)))
 * (hm_lines_modulation, hm_sps_modulation with hm_sps_lines' set), out[2k] =
 * d omega(k)/d omega_0 with omega_0 = 2 pi tune and out[2k+1] = d omega(k)/d eps. */
int hm_lines_modulation_derivatives(double tune, double epsilon, const double *coef, const double *freq,
                                    int n_lines, int start, int end, double *out);

/* ---- instrumentation ----------------------------------------------------- */
/* When enabled, every kernel launch of the library is bracketed by HIP events
 * on its own stream; hm_profile_read synchronises them and returns the summed
 * kernel time (ms) and launch count for kernels whose name contains `name`. */
int hm_profile_enable(int on);
int hm_profile_reset(void);
int hm_profile_read(const char *name, double *total_ms, uint64_t *launches);

#ifdef __cplusplus
}
#endif

#endif /* HENON_HIP_H_ */
