/*
 * henon_hip.h -- C ABI of the MI355X (gfx950) 4D modulated Henon-map tracker.
 *
 * Plain C: pointers, sizes and status codes; no C++ or torch types. Every entry
 * point returns HM_OK (0) or an HM_ERR_* code; hm_last_error() gives the
 * message of the calling thread's last failure. Device pointers are HIP device
 * allocations (hm_malloc, or any hipMalloc'd / torch CUDA tensor memory); a
 * `stream` argument is a hipStream_t passed as void* (NULL = the null stream).
 * Kernel launches are asynchronous with respect to the host unless stated.
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

enum { HM_MODE_EXACT = 0, HM_MODE_FAST = 1 };

/* Last error message of the calling thread (never NULL).
 * Replaces check_cuda_errors' print-and-exit(1) (src/henon.cu:17-25). */
const char *hm_last_error(void);

/* Library version string. */
const char *hm_version(void);

/* ---- device and memory plumbing ---------------------------------------- */
/* Number of visible HIP devices (0 when none). Replaces
 * is_cuda_device_available (src/bind.cpp:28-39). */
int hm_device_count(int *count);
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
    int32_t block;          /* threads per workgroup (64..1024, multiple of 64) */
    int32_t compact_every;  /* turns between loss compactions (0 = default, <0 = never) */
    int32_t unroll;         /* reserved */
    int32_t reserved;
} hm_launch_cfg;

/* Track n particles for n_turns turns of the modulated map.
 * Replaces henon_tracker_gpu::track (src/henon.cu:3155-3185) + gpu_henon_track
 * (:636-691) and the CPU henon_tracker::track (:2701-2752): one persistent launch
 * per compaction segment instead of one launch per turn.
 *   x, px, y, py, steps : device SoA, updated in place (steps = survival turns)
 *   table, table_len    : device table[table_len][4]
 *   mu                  : octupole strength (the octupole kick is skipped when 0)
 *   barrier             : loss radius; lost when x^2+px^2+y^2+py^2 >= barrier^2,
 *                         then all four coordinates become NaN
 *   kick_module         : must be NaN (the random kick is not available)
 *   inverse             : 0 forward (table index steps), 1 inverse (steps-1)
 *   mode                : HM_MODE_EXACT (bit-exact with the reference CPU path)
 *                         or HM_MODE_FAST (fused multiply-adds)
 *   cfg                 : NULL for the defaults
 * The caller owns the global turn counter and its range check (the Python/C++
 * layer raises "The number of turns is too large."). */
int hm_track(double *x, double *px, double *y, double *py, uint32_t *steps, size_t n,
             const double *table, size_t table_len, uint32_t n_turns, double mu, double barrier,
             double kick_module, int inverse, int mode, const hm_launch_cfg *cfg, void *stream);

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

/* Tangent matrix of one turn at the current points, table index table_index:
 * J (reverse=0) or the analytic inverse J^-1 (reverse=1); particles with
 * valid[i]==0 get a NaN matrix. out: device double[n][16].
 * Replaces henon_tracker_gpu::get_tangent_matrix (src/henon.cu:3187-3219),
 * which computed it on the host. */
int hm_tangent_matrix(const double *x, const double *px, const double *y, const double *py,
                      const uint8_t *valid, size_t n, const double *table, size_t table_len,
                      uint32_t table_index, double mu, int reverse, double *out, void *stream);

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
