// Tunes from stored trajectories (replaces FFTW and the per-particle CPU loops
// of src/dynamic_indicator.cpp:32-361 and src/henon.cu:2890-3078, 3221-3463).
//  * k_history        : track n turns, writing every sample to HBM
//  * k_means          : per-coordinate mean of each particle's samples
//  * k_birkhoff_hist  : Birkhoff tunes of windows of the stored samples
//  * k_fft_tune_stockham / k_fft_tune_2048 : Hanning-windowed FFT tune,
//                       power-of-two lengths 8 .. 2048, one signal per workgroup
//  * other lengths    : whole-signal and four-step mixed radix
//                       (hm_fft_mixed.hip) or power-of-two chirp-z (hm_chirpz.hip)
// Two history layouts, addressed through strides (HistLayout):
//  * sample-major hist[sample][coord][particle] (coord = x, px, y, py): a
//    tracking wavefront writes one contiguous row per turn (k_history);
//  * particle rows rows[particle][coord][L]: each particle's samples contiguous
//    (k_history_rows buffers 16 turns in registers so that every store
//    completes 128-byte lines), which the FFT reads without a corner turn.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>

#include "hm_device.h"
#include "hm_internal.h"
#include "hm_tune.h"

namespace hm {

constexpr int kMaxWindows = 64;


// ---------------------------------------------------------------- history
template <bool INV, bool OCT>
__global__ void __launch_bounds__(256)
k_history(double *__restrict__ gx, double *__restrict__ gpx, double *__restrict__ gy, double *__restrict__ gpy,
          uint32_t *__restrict__ gsteps, uint32_t n, const TabEntry *__restrict__ tab, uint32_t tab_len,
          uint32_t n_turns, double mu, double b2, double *__restrict__ hist, uint32_t *__restrict__ status)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = gx[i], px = gpx[i], y = gy[i], py = gpy[i];
    uint32_t s = gsteps[i];
    const size_t nn = n;
    hist[0 * nn + i] = x; hist[1 * nn + i] = px; hist[2 * nn + i] = y; hist[3 * nn + i] = py;
    bool alive = !any_nan(x, px, y, py);
    const double m3mu = -3.0 * mu;
    for (uint32_t k = 0; k < n_turns; ++k) {
        if (alive) {
            const int64_t ti = INV ? (int64_t)s - 1 : (int64_t)s;
            if (ti < 0 || ti >= (int64_t)tab_len) {
                atomicOr(status, 1u);
                alive = false;
            } else {
                map_step<kExact, INV, OCT>(x, px, y, py, mu, m3mu, tab[ti]);
                if (radius2<kExact>(x, px, y, py) < b2) {
                    s = INV ? s - 1 : s + 1;
                } else {
                    x = px = y = py = __builtin_nan("");
                    alive = false;
                }
            }
        }
        const size_t base = (size_t)(k + 1) * 4 * nn;
        hist[base + 0 * nn + i] = x; hist[base + 1 * nn + i] = px;
        hist[base + 2 * nn + i] = y; hist[base + 3 * nn + i] = py;
    }
    gx[i] = x; gpx[i] = px; gy[i] = y; gpy[i] = py; gsteps[i] = s;
}

// Same tracking and samples as k_history, written as particle rows
// rows[particle][coord][L] (L a multiple of 16, samples 0..n_turns, the tail of
// the last block padded with the final state): 16 samples per coordinate are
// kept in registers; per coordinate the wavefront passes them through LDS so
// that each store instruction writes 8 whole 128-byte lines (8 particles x 16
// samples) instead of touching 64 lines (one per particle). With MEANS each
// coordinate's samples are also summed in sample order from 0.0 and divided by
// n_turns + 1: k_means' result on the same rows, without reading them again.
template <bool INV, bool OCT, bool MEANS>
__global__ void __launch_bounds__(256)
k_history_rows(double *__restrict__ gx, double *__restrict__ gpx, double *__restrict__ gy, double *__restrict__ gpy,
               uint32_t *__restrict__ gsteps, uint32_t n, const TabEntry *__restrict__ tab, uint32_t tab_len,
               uint32_t n_turns, double mu, double b2, double *__restrict__ rows, uint64_t L,
               double *__restrict__ means, uint32_t *__restrict__ status)
{
    // per wavefront: 64 particles x 16 samples, rows padded to 18 doubles
    // (16-byte aligned, spread over the LDS banks)
    constexpr int kPad = 18;
    __shared__ __align__(16) double tr_all[4 * 64 * kPad];
    double *const tr = tr_all + (threadIdx.x >> 6) * 64 * kPad;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lane = threadIdx.x & 63, wave0 = i - lane;
    if (wave0 >= n) return;  // whole wavefronts; lanes past n run along and store nothing
    const bool in = i < n;
    double x = __builtin_nan(""), px = x, y = x, py = x;
    uint32_t s = 0;
    if (in) {
        x = gx[i]; px = gpx[i]; y = gy[i]; py = gpy[i];
        s = gsteps[i];
    }
    bool alive = in && !any_nan(x, px, y, py);
    const double m3mu = -3.0 * mu;
    const uint32_t samples = n_turns + 1;
    double sum[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint32_t m0 = 0; m0 < samples; m0 += 16) {
        double buf[4][16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const uint32_t m = m0 + q;
            if (m > 0 && m < samples && alive) {
                const int64_t ti = INV ? (int64_t)s - 1 : (int64_t)s;
                if (ti < 0 || ti >= (int64_t)tab_len) {
                    atomicOr(status, 1u);
                    alive = false;
                } else {
                    map_step<kExact, INV, OCT>(x, px, y, py, mu, m3mu, tab[ti]);
                    if (radius2<kExact>(x, px, y, py) < b2) {
                        s = INV ? s - 1 : s + 1;
                    } else {
                        x = px = y = py = __builtin_nan("");
                        alive = false;
                    }
                }
            }
            buf[0][q] = x; buf[1][q] = px; buf[2][q] = y; buf[3][q] = py;
            if (MEANS && m < samples) {
                sum[0] = sum[0] + x; sum[1] = sum[1] + px; sum[2] = sum[2] + y; sum[3] = sum[3] + py;
            }
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            double2 *mine = reinterpret_cast<double2 *>(tr + lane * kPad);
#pragma unroll
            for (int q = 0; q < 8; ++q) mine[q] = double2{buf[c][2 * q], buf[c][2 * q + 1]};
            __builtin_amdgcn_wave_barrier();  // the LDS writes precede the reads (in order per wavefront)
            // store t: particles 8t..8t+7 of the wavefront, lane l the 16 bytes
            // (l % 8) of particle 8t + l / 8
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const uint32_t pw = 8 * t + (lane >> 3), part = lane & 7;
                const double2 v = reinterpret_cast<const double2 *>(tr + pw * kPad)[part];
                if (wave0 + pw < n)
                    reinterpret_cast<double2 *>(rows + (size_t)(wave0 + pw) * 4 * L + c * L + m0)[part] = v;
            }
            __builtin_amdgcn_wave_barrier();  // the reads precede the next coordinate's writes
        }
    }
    if (!in) return;
    gx[i] = x; gpx[i] = px; gy[i] = y; gpy[i] = py; gsteps[i] = s;
    if (MEANS) {
        const size_t nn = n;
#pragma unroll
        for (int c = 0; c < 4; c++) means[c * nn + i] = sum[c] / (double)samples;
    }
}

// std::accumulate order (henon.cu:3041-3051): sum from 0.0 in sample order, / L
__global__ void k_means(const double *__restrict__ hist, HistLayout h, uint32_t n, uint32_t L,
                        double *__restrict__ means)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t nn = n;
    for (int c = 0; c < 4; c++) {
        double s = 0.0;
        for (uint32_t m = 0; m < L; m++) s = s + hist[hat(h, m, c, i)];
        means[c * nn + i] = s / (double)L;
    }
}

struct WindowSet {
    uint32_t from[kMaxWindows], len[kMaxWindows], w_off[kMaxWindows];
    uint32_t count;
};

// dynamic_indicator.cpp:81-131 per window (len samples from `from`): angles of
// (x - mean) + i (px - mean), wrapped differences (+2pi when negative), weights
// birkhoff_weights(len-1); NaN when a sample of the window is NaN
// (dynamic_indicator.cpp:228-241). out: [window][plane][n].
// One thread per (particle, plane, window): enough wavefronts to hide the
// latency of the per-sample atan2 chains.
__global__ void k_birkhoff_hist(const double *__restrict__ hist, HistLayout h, uint32_t n, uint32_t p0, uint32_t nb,
                                const double *__restrict__ means, WindowSet ws, const double *__restrict__ w_flat,
                                double *__restrict__ out)
{
    // particles p0 .. p0 + nb - 1 of n
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t i = p0 + (uint32_t)(idx % nb);
    const uint64_t rest = idx / nb;
    const int p = (int)(rest % 2);
    const uint32_t w = (uint32_t)(rest / 2);
    if (w >= ws.count) return;
    const size_t nn = n;
    const int cx = 2 * p, cp = 2 * p + 1;
    double mc[4] = {0.0, 0.0, 0.0, 0.0};
    if (means)
        for (int q = 0; q < 4; q++) mc[q] = means[q * nn + i];
    const double mx = mc[cx], mp = mc[cp];
    double acc = 0.0, th_prev = 0.0;
    bool bad = false;
    for (uint32_t m = 0; m < ws.len[w]; m++) {
        const uint64_t sm = ws.from[w] + m;
        const double a0 = hist[hat(h, sm, 0, i)], a1 = hist[hat(h, sm, 1, i)];
        const double a2 = hist[hat(h, sm, 2, i)], a3 = hist[hat(h, sm, 3, i)];
        // any NaN among the four coordinates of the window after the mean
        // removal (dynamic_indicator.cpp:228-241 on the mean-removed vectors)
        bad |= any_nan(a0 - mc[0], a1 - mc[1], a2 - mc[2], a3 - mc[3]);
        const double xv = (p ? a2 : a0) - mx;
        const double pv = (p ? a3 : a1) - mp;
        const double th = atan2(pv, xv);
        if (m > 0) {
            double d = th - th_prev;
            if (d < 0.0) d += 2.0 * M_PI;
            acc = acc + w_flat[ws.w_off[w] + m - 1] * d;
        }
        th_prev = th;
    }
    out[((size_t)w * 2 + p) * nn + i] = bad ? __builtin_nan("") : 1 - acc / (2 * M_PI);
}

// Birkhoff tunes of particle rows: one wavefront per (particle, plane,
// window); lane l takes samples m0 + l of each 64-sample step (coalesced along
// the row), the angle of the sample before comes from the neighbouring lane
// (from the previous step's lane 63 for lane 0), and the weighted wrapped
// differences are summed per lane and then across the wavefront in a fixed
// order. Same terms as k_birkhoff_hist, summed in that order instead of
// sample order (within rounding of it).
__global__ void __launch_bounds__(256) k_birkhoff_rows(const double *__restrict__ rows, uint64_t L, uint32_t n,
                                                       uint32_t p0, uint32_t nb, const double *__restrict__ means,
                                                       WindowSet ws, const double *__restrict__ w_flat,
                                                       double *__restrict__ out)
{
    // particles p0 .. p0 + nb - 1 of n
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t i = p0 + (uint32_t)(wave % nb);
    const uint64_t rest = wave / nb;
    const int p = (int)(rest % 2);
    const uint32_t w = (uint32_t)(rest / 2);
    if (w >= ws.count) return;  // whole wavefronts
    const size_t nn = n;
    const int cx = 2 * p, cp = 2 * p + 1;
    const double mx = means ? means[cx * nn + i] : 0.0;
    const double mp = means ? means[cp * nn + i] : 0.0;
    const double *r = rows + (size_t)i * 4 * L + ws.from[w];
    const uint32_t len = ws.len[w];
    // NaN rule on the mean-removed samples (as k_birkhoff_hist): a NaN mean
    // makes every sample of its coordinate NaN, so the tune is NaN without
    // reading the rows; four finite means prove that no sample is NaN, and
    // only this plane's two rows are read
    const double mo0 = means ? means[(2 - cx) * nn + i] : 0.0, mo1 = means ? means[(3 - cx) * nn + i] : 0.0;
    if (any_nan(mx, mp, mo0, mo1)) {
        if (lane == 0) out[((size_t)w * 2 + p) * nn + i] = __builtin_nan("");
        return;  // whole wavefronts
    }
    const bool scan_other = !means || !(isfinite(mx) && isfinite(mp) && isfinite(mo0) && isfinite(mo1));
    const double *rx = r + cx * L, *rp = r + cp * L, *ro0 = r + (2 - cx) * L, *ro1 = r + (3 - cx) * L;
    double acc = 0.0, carry = 0.0;
    bool bad = false;
    // eight 64-sample steps per iteration: their loads are issued before the
    // first angle, so several are in flight per wavefront
    constexpr uint32_t U = 8;
    for (uint32_t m00 = 0; m00 < len; m00 += 64 * U) {
        double a0[U], a1[U];
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            const uint32_t m = m00 + 64 * u + lane;
            a0[u] = a1[u] = 0.0;
            if (m < len) {
                a0[u] = rx[m];
                a1[u] = rp[m];
                if (scan_other) bad |= any_nan(ro0[m] - mo0, ro1[m] - mo1, 0.0, 0.0);
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < U; ++u) {
            const uint32_t m = m00 + 64 * u + lane;
            const bool valid = m < len;
            double th = 0.0;
            if (valid) {
                bad |= any_nan(a0[u] - mx, a1[u] - mp, 0.0, 0.0);
                th = atan2(a1[u] - mp, a0[u] - mx);
            }
            double prev = __shfl_up(th, 1);
            if (lane == 0) prev = carry;
            carry = __shfl(th, 63);
            if (valid && m > 0) {
                double d = th - prev;
                if (d < 0.0) d += 2.0 * M_PI;
                acc = acc + w_flat[ws.w_off[w] + m - 1] * d;
            }
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc = acc + __shfl_xor(acc, off);
    bad = __any(bad);
    if (lane == 0) out[((size_t)w * 2 + p) * nn + i] = bad ? __builtin_nan("") : 1 - acc / (2 * M_PI);
}

// ---------------------------------------------------------------- FFT tune
// Corner turn for the FFT: samples from..from+L-1 of the history rows
// hist[sample][coord][particle] (particle-minor: what tracking writes) into
// signal rows rows[particle - p0][coord][sample] minus the particle's means, so
// that a workgroup reads its particle's signals contiguously. 64 x 64 tiles
// through LDS, coalesced on both sides. grid: (particles/64, L/64, 4 coords).
__global__ void __launch_bounds__(256) k_signal_rows(const double *__restrict__ hist, uint32_t n,
                                                     const double *__restrict__ means, uint32_t from, uint32_t L,
                                                     uint32_t p0, uint32_t nb, double *__restrict__ rows)
{
    __shared__ double tile[64][65];
    const uint32_t tp = blockIdx.x * 64, ts = blockIdx.y * 64, c = blockIdx.z;
    const uint32_t tx = threadIdx.x & 63u, ty = threadIdx.x >> 6;
    const size_t nn = n;
    {
        const uint32_t q = tp + tx;  // batch-relative particle
        const double m = (means && q < nb) ? means[c * nn + p0 + q] : 0.0;
        for (uint32_t r = ty; r < 64; r += 4) {
            const uint32_t sm = ts + r;
            if (q < nb && sm < L) tile[r][tx] = hist[(size_t)(from + sm) * 4 * nn + c * nn + p0 + q] - m;
        }
    }
    __syncthreads();
    for (uint32_t r = ty; r < 64; r += 4) {
        const uint32_t q = tp + r, sm = ts + tx;
        if (q < nb && sm < L) rows[((size_t)q * 4 + c) * L + sm] = tile[tx][r];
    }
}

struct FftRowArgs {
    const double *rows;    // signal of block b, coordinate c, sample m at
                           // rows[(row0 + b) * pstride + c * cstride + from + m]
    uint64_t pstride, cstride, row0, from;
    const double *means;   // [4][n] subtracted here, or NULL (already removed)
    uint32_t nb, p0, n, N;
    const double *hann;    // [N]
    const double *tw;      // [N] complex (cos, -sin)
    double *out;           // [window][plane][n]
    uint32_t window;
    const double *bw;      // Birkhoff weights [N - 1] (fused kernel only)
    double *bout;          // Birkhoff tunes [window][plane][n] (fused kernel only)
};

// ---------------------------------------------------------------- Stockham FFT
// Self-sorting (Stockham) FFT of one signal per workgroup, radix-8 passes with
// a final radix-2 or radix-4 pass (N = 2048: 8 8 8 4), butterflies in
// registers, the signal in LDS between passes. Pass with radix R and span Ns
// (the product of the earlier radices): butterfly j reads x[j + r N/R],
// multiplies by w^(k r N/(Ns R)) (k = j mod Ns, w = exp(-2 pi i / N); one
// twiddle from the table, its powers by products), takes the R-point DFT and
// writes element r to
// (j / Ns) Ns R + k + r Ns; after the last pass the spectrum is in natural
// order.

constexpr int kStockThreads = 256;

// LDS position of spectrum entry i: one pad entry after every 16, so that the
// stores of the first pass (eight entries of 16 bytes apart per thread) fall
// into distinct banks.
__device__ __forceinline__ uint32_t stk_pos(uint32_t i) { return i + (i >> 4); }

// One pass: butterflies j = t, t + 256, ... of N / R; inputs from LDS (or, for
// the first pass, from `load`), outputs to LDS, a barrier between the two.
template <int LOGN, int R, bool FIRST, class Load>
__device__ __forceinline__ void stockham_pass(Cplx *buf, uint32_t Ns, const double *__restrict__ tw, Load load)
{
    constexpr uint32_t N = 1u << LOGN, B = N / R;
    constexpr uint32_t PER = (B + kStockThreads - 1) / kStockThreads;
    Cplx v[PER][R];
#pragma unroll
    for (uint32_t q = 0; q < PER; ++q) {
        const uint32_t j = threadIdx.x + q * kStockThreads;
        if (j < B) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = FIRST ? load(j + r * B) : buf[stk_pos(j + r * B)];
            const uint32_t k = j % Ns;
            if (!FIRST && k != 0) {
                // w^r for w = exp(-2 pi i step / N): one table load, the higher
                // powers by products
                const uint32_t step = k * (N / (Ns * R));
                const Cplx w1{tw[2 * step], tw[2 * step + 1]};
                Cplx wr = w1;
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    v[q][r] = cmul(v[q][r], wr);
                    if (r + 1 < R) wr = cmul(wr, w1);
                }
            }
            dft_r<R>(v[q]);
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < PER; ++q) {
        const uint32_t j = threadIdx.x + q * kStockThreads;
        if (j < B) {
            const uint32_t k = j % Ns, base = (j / Ns) * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) buf[stk_pos(base + r * Ns)] = v[q][r];
        }
    }
    __syncthreads();
}

// First maximum (lowest index among equal values) of (value, index) over the
// workgroup: shuffles within each wavefront, then across the four.
__device__ __forceinline__ void argmax_first(double v, uint32_t i, double *sv, uint32_t *si, uint32_t &best_i)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(v, off);
        const uint32_t oi = __shfl_xor(i, off);
        if (ov > v || (ov == v && oi < i)) { v = ov; i = oi; }
    }
    const uint32_t w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sv[w] = v; si[w] = i; }
    __syncthreads();
    double bv = sv[0];
    uint32_t bi = si[0];
    for (uint32_t q = 1; q < kStockThreads / 64; ++q)
        if (sv[q] > bv || (sv[q] == bv && si[q] < bi)) { bv = sv[q]; bi = si[q]; }
    best_i = bi;
}

// One workgroup per (particle, plane) of a batch of signal rows (FftRowArgs):
// window, mean removal and Hanning at the load, the FFT, magnitudes, first
// maximum, interpolated tune. N = 2^LOGN, 8 <= N <= 1024 (2048: k_fft_tune_2048).
// BIRK: the Birkhoff tune of the same window from the same read of the rows
// (dynamic_indicator.cpp:81-131 as k_birkhoff_rows): the first pass's loads
// also give each sample's angle, the angles go through the LDS buffer before
// the transform uses it (the angle of sample m - 1 comes from its neighbour),
// and the weighted wrapped differences are summed per thread, then across the
// workgroup in a fixed order (within rounding of the sample-order sum).
template <int LOGN, bool BIRK>
__global__ void __launch_bounds__(kStockThreads) k_fft_tune_stockham(FftRowArgs a)
{
    constexpr uint32_t N = 1u << LOGN;
    __shared__ Cplx buf[N + (N - 1) / 16 + 1];  // entry i at stk_pos(i)
    __shared__ double sv[kStockThreads / 64];
    __shared__ uint32_t si[kStockThreads / 64];
    __shared__ double sb[kStockThreads / 64];
    __shared__ int any_nan_flag;
    const uint32_t b = blockIdx.x >> 1, plane = blockIdx.x & 1;
    if (b >= a.nb) return;
    const double *base = a.rows + (size_t)((a.row0 + b) * a.pstride + a.from);
    const double *rx = base + (2 * plane) * a.cstride, *rp = base + (2 * plane + 1) * a.cstride;
    const double *oth0 = base + (2 * (1 - plane)) * a.cstride, *oth1 = base + (2 * (1 - plane) + 1) * a.cstride;
    const size_t nn = a.n, pp = a.p0 + b;
    const double mx = a.means ? a.means[(2 * plane) * nn + pp] : 0.0;
    const double mp = a.means ? a.means[(2 * plane + 1) * nn + pp] : 0.0;
    const double mo0 = a.means ? a.means[(2 * (1 - plane)) * nn + pp] : 0.0;
    const double mo1 = a.means ? a.means[(2 * (1 - plane) + 1) * nn + pp] : 0.0;
    double *o = a.out + ((size_t)a.window * 2 + plane) * nn + pp;
    double *ob = BIRK ? a.bout + ((size_t)a.window * 2 + plane) * nn + pp : nullptr;
    if (any_nan(mx, mp, mo0, mo1)) {
        // a NaN mean makes every sample of its coordinate NaN after the mean
        // removal: the tune is NaN, nothing needs reading
        if (threadIdx.x == 0) {
            *o = __builtin_nan("");
            if constexpr (BIRK) *ob = __builtin_nan("");
        }
        return;  // the whole workgroup
    }
    if (threadIdx.x == 0) any_nan_flag = 0;
    __syncthreads();
    // A NaN in any of the four coordinates of the window marks both planes.
    // With means given (over every sample of the rows), four finite means
    // prove that no sample of the particle is NaN or infinite, and the other
    // plane's rows need not be read.
    const bool scan_other = !a.means || !(isfinite(mx) && isfinite(mp) && isfinite(mo0) && isfinite(mo1));
    bool nan_seen = false;
    auto load4 = [&](uint32_t m) -> Cplx {
        const double vx = rx[m] - mx, vp = rp[m] - mp;
        nan_seen |= any_nan(vx, vp, oth0[m] - mo0, oth1[m] - mo1);
        const double h = a.hann[m];
        return Cplx{vx * h, vp * h};
    };
    auto load2 = [&](uint32_t m) -> Cplx {
        const double vx = rx[m] - mx, vp = rp[m] - mp;
        nan_seen |= any_nan(vx, vp, 0.0, 0.0);
        const double h = a.hann[m];
        return Cplx{vx * h, vp * h};
    };
    constexpr int R8 = LOGN / 3, REM = LOGN % 3;
    constexpr int R1 = R8 > 0 ? 8 : (REM == 2 ? 4 : 2);
    double bacc = 0.0;
    if constexpr (BIRK) {
        // first radix-8 pass (Ns = 1, no twiddles) with the angles: thread j
        // takes samples m = j + r B (B = N / 8 <= 256, one butterfly per thread)
        static_assert(R1 == 8 && N / 8 <= kStockThreads, "fused kernel: one radix-8 butterfly per thread");
        constexpr uint32_t B = N / 8;
        double *th = reinterpret_cast<double *>(buf);  // N angles; buf is not in use yet
        const uint32_t j = threadIdx.x;
        Cplx v[8];
        double hw[8], bw[8];
        if (j < B) {
            // every load of the pass first (samples, Hanning factors, weights)
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                const uint32_t m = j + r * B;
                v[r] = Cplx{rx[m], rp[m]};
                hw[r] = a.hann[m];
                bw[r] = m > 0 ? a.bw[m - 1] : 0.0;
                if (scan_other) nan_seen |= any_nan(oth0[m] - mo0, oth1[m] - mo1, 0.0, 0.0);
            }
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                const double vx = v[r].re - mx, vp = v[r].im - mp;
                nan_seen |= any_nan(vx, vp, 0.0, 0.0);
                th[j + r * B] = atan2(vp, vx);
                v[r] = Cplx{vx, vp};
            }
        }
        __syncthreads();
        if (j < B) {
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                const uint32_t m = j + r * B;
                if (m > 0) {
                    double d = th[m] - th[m - 1];
                    if (d < 0.0) d += 2.0 * M_PI;
                    bacc = bacc + bw[r] * d;
                }
            }
        }
        __syncthreads();  // every angle read before the transform overwrites buf
        if (j < B) {
#pragma unroll
            for (int r = 0; r < 8; ++r) v[r] = Cplx{v[r].re * hw[r], v[r].im * hw[r]};
            dft_r<8>(v);
#pragma unroll
            for (int r = 0; r < 8; ++r) buf[stk_pos(j * 8 + r)] = v[r];
        }
        __syncthreads();
    } else {
        // first pass from memory (the same pass twice: a uniform branch)
        if (scan_other) stockham_pass<LOGN, R1, true>(buf, 1, a.tw, load4);
        else stockham_pass<LOGN, R1, true>(buf, 1, a.tw, load2);
    }
    uint32_t Ns = R1;
    auto &load = load2;  // unused by the later passes
    if (nan_seen) atomicOr(&any_nan_flag, 1);
#pragma unroll
    for (int p = 1; p < R8; ++p) {
        stockham_pass<LOGN, 8, false>(buf, Ns, a.tw, load);
        Ns *= 8;
    }
    if constexpr (R8 > 0 && REM == 2) stockham_pass<LOGN, 4, false>(buf, Ns, a.tw, load);
    if constexpr (R8 > 0 && REM == 1) stockham_pass<LOGN, 2, false>(buf, Ns, a.tw, load);
    if (any_nan_flag) {
        if (threadIdx.x == 0) {
            *o = __builtin_nan("");
            if constexpr (BIRK) *ob = __builtin_nan("");
        }
        return;
    }
    if constexpr (BIRK) {
        // wavefront sums; argmax_first's barrier publishes them
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) bacc = bacc + __shfl_xor(bacc, off);
        if ((threadIdx.x & 63) == 0) sb[threadIdx.x >> 6] = bacc;
    }
    // magnitudes and their first maximum; the interpolation recomputes the
    // three magnitudes it needs from the spectrum, which stays in LDS
    auto magnitude = [&](uint32_t k) {
        const Cplx z = buf[stk_pos(k)];
        return sqrt(z.re * z.re + z.im * z.im);
    };
    double best = -1.0;
    uint32_t best_i = 0xffffffffu;
#pragma unroll
    for (uint32_t q = 0; q < (N + kStockThreads - 1) / kStockThreads; ++q) {
        const uint32_t k = threadIdx.x + q * kStockThreads;
        if (k < N) {
            const double mg = magnitude(k);
            if (mg > best || best_i == 0xffffffffu) { best = mg; best_i = k; }
        }
    }
    uint32_t bi;
    argmax_first(best, best_i, sv, si, bi);
    if (threadIdx.x == 0) {
        *o = interp_tune_at(magnitude, N, bi);
        if constexpr (BIRK) {
            double acc = sb[0];
            for (uint32_t q = 1; q < kStockThreads / 64; ++q) acc = acc + sb[q];
            *ob = 1 - acc / (2 * M_PI);
        }
    }
}

// 2048-point windows (the chaos-indicator workload's), four-step with one
// wavefront per quarter: N = 4 x 512, index m = n2 + 512 n1 in, k = k1 + 4 k2
// out,
//   X[k1 + 4 k2] = sum_{n2} W_512^{n2 k2} [W_2048^{n2 k1} sum_{n1} x[n2 + 512 n1] W_4^{n1 k1}].
// Step 1, straight from the rows: thread t takes n2 = t and t + 256, the radix-4
// DFT over n1 and the twiddle, and stores entry k1 in quarter k1 of LDS. One
// barrier. Step 2: wavefront w runs the 512-point transform of quarter w (three
// radix-8 Stockham passes of one butterfly per lane) with no workgroup barrier:
// a wavefront's LDS operations take effect in program order. Step 3: the
// magnitudes, each wavefront's first maximum in the spectrum's order, one
// barrier, the interpolation. Two barriers against nine of the general kernel.
// BIRK: the Birkhoff tune from the same loads, as k_fft_tune_stockham<.., true>.
constexpr uint32_t kQuarter = 512 + 32 + 1;  // LDS entries of a quarter (stk_pos(511) + 1, odd)

template <bool BIRK>
__global__ void __launch_bounds__(kStockThreads) k_fft_tune_2048(FftRowArgs a)
{
    constexpr uint32_t N = 2048;
    __shared__ Cplx buf[4 * kQuarter];
    __shared__ double sv[kStockThreads / 64];
    __shared__ uint32_t si[kStockThreads / 64];
    __shared__ double sb[kStockThreads / 64];
    __shared__ int any_nan_flag;
    const uint32_t b = blockIdx.x >> 1, plane = blockIdx.x & 1;
    if (b >= a.nb) return;
    const double *base = a.rows + (size_t)((a.row0 + b) * a.pstride + a.from);
    const double *rx = base + (2 * plane) * a.cstride, *rp = base + (2 * plane + 1) * a.cstride;
    const double *oth0 = base + (2 * (1 - plane)) * a.cstride, *oth1 = base + (2 * (1 - plane) + 1) * a.cstride;
    const size_t nn = a.n, pp = a.p0 + b;
    const double mx = a.means ? a.means[(2 * plane) * nn + pp] : 0.0;
    const double mp = a.means ? a.means[(2 * plane + 1) * nn + pp] : 0.0;
    const double mo0 = a.means ? a.means[(2 * (1 - plane)) * nn + pp] : 0.0;
    const double mo1 = a.means ? a.means[(2 * (1 - plane) + 1) * nn + pp] : 0.0;
    double *o = a.out + ((size_t)a.window * 2 + plane) * nn + pp;
    double *ob = BIRK ? a.bout + ((size_t)a.window * 2 + plane) * nn + pp : nullptr;
    if (any_nan(mx, mp, mo0, mo1)) {
        // a NaN mean makes every sample of its coordinate NaN: nothing to read
        if (threadIdx.x == 0) {
            *o = __builtin_nan("");
            if constexpr (BIRK) *ob = __builtin_nan("");
        }
        return;  // the whole workgroup
    }
    if (threadIdx.x == 0) any_nan_flag = 0;
    __syncthreads();
    // four finite means: no sample of the particle is NaN, the other plane's
    // rows need not be read
    const bool scan_other = !a.means || !(isfinite(mx) && isfinite(mp) && isfinite(mo0) && isfinite(mo1));
    const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
    bool nan_seen = false;
    // ---- step 1: samples m = n2 + 512 n1 (n2 = t + 256 h), loads first
    Cplx v[2][4];
    double hw[2][4], bw[2][4];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int n1 = 0; n1 < 4; ++n1) {
            const uint32_t m = t + 256 * h + 512 * n1;
            v[h][n1] = Cplx{rx[m], rp[m]};
            hw[h][n1] = a.hann[m];
            bw[h][n1] = (BIRK && m > 0) ? a.bw[m - 1] : 0.0;
            if (scan_other) nan_seen |= any_nan(oth0[m] - mo0, oth1[m] - mo1, 0.0, 0.0);
        }
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int n1 = 0; n1 < 4; ++n1) {
            v[h][n1] = Cplx{v[h][n1].re - mx, v[h][n1].im - mp};
            nan_seen |= any_nan(v[h][n1].re, v[h][n1].im, 0.0, 0.0);
        }
    double bacc = 0.0;
    if constexpr (BIRK) {
        // the angles through LDS (not in use yet): the angle of sample m - 1
        double *th = reinterpret_cast<double *>(buf);
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int n1 = 0; n1 < 4; ++n1) th[t + 256 * h + 512 * n1] = atan2(v[h][n1].im, v[h][n1].re);
        __syncthreads();
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int n1 = 0; n1 < 4; ++n1) {
                const uint32_t m = t + 256 * h + 512 * n1;
                if (m > 0) {
                    double d = th[m] - th[m - 1];
                    if (d < 0.0) d += 2.0 * M_PI;
                    bacc = bacc + bw[h][n1] * d;
                }
            }
        __syncthreads();  // every angle read before the transform overwrites buf
    }
    if (nan_seen) atomicOr(&any_nan_flag, 1);
    // ---- step 1
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const uint32_t n2 = t + 256 * h;
        Cplx *x4 = v[h];
#pragma unroll
        for (int n1 = 0; n1 < 4; ++n1) x4[n1] = Cplx{x4[n1].re * hw[h][n1], x4[n1].im * hw[h][n1]};
        dft4(x4[0], x4[1], x4[2], x4[3]);
        // W_2048^{n2 k1}: one table load, its square and cube by products
        const Cplx w1{a.tw[2 * n2], a.tw[2 * n2 + 1]}, w2 = cmul(w1, w1);
        x4[1] = cmul(x4[1], w1);
        x4[2] = cmul(x4[2], w2);
        x4[3] = cmul(x4[3], cmul(w2, w1));
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) buf[k1 * kQuarter + stk_pos(n2)] = x4[k1];
    }
    __syncthreads();
    if (any_nan_flag) {
        if (t == 0) {
            *o = __builtin_nan("");
            if constexpr (BIRK) *ob = __builtin_nan("");
        }
        return;
    }
    // ---- step 2: wavefront w, the 512-point transform of quarter w
    Cplx *qb = buf + w * kQuarter;
    uint32_t Ns = 1;
#pragma unroll
    for (int pass = 0; pass < 3; ++pass) {
        const uint32_t j = lane, k = j % Ns;
        Cplx u[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) u[r] = qb[stk_pos(j + 64 * r)];
        if (k != 0) {
            // W_512^{k r 512 / (8 Ns)} = W_2048^{4 k r 64 / Ns}: one table load,
            // the higher powers by products (relative error below 8 ulps)
            const uint32_t step = 4 * k * (64 / Ns);
            const Cplx w1{a.tw[2 * step], a.tw[2 * step + 1]};
            Cplx wr = w1;
#pragma unroll
            for (int r = 1; r < 8; ++r) {
                u[r] = cmul(u[r], wr);
                if (r + 1 < 8) wr = cmul(wr, w1);
            }
        }
        dft_r<8>(u);
        __builtin_amdgcn_wave_barrier();  // this pass's reads before its writes
        const uint32_t dst = (j - k) * 8 + k;
#pragma unroll
        for (int r = 0; r < 8; ++r) qb[stk_pos(dst + r * Ns)] = u[r];
        __builtin_amdgcn_wave_barrier();  // the writes before the next pass's reads
        Ns *= 8;
    }
    // ---- step 3: quarter w holds X[w + 4 k2] at stk_pos(k2)
    auto magnitude = [&](uint32_t kk) {
        const Cplx z = buf[(kk & 3) * kQuarter + stk_pos(kk >> 2)];
        return sqrt(z.re * z.re + z.im * z.im);
    };
    double best = -1.0;
    uint32_t best_i = 0xffffffffu;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const uint32_t k2 = lane + 64 * q;
        const Cplx z = qb[stk_pos(k2)];
        const double mg = sqrt(z.re * z.re + z.im * z.im);
        const uint32_t kk = w + 4 * k2;  // ascending in q: the first of equal values stays
        if (mg > best || best_i == 0xffffffffu) { best = mg; best_i = kk; }
    }
    if constexpr (BIRK) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) bacc = bacc + __shfl_xor(bacc, off);
        if (lane == 0) sb[w] = bacc;
    }
    uint32_t bi;
    argmax_first(best, best_i, sv, si, bi);  // its barrier also publishes every quarter
    if (t == 0) {
        *o = interp_tune_at(magnitude, N, bi);
        if constexpr (BIRK) {
            double acc = sb[0];
            for (uint32_t q = 1; q < kStockThreads / 64; ++q) acc = acc + sb[q];
            *ob = 1 - acc / (2 * M_PI);
        }
    }
}

// Host tables per length N: Hanning window (dynamic_indicator.cpp:47-52, glibc)
// and twiddles exp(-2 pi i q / N) in long double, cached on the device.
struct FftTables {
    double *hann = nullptr, *tw = nullptr;
    double *bw = nullptr;  // Birkhoff weights of N - 1 differences (the fused kernel's)
};

static void host_birkhoff_weights(uint32_t n, double *out);

// Signal rows staged per launch pair (bytes); particles are taken in batches that fit.
static constexpr uint64_t kRowBudget = 4ull << 30;
// HM_FFT_ROW_BYTES overrides the budget (small values force several batches).
static uint64_t row_budget()
{
    const char *e = std::getenv("HM_FFT_ROW_BYTES");
    if (e && *e) {
        const unsigned long long v = std::strtoull(e, nullptr, 10);
        if (v > 0) return v;
    }
    return kRowBudget;
}

static FftTables fft_tables(uint32_t N, hipStream_t s)
{
    static std::mutex mu;
    static std::map<std::pair<int, uint32_t>, FftTables> cache;
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find({dev, N});
    if (it != cache.end()) return it->second;
    // the twiddles serve the power-of-two kernels only (the chirp-z transform
    // keeps its own per length)
    const bool pow2 = (N & (N - 1)) == 0 && N <= 8192;
    std::vector<double> h(N), tw(pow2 ? 2 * (size_t)N : 0);
    for (uint32_t i = 0; i < N; i++) h[i] = 2 * std::pow((std::sin(M_PI * i / (N))), 2);
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (uint32_t q = 0; pow2 && q < N; q++) {
        const long double ang = two_pi * (long double)q / (long double)N;
        tw[2 * q] = (double)cosl(ang);
        tw[2 * q + 1] = (double)-sinl(ang);
    }
    FftTables t;
    hip_check(hipMalloc(&t.hann, N * sizeof(double)), "hipMalloc(hann)");
    hip_check(hipMemcpyAsync(t.hann, h.data(), N * sizeof(double), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    if (pow2) {
        hip_check(hipMalloc(&t.tw, 2 * (size_t)N * sizeof(double)), "hipMalloc(twiddles)");
        hip_check(hipMemcpyAsync(t.tw, tw.data(), 2 * (size_t)N * sizeof(double), hipMemcpyHostToDevice, s),
                  "hipMemcpyAsync");
    }
    if (pow2 && N >= 8 && N <= 2048) {
        std::vector<double> bw(N - 1);
        host_birkhoff_weights(N - 1, bw.data());
        hip_check(hipMalloc(&t.bw, (N - 1) * sizeof(double)), "hipMalloc(Birkhoff weights)");
        hip_check(hipMemcpyAsync(t.bw, bw.data(), (N - 1) * sizeof(double), hipMemcpyHostToDevice, s),
                  "hipMemcpyAsync");
    }
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    cache[{dev, N}] = t;
    return t;
}

static void host_birkhoff_weights(uint32_t n, double *out)
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

}  // namespace hm

using namespace hm;

extern "C" int hm_track_history(double *x, double *px, double *y, double *py, uint32_t *steps, size_t n,
                                const double *table, size_t table_len, uint32_t n_turns, double mu, double barrier,
                                int inverse, double *hist, void *stream)
{
    return guarded([&] {
        if (n == 0) return;
        if (!x || !px || !y || !py || !steps || !table || !hist) throw Error(HM_ERR_ARG, "history: NULL pointer");
        hipStream_t s = as_stream(stream);
        const uint32_t grid = (uint32_t)((n + 255) / 256);
        const double b2 = barrier * barrier;
        const TabEntry *tab = reinterpret_cast<const TabEntry *>(table);
        LaunchTimer timer("k_history", s);
        const bool oct = mu != 0.0;
        if (inverse) {
            if (oct) hipLaunchKernelGGL((k_history<true, true>), dim3(grid), dim3(256), 0, s, x, px, y, py, steps,
                                        (uint32_t)n, tab, (uint32_t)table_len, n_turns, mu, b2, hist, status_word());
            else hipLaunchKernelGGL((k_history<true, false>), dim3(grid), dim3(256), 0, s, x, px, y, py, steps,
                                    (uint32_t)n, tab, (uint32_t)table_len, n_turns, mu, b2, hist, status_word());
        } else {
            if (oct) hipLaunchKernelGGL((k_history<false, true>), dim3(grid), dim3(256), 0, s, x, px, y, py, steps,
                                        (uint32_t)n, tab, (uint32_t)table_len, n_turns, mu, b2, hist, status_word());
            else hipLaunchKernelGGL((k_history<false, false>), dim3(grid), dim3(256), 0, s, x, px, y, py, steps,
                                    (uint32_t)n, tab, (uint32_t)table_len, n_turns, mu, b2, hist, status_word());
        }
        launch_check("k_history");
    });
}

extern "C" int hm_track_rows(double *x, double *px, double *y, double *py, uint32_t *steps, size_t n,
                             const double *table, size_t table_len, uint32_t n_turns, double mu, double barrier,
                             int inverse, double *rows, size_t row_len, double *means, void *stream)
{
    return guarded([&] {
        if (n == 0) return;
        if (!x || !px || !y || !py || !steps || !table || !rows) throw Error(HM_ERR_ARG, "history: NULL pointer");
        if (row_len % 16 != 0 || row_len < (size_t)n_turns + 1)
            throw Error(HM_ERR_ARG, "history rows: row_len must be a multiple of 16 of at least n_turns + 1");
        if (reinterpret_cast<uintptr_t>(rows) % 16 != 0) throw Error(HM_ERR_ARG, "history rows: rows not 16-byte aligned");
        if (n > UINT32_MAX) throw Error(HM_ERR_RANGE, "history rows: too many particles");
        hipStream_t s = as_stream(stream);
        const uint32_t grid = (uint32_t)((n + 255) / 256);
        const double b2 = barrier * barrier;
        const TabEntry *tab = reinterpret_cast<const TabEntry *>(table);
        LaunchTimer timer("k_history_rows", s);
        const bool oct = mu != 0.0, inv = inverse != 0, with_means = means != nullptr;
        using K = void (*)(double *, double *, double *, double *, uint32_t *, uint32_t, const TabEntry *, uint32_t,
                           uint32_t, double, double, double *, uint64_t, double *, uint32_t *);
        K k;
        if (inv) {
            if (oct) k = with_means ? k_history_rows<true, true, true> : k_history_rows<true, true, false>;
            else k = with_means ? k_history_rows<true, false, true> : k_history_rows<true, false, false>;
        } else {
            if (oct) k = with_means ? k_history_rows<false, true, true> : k_history_rows<false, true, false>;
            else k = with_means ? k_history_rows<false, false, true> : k_history_rows<false, false, false>;
        }
        hipLaunchKernelGGL(k, dim3(grid), dim3(256), 0, s, x, px, y, py, steps, (uint32_t)n, tab, (uint32_t)table_len,
                           n_turns, mu, b2, rows, (uint64_t)row_len, means, status_word());
        launch_check("k_history_rows");
    });
}

static void means_impl(const double *hist, HistLayout lay, size_t n, uint32_t n_samples, double *means,
                       hipStream_t s)
{
    if (n == 0) return;
    if (!hist || !means || n_samples == 0) throw Error(HM_ERR_ARG, "means: bad arguments");
    hipLaunchKernelGGL(k_means, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, hist, lay, (uint32_t)n, n_samples,
                       means);
    launch_check("k_means");
}

extern "C" int hm_history_means(const double *hist, size_t n, uint32_t n_samples, double *means, void *stream)
{
    return guarded([&] { means_impl(hist, sample_major(n), n, n_samples, means, as_stream(stream)); });
}

extern "C" int hm_rows_means(const double *rows, size_t n, size_t row_len, uint32_t n_samples, double *means,
                             void *stream)
{
    return guarded([&] {
        if (n_samples > row_len) throw Error(HM_ERR_RANGE, "means: more samples than row_len");
        means_impl(rows, particle_rows(row_len), n, n_samples, means, as_stream(stream));
    });
}

static void birkhoff_impl(const double *hist, HistLayout lay, size_t n, uint32_t n_samples, const double *means,
                          const uint32_t *from, const uint32_t *to, size_t n_windows, double *out, hipStream_t s)
{
    if (n == 0 || n_windows == 0) return;
    if (!hist || !out || !from || !to) throw Error(HM_ERR_ARG, "birkhoff: NULL pointer");
    if (n > 0xffffffffull) throw Error(HM_ERR_ARG, "birkhoff: too many particles");
    for (size_t w0 = 0; w0 < n_windows; w0 += kMaxWindows) {
        WindowSet ws{};
        ws.count = (uint32_t)std::min<size_t>(kMaxWindows, n_windows - w0);
        size_t total = 0;
        for (uint32_t w = 0; w < ws.count; w++) {
            const uint32_t f = from[w0 + w], t = to[w0 + w];
            if (f >= n_samples || t >= n_samples) throw Error(HM_ERR_RANGE, "compute_tune: from and to must be valid indices");
            if (t < f + 1) throw Error(HM_ERR_ARG, "birkhoff: a window needs at least two samples");
            ws.from[w] = f;
            ws.len[w] = t - f + 1;
            ws.w_off[w] = (uint32_t)total;
            total += ws.len[w] - 1;
        }
        std::vector<double> wts(total);
        for (uint32_t w = 0; w < ws.count; w++) host_birkhoff_weights(ws.len[w] - 1, wts.data() + ws.w_off[w]);
        double *dw = static_cast<double *>(workspace(std::max<size_t>(total, 1) * sizeof(double), 2));
        hip_check(hipMemcpyAsync(dw, wts.data(), total * sizeof(double), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
        LaunchTimer timer("k_birkhoff_hist", s);
        // particle ranges of at most 2^22 / windows: grids below 2^31 threads
        const uint32_t pb = std::max<uint32_t>(1, (1u << 22) / ws.count);
        for (uint32_t p0 = 0; p0 < (uint32_t)n; p0 += pb) {
            const uint32_t nb = std::min<uint32_t>(pb, (uint32_t)n - p0);
            if (lay.ms == 1) {
                // particle rows: a wavefront per (particle, plane, window)
                const uint64_t waves = (uint64_t)nb * 2 * ws.count;
                hipLaunchKernelGGL(k_birkhoff_rows, dim3((uint32_t)((waves + 3) / 4)), dim3(256), 0, s, hist, lay.cs,
                                   (uint32_t)n, p0, nb, means, ws, (const double *)dw, out + w0 * 2 * n);
            } else {
                hipLaunchKernelGGL(k_birkhoff_hist, dim3((uint32_t)(((uint64_t)nb * 2 * ws.count + 255) / 256)),
                                   dim3(256), 0, s, hist, lay, (uint32_t)n, p0, nb, means, ws, (const double *)dw,
                                   out + w0 * 2 * n);
            }
            launch_check("k_birkhoff_hist");
        }
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    }
}

extern "C" int hm_birkhoff_tunes_history(const double *hist, size_t n, uint32_t n_samples, const double *means,
                                         const uint32_t *from, const uint32_t *to, size_t n_windows, double *out,
                                         void *stream)
{
    return guarded([&] {
        birkhoff_impl(hist, sample_major(n), n, n_samples, means, from, to, n_windows, out, as_stream(stream));
    });
}

extern "C" int hm_birkhoff_tunes_rows(const double *rows, size_t n, size_t row_len, uint32_t n_samples,
                                      const double *means, const uint32_t *from, const uint32_t *to, size_t n_windows,
                                      double *out, void *stream)
{
    return guarded([&] {
        if (n_samples > row_len) throw Error(HM_ERR_RANGE, "birkhoff: more samples than row_len");
        birkhoff_impl(rows, particle_rows(row_len), n, n_samples, means, from, to, n_windows, out, as_stream(stream));
    });
}

// Launches the FFT of a batch of signal rows: the Stockham kernel for
// 8 <= N <= 1024 and the four-step kernel for 2048 (the signal fits the static
// LDS), else the radix-2 row kernel.
// Particles per launch of the row kernels (one or two workgroups of 256 threads
// each): grids stay below 2^31 threads.
static constexpr uint32_t kRowLaunchParticles = 1u << 22;

template <bool BIRK>
static void launch_stockham_part(const FftRowArgs &r, uint32_t logn, hipStream_t s)
{
    const dim3 g2(2 * r.nb);
    switch (logn) {
    case 3: hipLaunchKernelGGL((k_fft_tune_stockham<3, BIRK>), g2, dim3(kStockThreads), 0, s, r); break;
    case 4: hipLaunchKernelGGL((k_fft_tune_stockham<4, BIRK>), g2, dim3(kStockThreads), 0, s, r); break;
    case 5: hipLaunchKernelGGL((k_fft_tune_stockham<5, BIRK>), g2, dim3(kStockThreads), 0, s, r); break;
    case 6: hipLaunchKernelGGL((k_fft_tune_stockham<6, BIRK>), g2, dim3(kStockThreads), 0, s, r); break;
    case 7: hipLaunchKernelGGL((k_fft_tune_stockham<7, BIRK>), g2, dim3(kStockThreads), 0, s, r); break;
    case 8: hipLaunchKernelGGL((k_fft_tune_stockham<8, BIRK>), g2, dim3(kStockThreads), 0, s, r); break;
    case 9: hipLaunchKernelGGL((k_fft_tune_stockham<9, BIRK>), g2, dim3(kStockThreads), 0, s, r); break;
    case 10: hipLaunchKernelGGL((k_fft_tune_stockham<10, BIRK>), g2, dim3(kStockThreads), 0, s, r); break;
    default: hipLaunchKernelGGL((k_fft_tune_2048<BIRK>), g2, dim3(kStockThreads), 0, s, r); break;
    }
}

// The launches of `r` in particle ranges of at most kRowLaunchParticles.
template <class F>
static void for_row_parts(const FftRowArgs &r, F launch)
{
    for (uint32_t b0 = 0; b0 < r.nb; b0 += kRowLaunchParticles) {
        FftRowArgs q = r;
        q.row0 = r.row0 + b0;
        q.p0 = r.p0 + b0;
        q.nb = std::min<uint32_t>(kRowLaunchParticles, r.nb - b0);
        launch(q);
    }
}

template <bool BIRK>
static void launch_stockham(const FftRowArgs &r, uint32_t logn, hipStream_t s)
{
    for_row_parts(r, [&](const FftRowArgs &q) { launch_stockham_part<BIRK>(q, logn, s); });
}

// Power-of-two windows of 2 .. 2048 samples (the Stockham kernels; the timer
// keeps the name the row kernels had).
static void launch_row_fft(const FftRowArgs &r, uint32_t N, hipStream_t s)
{
    const uint32_t logn = 31 - __builtin_clz(N);
    LaunchTimer timer("k_fft_tune_rows", s);
    launch_stockham<false>(r, logn, s);
    launch_check("k_fft_tune_rows");
}

// rows_input: hist is particle rows (row_len = lay.cs); else sample-major.
static void fft_impl(const double *hist, HistLayout lay, bool rows_input, size_t n, uint32_t n_samples,
                     const double *means, const uint32_t *from, const uint32_t *len, size_t n_windows, double *out,
                     hipStream_t s)
{
    if (n == 0 || n_windows == 0) return;
    if (!hist || !out || !from || !len) throw Error(HM_ERR_ARG, "fft_tunes: NULL pointer");
    if (n > 0x3fffffffull) throw Error(HM_ERR_ARG, "fft_tunes: too many particles");
    // group windows by length: one launch per (length, up to 64 windows)
    std::map<uint32_t, std::vector<uint32_t>> by_len;
    for (size_t w = 0; w < n_windows; w++) {
        if (len[w] < 2) throw Error(HM_ERR_ARG, "fft_tunes: a window needs at least two samples");
        if ((uint64_t)from[w] + len[w] > n_samples)
            throw Error(HM_ERR_RANGE, "compute_tune: from and to must be valid indices");
        by_len[len[w]].push_back((uint32_t)w);
    }
    for (auto &kv : by_len) {
        const uint32_t N = kv.first;
        const FftTables t = fft_tables(N, s);
        // powers of two from 8 to 2048: the Stockham kernels; every other
        // length (2, 4, 4096 and 8192 included) the whole-signal or four-step
        // transforms of fft_tune_mixed, past 2^21 samples power-of-two chirp-z
        const bool stockham = (N & (N - 1)) == 0 && N >= 8 && N <= 2048;
        const auto &ws = kv.second;
        if (!stockham) {
            for (uint32_t w : ws)
                if (!fft_tune_mixed(hist, lay, (uint32_t)n, means, from[w], N, t.hann, out, w, s))
                    fft_tune_chirpz(hist, lay, (uint32_t)n, means, from[w], N, t.hann, out, w, s);
            continue;
        }
        if (rows_input) {
            // one workgroup per particle and plane, reading its rows in place
            for (uint32_t w : ws) {
                FftRowArgs r{hist, lay.ps, lay.cs, 0, from[w], means, (uint32_t)n, 0, (uint32_t)n, N,
                             t.hann, t.tw, out, w};
                launch_row_fft(r, N, s);
            }
            continue;
        }
        // sample-major history, per window: signal rows for a batch of
        // particles (corner turn), then the row kernels reading them
        const uint64_t row_bytes = 4ull * N * sizeof(double);
        const uint32_t batch = (uint32_t)std::min<uint64_t>(n, std::max<uint64_t>(64, row_budget() / row_bytes));
        double *rows = static_cast<double *>(workspace((size_t)batch * row_bytes, 8));
        for (uint32_t w : ws) {
            for (uint32_t p0 = 0; p0 < (uint32_t)n; p0 += batch) {
                const uint32_t nb = std::min<uint32_t>(batch, (uint32_t)n - p0);
                {
                    LaunchTimer timer("k_signal_rows", s);
                    hipLaunchKernelGGL(k_signal_rows, dim3((nb + 63) / 64, (N + 63) / 64, 4), dim3(256), 0, s, hist,
                                       (uint32_t)n, means, from[w], N, p0, nb, rows);
                    launch_check("k_signal_rows");
                }
                FftRowArgs r{rows, 4ull * N, N, 0, 0, nullptr, nb, p0, (uint32_t)n, N, t.hann, t.tw, out, w};
                launch_row_fft(r, N, s);
            }
        }
    }
}

extern "C" int hm_fft_tunes(const double *hist, size_t n, uint32_t n_samples, const double *means,
                            const uint32_t *from, const uint32_t *len, size_t n_windows, double *out, void *stream)
{
    return guarded([&] {
        fft_impl(hist, sample_major(n), false, n, n_samples, means, from, len, n_windows, out, as_stream(stream));
    });
}

extern "C" int hm_fft_tunes_rows(const double *rows, size_t n, size_t row_len, uint32_t n_samples,
                                 const double *means, const uint32_t *from, const uint32_t *len, size_t n_windows,
                                 double *out, void *stream)
{
    return guarded([&] {
        if (n_samples > row_len) throw Error(HM_ERR_RANGE, "fft_tunes: more samples than row_len");
        fft_impl(rows, particle_rows(row_len), true, n, n_samples, means, from, len, n_windows, out,
                 as_stream(stream));
    });
}

extern "C" int hm_tunes_rows(const double *rows, size_t n, size_t row_len, uint32_t n_samples, const double *means,
                             const uint32_t *from, const uint32_t *to, size_t n_windows, double *birkhoff_out,
                             double *fft_out, void *stream)
{
    return guarded([&] {
        if (n == 0 || n_windows == 0) return;
        if (!rows || !from || !to || !birkhoff_out || !fft_out) throw Error(HM_ERR_ARG, "tunes: NULL pointer");
        if (n_samples > row_len) throw Error(HM_ERR_RANGE, "tunes: more samples than row_len");
        if (n > 0x3fffffffull) throw Error(HM_ERR_ARG, "tunes: too many particles");
        hipStream_t s = as_stream(stream);
        const HistLayout lay = particle_rows(row_len);
        for (size_t w = 0; w < n_windows; w++) {
            const uint32_t f = from[w], t = to[w];
            if (f >= n_samples || t >= n_samples) throw Error(HM_ERR_RANGE, "compute_tune: from and to must be valid indices");
            if (t < f + 1) throw Error(HM_ERR_ARG, "tunes: a window needs at least two samples");
            const uint32_t N = t - f + 1;
            if ((N & (N - 1)) == 0 && N >= 8 && N <= 2048) {
                // both tunes from one read of the window
                const FftTables tb = fft_tables(N, s);
                FftRowArgs r{rows, lay.ps, lay.cs, 0, f, means, (uint32_t)n, 0, (uint32_t)n, N, tb.hann, tb.tw,
                             fft_out, (uint32_t)w, tb.bw, birkhoff_out};
                LaunchTimer timer("k_tunes_fused", s);
                launch_stockham<true>(r, 31 - __builtin_clz(N), s);
                launch_check("k_tunes_fused");
            } else {
                birkhoff_impl(rows, lay, n, n_samples, means, &from[w], &to[w], 1, birkhoff_out + w * 2 * n, s);
                fft_impl(rows, lay, true, n, n_samples, means, &from[w], &N, 1, fft_out + w * 2 * n, s);
            }
        }
    });
}
