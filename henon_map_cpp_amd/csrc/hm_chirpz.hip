// FFT tunes of windows of any length (src/dynamic_indicator.cpp:32-61,
// 178-196, 255-331, where FFTW transforms any N): Bluestein's chirp-z
// transform. With c_m = exp(-i pi m^2 / N),
//   X_k = sum_m x_m W_N^{mk} = c_k sum_m (x_m c_m) conj(c_{k-m}),
// a cyclic convolution of length M = 2^k >= 2N - 1, taken as
//   conv = IDFT_M(DFT_M(a) . B),   a_m = x_m c_m (m < N, else 0),
//   B = DFT_M(b),  b_j = conj(c_j) for |j| < N (j mod M), else 0,
// so |X_k| = |conv_k| / M. The inverse transform is the forward one applied to
// the conjugate, and the conjugation is carried through the second half, so
// only forward transforms run. Each length-M transform is split M = M1 x M2
// (index n = M2 n1 + n2 in, k = k1 + M1 k2 out):
//   X[k1 + M1 k2] = sum_{n2} W_{M2}^{n2 k2} W_M^{n2 k1} sum_{n1} x[M2 n1 + n2] W_{M1}^{n1 k1}
// Four launches per batch of signals, every sub-transform held in LDS:
//   k_cz_columns : load, window and chirp the samples of a few columns n2,
//                  M1-point transforms, twiddle W_M^{n2 k1}, store rows k1
//   k_cz_rows    : per row k1: M2-point transform (the spectrum), times B,
//                  conjugate, M2-point transform, twiddle, store
//   k_cz_final   : per column n2: M1-point transform, magnitudes / M of the
//                  N wanted entries
//   k_cz_tune    : per signal: first maximum of the magnitudes and the
//                  interpolated tune (dynamic_indicator.cpp:139-176)
// HBM traffic per signal: the samples once, 3 x 16 M bytes of scratch each
// way, 8 N bytes of magnitudes each way.
#include <hip/hip_runtime.h>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>

#include "hm_device.h"
#include "hm_internal.h"
#include "hm_tune.h"

namespace hm {

constexpr uint32_t kCzThreads = 256;
// Largest length N of a window (M = 2^26 at most: both factors fit an 8192-point
// workgroup transform).
constexpr uint32_t kCzMaxN = 1u << 25;

struct CzArgs {
    const double *hist;
    HistLayout lay;
    const double *means;  // [4][n] or NULL
    uint32_t n, p0;       // particles in total, first particle of the batch
    uint32_t nsig;        // signals of the batch: particle (p0 + sig / 2), plane sig % 2
    uint32_t from, N;
    uint32_t logM, logM1, logM2;
    uint32_t logC;        // columns per workgroup (k_cz_columns / k_cz_final)
    uint32_t logRR;       // rows per workgroup (k_cz_rows)
    const double *hann;   // [N]
    const Cplx *tw;       // [M]: exp(-2 pi i q / M)
    const Cplx *chirp;    // [N]: exp(-i pi (m^2 mod 2N) / N)
    const Cplx *bhat;     // [M1][M2]: B[k1 + M1 k2] at k1 M2 + k2
    const Cplx *twm;      // [M1][M2]: W_M^{k1 n2} at k1 M2 + n2 (the step twiddles, coalesced)
    Cplx *Y;              // scratch [nsig][M1][M2]
    double *mags;         // scratch [nsig][N]
    uint32_t *flags;      // [nsig]: 1 when a sample of the window is NaN
    double *out;          // [window][plane][n]
    uint32_t window;
};

// Workgroup b's virtual index: the hardware hands workgroup b to XCD b % 8, so
// XCD x gets the contiguous range [x per, (x+1) per) and a signal's workgroups
// share one L2. The grid is 8 per workgroups.
__device__ __forceinline__ uint32_t xcd_virtual(uint32_t b, uint32_t per) { return (b % 8) * per + b / 8; }

// One radix-R Stockham pass over G signals of length L = 2^logL in LDS (signal
// g at buf[g GS ..]); Ns = product of the earlier radices. Butterfly j of a
// signal reads x[j + r L/R], twiddles by W_L^{k r L/(Ns R)} (k = j mod Ns),
// takes the R-point DFT and writes element r to (j - k) R + k + r Ns.
// W_L^q = tw[q M / L].
template <int R, uint32_t PMAX>
__device__ __forceinline__ void cz_pass(Cplx *buf, uint32_t G, uint32_t logL, uint32_t GS, uint32_t Ns,
                                        uint32_t lognsr, const Cplx *__restrict__ tw, uint32_t logM)
{
    constexpr uint32_t PER = (PMAX / R + kCzThreads - 1) / kCzThreads;
    const uint32_t logB = logL - (R == 8 ? 3 : R == 4 ? 2 : 1);
    const uint32_t B = 1u << logB, total = G << logB;
    Cplx v[PER][R];
#pragma unroll
    for (uint32_t q = 0; q < PER; ++q) {
        const uint32_t J = threadIdx.x + q * kCzThreads;
        if (J < total) {
            const uint32_t g = J >> logB, j = J & (B - 1);
            const Cplx *src = buf + g * GS + j;
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = src[r * B];
            const uint32_t k = j & (Ns - 1);
            if (k != 0) {
                // one table load; the higher powers by products (relative
                // error below R ulps)
                const Cplx w1 = tw[k << (logM - lognsr)];
                Cplx w = w1;
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    v[q][r] = cmul(v[q][r], w);
                    if (r + 1 < R) w = cmul(w, w1);
                }
            }
            dft_r<R>(v[q]);
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < PER; ++q) {
        const uint32_t J = threadIdx.x + q * kCzThreads;
        if (J < total) {
            const uint32_t g = J >> logB, j = J & (B - 1), k = j & (Ns - 1);
            Cplx *dst = buf + g * GS + (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) dst[r * Ns] = v[q][r];
        }
    }
    __syncthreads();
}

// Forward DFT of G signals of length 2^logL in LDS: radix-8 passes, then one
// radix-4 or radix-2 pass; the result is in natural order.
template <uint32_t PMAX>
__device__ void cz_fft(Cplx *buf, uint32_t G, uint32_t logL, uint32_t GS, const Cplx *__restrict__ tw,
                       uint32_t logM)
{
    uint32_t Ns = 1, logNs = 0;
    while (logNs < logL) {
        const uint32_t rem = logL - logNs;
        if (rem >= 3) {
            cz_pass<8, PMAX>(buf, G, logL, GS, Ns, logNs + 3, tw, logM);
            logNs += 3;
        } else if (rem == 2) {
            cz_pass<4, PMAX>(buf, G, logL, GS, Ns, logNs + 2, tw, logM);
            logNs += 2;
        } else {
            cz_pass<2, PMAX>(buf, G, logL, GS, Ns, logNs + 1, tw, logM);
            logNs += 1;
        }
        Ns = 1u << logNs;
    }
}

// Columns n2 in [c0, c0 + C) of signal `sig`; LDS holds column c at
// buf[c (M1 + 1) ..] (one pad entry against bank conflicts).
template <uint32_t PMAX>
__global__ void __launch_bounds__(kCzThreads, PMAX <= 2048 ? 2 : 1) k_cz_columns(CzArgs a)
{
    extern __shared__ Cplx buf[];
    const uint32_t per_sig = 1u << (a.logM2 - a.logC);
    const uint32_t total = a.nsig * per_sig;
    const uint32_t v = xcd_virtual(blockIdx.x, gridDim.x / 8);
    if (v >= total) return;
    const uint32_t sig = v / per_sig, c0 = (v % per_sig) << a.logC;
    const uint32_t C = 1u << a.logC, M1 = 1u << a.logM1, GS = M1 + 1;
    const uint32_t particle = a.p0 + sig / 2, plane = sig % 2;
    const size_t nn = a.n;
    double mc[4] = {0.0, 0.0, 0.0, 0.0};
    if (a.means)
        for (int c = 0; c < 4; ++c) mc[c] = a.means[c * nn + particle];
    if (any_nan(mc[0], mc[1], mc[2], mc[3])) {
        // a NaN mean makes every sample of its coordinate NaN: nothing to read
        if (threadIdx.x == 0) atomicOr(a.flags + sig, 1u);
        return;
    }
    // four finite means prove that no sample of the particle is NaN or
    // infinite: the other plane need not be read
    const bool scan_other =
        !a.means || !(isfinite(mc[0]) && isfinite(mc[1]) && isfinite(mc[2]) && isfinite(mc[3]));
    const int cx = 2 * plane, cp = cx + 1, co0 = 2 * (1 - plane), co1 = co0 + 1;
    bool nan_seen = false;
    for (uint32_t e = threadIdx.x; e < (C << a.logM1); e += kCzThreads) {
        const uint32_t c = e & (C - 1), n1 = e >> a.logC;
        const uint32_t m = (n1 << a.logM2) + c0 + c;
        Cplx z{0.0, 0.0};
        if (m < a.N) {
            const uint64_t sm = (uint64_t)a.from + m;
            const double vx = a.hist[hat(a.lay, sm, cx, particle)] - mc[cx];
            const double vp = a.hist[hat(a.lay, sm, cp, particle)] - mc[cp];
            nan_seen |= __builtin_isnan(vx) || __builtin_isnan(vp);
            if (scan_other)
                nan_seen |= __builtin_isnan(a.hist[hat(a.lay, sm, co0, particle)] - mc[co0]) ||
                            __builtin_isnan(a.hist[hat(a.lay, sm, co1, particle)] - mc[co1]);
            const double h = a.hann[m];
            z = cmul(Cplx{vx * h, vp * h}, a.chirp[m]);
        }
        buf[c * GS + n1] = z;
    }
    if (__any(nan_seen) && (threadIdx.x & 63) == 0) atomicOr(a.flags + sig, 1u);
    __syncthreads();
    cz_fft<PMAX>(buf, C, a.logM1, GS, a.tw, a.logM);
    Cplx *Y = a.Y + ((size_t)sig << a.logM);
    for (uint32_t e = threadIdx.x; e < (C << a.logM1); e += kCzThreads) {
        const uint32_t c = e & (C - 1), k1 = e >> a.logC, n2 = c0 + c;
        const size_t o = ((size_t)k1 << a.logM2) + n2;
        Y[o] = cmul(buf[c * GS + k1], a.twm[o]);
    }
}

// Rows k1 in [r0, r0 + RR) of signal `sig`, contiguous in LDS.
template <uint32_t PMAX>
__global__ void __launch_bounds__(kCzThreads, PMAX <= 2048 ? 2 : 1) k_cz_rows(CzArgs a)
{
    extern __shared__ Cplx buf[];
    const uint32_t per_sig = 1u << (a.logM1 - a.logRR);
    const uint32_t total = a.nsig * per_sig;
    const uint32_t v = xcd_virtual(blockIdx.x, gridDim.x / 8);
    if (v >= total) return;
    const uint32_t sig = v / per_sig, r0 = (v % per_sig) << a.logRR;
    const uint32_t P = 1u << (a.logRR + a.logM2), M2 = 1u << a.logM2;
    Cplx *Y = a.Y + ((size_t)sig << a.logM) + ((size_t)r0 << a.logM2);
    const Cplx *bh = a.bhat + ((size_t)r0 << a.logM2);
    for (uint32_t e = threadIdx.x; e < P; e += kCzThreads) buf[e] = Y[e];
    __syncthreads();
    cz_fft<PMAX>(buf, 1u << a.logRR, a.logM2, M2, a.tw, a.logM);
    // spectrum entry k1 + M1 k2 times B, conjugated: the inverse transform
    // becomes a forward one
    for (uint32_t e = threadIdx.x; e < P; e += kCzThreads) {
        const Cplx z = cmul(buf[e], bh[e]);
        buf[e] = Cplx{z.re, -z.im};
    }
    __syncthreads();
    cz_fft<PMAX>(buf, 1u << a.logRR, a.logM2, M2, a.tw, a.logM);
    const Cplx *tm = a.twm + ((size_t)r0 << a.logM2);
    for (uint32_t e = threadIdx.x; e < P; e += kCzThreads) Y[e] = cmul(buf[e], tm[e]);
}

// Columns n2 in [c0, c0 + C): the last M1-point transforms and the magnitudes
// of entries m = n2 + M2 n1 < N.
template <uint32_t PMAX>
__global__ void __launch_bounds__(kCzThreads, PMAX <= 2048 ? 2 : 1) k_cz_final(CzArgs a)
{
    extern __shared__ Cplx buf[];
    const uint32_t per_sig = 1u << (a.logM2 - a.logC);
    const uint32_t total = a.nsig * per_sig;
    const uint32_t v = xcd_virtual(blockIdx.x, gridDim.x / 8);
    if (v >= total) return;
    const uint32_t sig = v / per_sig, c0 = (v % per_sig) << a.logC;
    const uint32_t C = 1u << a.logC, M1 = 1u << a.logM1, GS = M1 + 1;
    const Cplx *Y = a.Y + ((size_t)sig << a.logM);
    for (uint32_t e = threadIdx.x; e < (C << a.logM1); e += kCzThreads) {
        const uint32_t c = e & (C - 1), k1 = e >> a.logC;
        buf[c * GS + k1] = Y[((size_t)k1 << a.logM2) + c0 + c];
    }
    __syncthreads();
    cz_fft<PMAX>(buf, C, a.logM1, GS, a.tw, a.logM);
    const double inv_m = ldexp(1.0, -(int)a.logM);  // exact
    double *mags = a.mags + (size_t)sig * a.N;
    for (uint32_t e = threadIdx.x; e < (C << a.logM1); e += kCzThreads) {
        const uint32_t c = e & (C - 1), n1 = e >> a.logC;
        const uint32_t m = c0 + c + (n1 << a.logM2);
        if (m < a.N) {
            const Cplx z = buf[c * GS + n1];
            mags[m] = sqrt(z.re * z.re + z.im * z.im) * inv_m;
        }
    }
}

// One workgroup per signal: first maximum of the magnitudes, interpolated tune.
__global__ void __launch_bounds__(kCzThreads) k_cz_tune(CzArgs a)
{
    __shared__ double sv[kCzThreads / 64];
    __shared__ uint32_t si[kCzThreads / 64];
    const uint32_t sig = blockIdx.x;
    if (sig >= a.nsig) return;
    const uint32_t particle = a.p0 + sig / 2, plane = sig % 2;
    double *o = a.out + ((size_t)a.window * 2 + plane) * a.n + particle;
    if (a.flags[sig]) {
        if (threadIdx.x == 0) *o = __builtin_nan("");
        return;
    }
    const double *mags = a.mags + (size_t)sig * a.N;
    double v = -1.0;
    uint32_t i = 0xffffffffu;
    for (uint32_t k = threadIdx.x; k < a.N; k += kCzThreads) {
        const double mg = mags[k];
        if (mg > v || i == 0xffffffffu) { v = mg; i = k; }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double ov = __shfl_xor(v, off);
        const uint32_t oi = __shfl_xor(i, off);
        if (ov > v || (ov == v && oi < i)) { v = ov; i = oi; }
    }
    if ((threadIdx.x & 63) == 0) { sv[threadIdx.x >> 6] = v; si[threadIdx.x >> 6] = i; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double bv = sv[0];
        uint32_t bi = si[0];
        for (uint32_t q = 1; q < kCzThreads / 64; ++q)
            if (sv[q] > bv || (sv[q] == bv && si[q] < bi)) { bv = sv[q]; bi = si[q]; }
        *o = interp_tune_at([&](uint32_t k) { return mags[k]; }, a.N, bi);
    }
}

__global__ void k_cz_zero(uint32_t *__restrict__ p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------- host plan
struct CzPlan {
    uint32_t logM = 0, logM1 = 0, logM2 = 0;
    Cplx *tw = nullptr, *bhat = nullptr, *chirp = nullptr, *twm = nullptr;
};

// In-place forward DFT (radix 2, decimation in time) in long double, for the
// filter spectrum B; roots from cosl/sinl.
static void host_fft(std::vector<std::complex<long double>> &x)
{
    const size_t M = x.size();
    for (size_t i = 1, j = 0; i < M; ++i) {
        size_t bit = M >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(x[i], x[j]);
    }
    const long double two_pi = 6.283185307179586476925286766559005768L;
    std::vector<std::complex<long double>> root(M / 2);
    for (size_t q = 0; q < M / 2; ++q) {
        const long double ang = two_pi * (long double)q / (long double)M;
        root[q] = std::complex<long double>(cosl(ang), -sinl(ang));
    }
    for (size_t len = 2; len <= M; len <<= 1) {
        const size_t half = len / 2, step = M / len;
        for (size_t i = 0; i < M; i += len)
            for (size_t k = 0; k < half; ++k) {
                const std::complex<long double> t = x[i + k + half] * root[k * step];
                x[i + k + half] = x[i + k] - t;
                x[i + k] += t;
            }
    }
}

static CzPlan cz_plan(uint32_t N, hipStream_t s)
{
    static std::mutex mu;
    static std::map<std::pair<int, uint32_t>, CzPlan> cache;
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find({dev, N});
    if (it != cache.end()) return it->second;
    CzPlan p;
    while ((1ull << p.logM) < 2ull * N - 1) ++p.logM;
    if (p.logM < 2) p.logM = 2;
    p.logM1 = p.logM / 2;
    p.logM2 = p.logM - p.logM1;
    const size_t M = size_t(1) << p.logM, M1 = size_t(1) << p.logM1, M2 = size_t(1) << p.logM2;
    const long double pi = 3.141592653589793238462643383279502884L;
    std::vector<Cplx> tw(M), chirp(N), bhat(M);
    for (size_t q = 0; q < M; ++q) {
        const long double ang = 2 * pi * (long double)q / (long double)M;
        tw[q] = Cplx{(double)cosl(ang), (double)-sinl(ang)};
    }
    std::vector<std::complex<long double>> b(M, std::complex<long double>(0, 0));
    for (uint64_t m = 0; m < N; ++m) {
        const uint64_t q = (m * m) % (2ull * N);  // m < 2^25: no overflow
        const long double ang = pi * (long double)q / (long double)N;
        const long double c = cosl(ang), sn = sinl(ang);
        chirp[m] = Cplx{(double)c, (double)-sn};
        b[m] = std::complex<long double>(c, sn);
        if (m > 0) b[M - m] = b[m];
    }
    host_fft(b);
    for (size_t k1 = 0; k1 < M1; ++k1)
        for (size_t k2 = 0; k2 < M2; ++k2) {
            const std::complex<long double> z = b[k1 + M1 * k2];
            bhat[k1 * M2 + k2] = Cplx{(double)z.real(), (double)z.imag()};
        }
    // step twiddles W_M^{k1 n2} (k1 < M1, n2 < M2), reduced exactly mod M
    std::vector<Cplx> twm(M);
    for (size_t k1 = 0; k1 < M1; ++k1)
        for (size_t n2 = 0; n2 < M2; ++n2) twm[k1 * M2 + n2] = tw[(k1 * n2) & (M - 1)];
    hip_check(hipMalloc(&p.twm, M * sizeof(Cplx)), "hipMalloc(chirp-z step twiddles)");
    hip_check(hipMemcpyAsync(p.twm, twm.data(), M * sizeof(Cplx), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    hip_check(hipMalloc(&p.tw, M * sizeof(Cplx)), "hipMalloc(chirp-z twiddles)");
    hip_check(hipMalloc(&p.bhat, M * sizeof(Cplx)), "hipMalloc(chirp-z filter)");
    hip_check(hipMalloc(&p.chirp, N * sizeof(Cplx)), "hipMalloc(chirp)");
    hip_check(hipMemcpyAsync(p.tw, tw.data(), M * sizeof(Cplx), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    hip_check(hipMemcpyAsync(p.bhat, bhat.data(), M * sizeof(Cplx), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    hip_check(hipMemcpyAsync(p.chirp, chirp.data(), N * sizeof(Cplx), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    cache[{dev, N}] = p;
    return p;
}

// Scratch per batch (bytes); HM_FFT_ROW_BYTES overrides it, as for the signal
// rows of the power-of-two kernels.
static uint64_t cz_budget()
{
    const char *e = std::getenv("HM_FFT_ROW_BYTES");
    if (e && *e) {
        const unsigned long long v = std::strtoull(e, nullptr, 10);
        if (v > 0) return v;
    }
    return 8ull << 30;  // a few percent of the device memory: large batches, few launches
}

template <uint32_t PMAX>
static void cz_launch(const CzArgs &a, size_t lds_cols, size_t lds_rows, hipStream_t s)
{
    const uint32_t cols = a.nsig << (a.logM2 - a.logC), rows = a.nsig << (a.logM1 - a.logRR);
    const uint32_t gc = (cols + 7) / 8 * 8, gr = (rows + 7) / 8 * 8;
    hip_check(hipFuncSetAttribute((const void *)k_cz_columns<PMAX>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds_cols), "hipFuncSetAttribute");
    hip_check(hipFuncSetAttribute((const void *)k_cz_rows<PMAX>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds_rows), "hipFuncSetAttribute");
    hip_check(hipFuncSetAttribute((const void *)k_cz_final<PMAX>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds_cols), "hipFuncSetAttribute");
    {
        LaunchTimer timer("k_cz_columns", s);
        hipLaunchKernelGGL(k_cz_columns<PMAX>, dim3(gc), dim3(kCzThreads), lds_cols, s, a);
        launch_check("k_cz_columns");
    }
    {
        LaunchTimer timer("k_cz_rows", s);
        hipLaunchKernelGGL(k_cz_rows<PMAX>, dim3(gr), dim3(kCzThreads), lds_rows, s, a);
        launch_check("k_cz_rows");
    }
    {
        LaunchTimer timer("k_cz_final", s);
        hipLaunchKernelGGL(k_cz_final<PMAX>, dim3(gc), dim3(kCzThreads), lds_cols, s, a);
        launch_check("k_cz_final");
    }
}

void fft_tune_chirpz(const double *hist, HistLayout lay, uint32_t n, const double *means, uint32_t from,
                     uint32_t N, const double *hann, double *out, uint32_t window, hipStream_t s)
{
    if (N < 2) throw Error(HM_ERR_ARG, "fft_tunes: a window needs at least two samples");
    if (N > kCzMaxN) throw Error(HM_ERR_UNSUPPORTED, "fft_tunes: window longer than 33554432 samples");
    if (n == 0) return;
    const CzPlan p = cz_plan(N, s);
    // 2048-point workgroups (32 KiB of LDS, four resident per CU) while both
    // factors fit, else 8192
    const uint32_t PMAX = std::max(p.logM1, p.logM2) <= 11 ? 2048u : 8192u;
    const uint32_t logPM = PMAX == 2048 ? 11 : 13;
    CzArgs a{};
    a.hist = hist; a.lay = lay; a.means = means; a.n = n; a.from = from; a.N = N;
    a.logM = p.logM; a.logM1 = p.logM1; a.logM2 = p.logM2;
    a.logC = std::min(p.logM2, logPM - p.logM1);
    a.logRR = std::min(p.logM1, logPM - p.logM2);
    a.hann = hann; a.tw = p.tw; a.chirp = p.chirp; a.bhat = p.bhat; a.twm = p.twm; a.out = out; a.window = window;
    const size_t lds_cols = ((size_t)1 << a.logC) * (((size_t)1 << p.logM1) + 1) * sizeof(Cplx);
    const size_t lds_rows = ((size_t)1 << (a.logRR + p.logM2)) * sizeof(Cplx);
    // scratch per particle (both planes): the M-point work arrays, magnitudes, flags
    const uint64_t per_particle = 2 * ((uint64_t)sizeof(Cplx) << p.logM) + 2 * 8ull * N + 8;
    // at most 2^22 particles per batch keeps every grid below 2^24 workgroups
    const uint32_t batch =
        (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>({(uint64_t)n, cz_budget() / per_particle, 1ull << 22}));
    const size_t nsig_max = 2 * (size_t)batch;
    const size_t y_bytes = nsig_max * ((size_t)sizeof(Cplx) << p.logM);
    const size_t mag_bytes = nsig_max * (size_t)N * 8;
    char *ws = static_cast<char *>(workspace(y_bytes + mag_bytes + nsig_max * 4 + 256, 3));
    a.Y = reinterpret_cast<Cplx *>(ws);
    a.mags = reinterpret_cast<double *>(ws + y_bytes);
    a.flags = reinterpret_cast<uint32_t *>(ws + y_bytes + mag_bytes);
    for (uint32_t p0 = 0; p0 < n; p0 += batch) {
        const uint32_t nb = std::min(batch, n - p0);
        a.p0 = p0;
        a.nsig = 2 * nb;
        hipLaunchKernelGGL(k_cz_zero, dim3((a.nsig + 255) / 256), dim3(256), 0, s, a.flags, a.nsig);
        launch_check("k_cz_zero");
        if (PMAX == 2048) cz_launch<2048>(a, lds_cols, lds_rows, s);
        else cz_launch<8192>(a, lds_cols, lds_rows, s);
        LaunchTimer timer("k_cz_tune", s);
        hipLaunchKernelGGL(k_cz_tune, dim3(a.nsig), dim3(kCzThreads), 0, s, a);
        launch_check("k_cz_tune");
    }
}

}  // namespace hm
