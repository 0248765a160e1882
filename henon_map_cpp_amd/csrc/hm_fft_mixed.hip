// FFT tunes of windows of any length (src/dynamic_indicator.cpp:32-61, 178-196,
// 255-331 for any N). Windows of up to about 10^4 points run in one kernel per
// signal with the whole signal in LDS (k_wh_tune, below the four-step
// kernels). Longer windows whose length has only the prime factors 2, 3, 5 and
// 7 run a mixed-radix four-step transform of the length itself, N = N1 x N2 (index n = N2 n1 + n2
// in, k = k1 + N1 k2 out):
//   X[k1 + N1 k2] = sum_{n2} W_{N2}^{n2 k2} W_N^{n2 k1} sum_{n1} x[N2 n1 + n2] W_{N1}^{n1 k1}
// with both factors LDS-resident, so no padding to a power of two and no chirp:
//   k_mr_columns : load, mean removal and Hanning of a few columns n2, N1-point
//                  FFTs, step twiddle W_N^{n2 k1}, store rows k1
//   k_mr_rows    : per row k1 the N2-point FFT, magnitudes (the spectrum's
//                  entry k1 + N1 k2 sits at k1 N2 + k2)
//   k_mr_tune    : per signal the first maximum in the spectrum's order and the
//                  three-point interpolation (dynamic_indicator.cpp:139-176)
// Sub-transforms are Stockham passes of radix 8, 4, 2, 3, 5, 7 (runtime plan)
// over several columns or rows per workgroup. HBM traffic per signal: the
// samples once, 16 N bytes of scratch each way, 8 N bytes of magnitudes each way.
//
// Other lengths run Bluestein's chirp-z transform over the smallest
// {2,3,5,7}-smooth length M >= 2N - 1 that splits the same way (with
// c_m = exp(-i pi m^2 / N): X_k = c_k sum_m (x_m c_m) conj(c_{k-m}), a cyclic
// convolution of length M with the fixed filter B = DFT_M(conj(c)), computed
// once per N on the host in long double). The same kernels with the chirp:
//   k_mr_columns<1> : chirped samples (zero from N on), M1-point FFTs, twiddle
//   k_mr_rows<1>    : per row k1 the M2-point FFT (the spectrum, entry
//                     k1 + M1 k2), times B, conjugate, M2-point FFT, twiddle
//                     W_M^{k1 j1}: the inverse transform as the conjugate of a
//                     forward one (input index k1 + M1 k2, output j1 + M2 j2)
//   k_mr_final      : per column j1 the M1-point FFT over k1 (giving j2), the
//                     magnitudes of the entries j = j1 + M2 j2 < N
//   k_mr_tune<1>    : first maximum and interpolation in natural order
// |X_k| = |conv_k| / M; the scale is left out since the first maximum and the
// interpolation (homogeneous of degree 0 in the magnitudes) do not see it.
// Windows too long for two 2048-point factors take the power-of-two chirp-z
// transform (hm_chirpz.hip).
#include <hip/hip_runtime.h>
#include <algorithm>
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

constexpr uint32_t kMrThreads = 512;
constexpr uint32_t kMrPmax = 2048;  // points of one workgroup's transforms (32 KiB)
constexpr int kMrMaxPasses = 12;

// Radix sequence of one sub-length, four bits per pass (pass s in bits 4s..4s+3):
// read by shifts, so no per-pass table is indexed at run time.
struct MrPlan {
    uint32_t L, n_pass;
    uint64_t radices;
};

struct MrArgs {
    const double *hist;
    HistLayout lay;
    const double *means;  // [4][n] or NULL
    uint32_t n, p0, nsig, from, N;
    uint32_t M;              // transform length: N, or the chirp-z length
    uint32_t N1, N2, C, RR;  // M = N1 x N2; C columns per workgroup (columns kernels), RR rows (rows kernel)
    MrPlan p1, p2;
    const double *hann;      // [N]
    const Cplx *tw1, *tw2;   // W_{N1}^q, W_{N2}^q
    const Cplx *twm;         // [N1][N2]: W_M^{k1 n2}
    const Cplx *chirp;       // chirp-z: [N] exp(-i pi (m^2 mod 2N) / N)
    const Cplx *bhat;        // chirp-z: [N1][N2], B[k1 + N1 k2] at k1 N2 + k2
    Cplx *Y;                 // scratch [nsig][N1][N2]
    double *mags;            // scratch [nsig][N]
    uint32_t *flags;         // [nsig]
    double *out;             // [window][plane][n]
    uint32_t window;
};

__device__ __forceinline__ void dft3(Cplx *v)
{
    // X1 = a + c (b + d) - i s (b - d), X2 = a + c (b + d) + i s (b - d)
    constexpr double c = -0.5, s = 0.86602540378443864676;  // cos, sin of 2 pi / 3
    const Cplx a = v[0], t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
    v[0] = cadd(a, t1);
    const Cplx m1 = Cplx{fma(c, t1.re, a.re), fma(c, t1.im, a.im)};
    const Cplx m2 = Cplx{s * t2.im, -(s * t2.re)};
    v[1] = cadd(m1, m2);
    v[2] = csub(m1, m2);
}

// cos and sin of 2 pi q / R for q = 1 .. (R-1)/2
template <int R>
__device__ __forceinline__ constexpr double root_cos(int q)
{
    if constexpr (R == 5) {
        constexpr double c[3] = {1.0, 0.30901699437494742410, -0.80901699437494742410};
        return c[q];
    } else {
        constexpr double c[4] = {1.0, 0.62348980185873353053, -0.22252093395631440429, -0.90096886790241912624};
        return c[q];
    }
}
template <int R>
__device__ __forceinline__ constexpr double root_sin(int q)
{
    if constexpr (R == 5) {
        constexpr double s[3] = {0.0, 0.95105651629515357212, 0.58778525229247312917};
        return s[q];
    } else {
        constexpr double s[4] = {0.0, 0.78183148246802980871, 0.97492791218182360702, 0.43388373911755812048};
        return s[q];
    }
}

// Forward DFT of prime size R = 5, 7 from the pairs m, R - m: with
// a_m = v_m + v_{R-m}, b_m = v_m - v_{R-m} and t = 2 pi m k / R,
//   X_k = v_0 + sum_m (a_m cos t - i b_m sin t),  X_{R-k} = v_0 + sum_m (a_m cos t + i b_m sin t)
// (m, k = 1 .. (R-1)/2): (R-1) constants instead of (R-1)^2 products of roots.
template <int R>
__device__ __forceinline__ void dft_prime(Cplx *v)
{
    constexpr int H = (R - 1) / 2;
    Cplx a[H], b[H];
#pragma unroll
    for (int m = 1; m <= H; ++m) {
        a[m - 1] = cadd(v[m], v[R - m]);
        b[m - 1] = csub(v[m], v[R - m]);
    }
    Cplx x0 = v[0];
#pragma unroll
    for (int m = 0; m < H; ++m) x0 = cadd(x0, a[m]);
#pragma unroll
    for (int k = 1; k <= H; ++k) {
        Cplx p = v[0], q{0.0, 0.0};
#pragma unroll
        for (int m = 1; m <= H; ++m) {
            const int r = (m * k) % R;
            const int f = r <= H ? r : R - r;  // cos t = cos(2 pi f / R), sin t = +-sin(2 pi f / R)
            const double c = root_cos<R>(f), sn = r <= H ? root_sin<R>(f) : -root_sin<R>(f);
            p.re = fma(a[m - 1].re, c, p.re);
            p.im = fma(a[m - 1].im, c, p.im);
            q.re = fma(b[m - 1].re, sn, q.re);
            q.im = fma(b[m - 1].im, sn, q.im);
        }
        // X_k = p - i q, X_{R-k} = p + i q
        v[k] = Cplx{p.re + q.im, p.im - q.re};
        v[R - k] = Cplx{p.re - q.im, p.im + q.re};
    }
    v[0] = x0;
}

template <int R>
__device__ __forceinline__ void dft_any(Cplx *v)
{
    if constexpr (R == 2 || R == 4 || R == 8) dft_r<R>(v);
    else if constexpr (R == 3) dft3(v);
    else dft_prime<R>(v);
}

// LDS position of entry i of a signal: one pad entry after every 16, so that
// the power-of-two strides of the early passes' stores (R entries of 16 bytes
// apart when Ns = 1) fall into distinct banks.
__device__ __forceinline__ uint32_t lp(uint32_t i) { return i + (i >> 4); }
// LDS entries a signal of L entries spans (lp(L - 1) + 1); the odd variant for
// the columns kernels, whose loads step across signals.
__host__ __device__ constexpr uint32_t lds_len(uint32_t L) { return L + (L - 1) / 16 + 1; }
__host__ __device__ constexpr uint32_t lds_len_odd(uint32_t L) { return lds_len(L) | 1u; }

// One radix-R Stockham pass over G signals of length L in LDS (signal g at
// buf[g GS ..], entry i at lp(i)); Ns = product of the earlier radices. Butterfly j reads
// x[j + r L/R], twiddles by W_L^{k r L/(Ns R)} (k = j mod Ns; one table load, the
// higher powers by products), takes the R-point DFT and writes element r to
// (j - k) R + k + r Ns.
template <int R>
__device__ __forceinline__ void mr_pass(Cplx *buf, uint32_t G, uint32_t L, uint32_t GS, uint32_t Ns,
                                        const Cplx *__restrict__ tw)
{
    constexpr uint32_t PER = (kMrPmax / R + kMrThreads - 1) / kMrThreads;
    const uint32_t B = L / R, total = G * B, step = L / (Ns * R);
    Cplx v[PER][R];
    uint32_t goff[PER], base[PER];
#pragma unroll
    for (uint32_t q = 0; q < PER; ++q) {
        const uint32_t J = threadIdx.x + q * kMrThreads;
        base[q] = 0xffffffffu;
        if (J < total) {
            const uint32_t g = J / B, j = J - g * B;
            const Cplx *src = buf + g * GS;
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = src[lp(j + r * B)];
            const uint32_t k = j % Ns;
            if (k != 0) {
                const Cplx w1 = tw[k * step];
                Cplx w = w1;
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    v[q][r] = cmul(v[q][r], w);
                    if (r + 1 < R) w = cmul(w, w1);
                }
            }
            dft_any<R>(v[q]);
            goff[q] = g * GS;
            base[q] = (j - k) * R + k;
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < PER; ++q) {
        if (base[q] != 0xffffffffu) {
            Cplx *dst = buf + goff[q];
            const uint32_t ns = Ns;
#pragma unroll
            for (int r = 0; r < R; ++r) dst[lp(base[q] + r * ns)] = v[q][r];
        }
    }
    __syncthreads();
}

// Forward DFT of G signals of length p.L in LDS, in natural order.
__device__ __forceinline__ void mr_fft(Cplx *buf, uint32_t G, const MrPlan &p, uint32_t GS, const Cplx *__restrict__ tw)
{
    uint32_t Ns = 1;
    for (uint32_t s = 0; s < p.n_pass; ++s) {
        const uint32_t R = (uint32_t)(p.radices >> (4 * s)) & 15u;
        switch (R) {
        case 8: mr_pass<8>(buf, G, p.L, GS, Ns, tw); break;
        case 4: mr_pass<4>(buf, G, p.L, GS, Ns, tw); break;
        case 2: mr_pass<2>(buf, G, p.L, GS, Ns, tw); break;
        case 3: mr_pass<3>(buf, G, p.L, GS, Ns, tw); break;
        case 5: mr_pass<5>(buf, G, p.L, GS, Ns, tw); break;
        default: mr_pass<7>(buf, G, p.L, GS, Ns, tw); break;
        }
        Ns *= R;
    }
}

__device__ __forceinline__ uint32_t mr_virtual(uint32_t b, uint32_t per) { return (b % 8) * per + b / 8; }

template <bool CZ>
__global__ void __launch_bounds__(kMrThreads, 4) k_mr_columns(MrArgs a)
{
    extern __shared__ Cplx buf[];
    const uint32_t per_sig = a.N2 / a.C;
    const uint32_t total = a.nsig * per_sig;
    const uint32_t v = mr_virtual(blockIdx.x, gridDim.x / 8);
    if (v >= total) return;
    const uint32_t sig = v / per_sig, c0 = (v % per_sig) * a.C;
    const uint32_t C = a.C, N1 = a.N1, N2 = a.N2, GS = lds_len_odd(N1);
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
    for (uint32_t e = threadIdx.x; e < C * N1; e += kMrThreads) {
        const uint32_t n1 = e / C, c = e - n1 * C;
        const uint32_t m = n1 * N2 + c0 + c;
        Cplx z{0.0, 0.0};
        if (!CZ || m < a.N) {
            const uint64_t sm = (uint64_t)a.from + m;
            const double vx = a.hist[hat(a.lay, sm, cx, particle)] - mc[cx];
            const double vp = a.hist[hat(a.lay, sm, cp, particle)] - mc[cp];
            nan_seen |= __builtin_isnan(vx) || __builtin_isnan(vp);
            if (scan_other)
                nan_seen |= __builtin_isnan(a.hist[hat(a.lay, sm, co0, particle)] - mc[co0]) ||
                            __builtin_isnan(a.hist[hat(a.lay, sm, co1, particle)] - mc[co1]);
            const double h = a.hann[m];
            z = Cplx{vx * h, vp * h};
            if constexpr (CZ) z = cmul(z, a.chirp[m]);
        }
        buf[c * GS + lp(n1)] = z;
    }
    if (__any(nan_seen) && (threadIdx.x & 63) == 0) atomicOr(a.flags + sig, 1u);
    __syncthreads();
    mr_fft(buf, C, a.p1, GS, a.tw1);
    Cplx *Y = a.Y + (size_t)sig * a.M;
    for (uint32_t e = threadIdx.x; e < C * N1; e += kMrThreads) {
        const uint32_t k1 = e / C, c = e - k1 * C;
        const size_t o = (size_t)k1 * N2 + c0 + c;
        Y[o] = cmul(buf[c * GS + lp(k1)], a.twm[o]);
    }
}

template <bool CZ>
__global__ void __launch_bounds__(kMrThreads, 4) k_mr_rows(MrArgs a)
{
    extern __shared__ Cplx buf[];
    const uint32_t per_sig = a.N1 / a.RR;
    const uint32_t total = a.nsig * per_sig;
    const uint32_t v = mr_virtual(blockIdx.x, gridDim.x / 8);
    if (v >= total) return;
    const uint32_t sig = v / per_sig, r0 = (v % per_sig) * a.RR;
    const uint32_t N2 = a.N2, P = a.RR * N2, GS = lds_len(N2);
    // row g entry i (global offset e = g N2 + i) at LDS g GS + lp(i)
    auto at = [&](uint32_t e) {
        const uint32_t g = e / N2;
        return g * GS + lp(e - g * N2);
    };
    Cplx *Y = a.Y + (size_t)sig * a.M + (size_t)r0 * N2;
    for (uint32_t e = threadIdx.x; e < P; e += kMrThreads) buf[at(e)] = Y[e];
    __syncthreads();
    mr_fft(buf, a.RR, a.p2, GS, a.tw2);
    if constexpr (CZ) {
        // spectrum entry k1 + N1 k2 times B, conjugated: the inverse transform
        // becomes a forward one over k2 (then k1, in k_mr_final)
        const Cplx *bh = a.bhat + (size_t)r0 * N2;
        for (uint32_t e = threadIdx.x; e < P; e += kMrThreads) {
            const uint32_t l = at(e);
            const Cplx z = cmul(buf[l], bh[e]);
            buf[l] = Cplx{z.re, -z.im};
        }
        __syncthreads();
        mr_fft(buf, a.RR, a.p2, GS, a.tw2);
        const Cplx *tm = a.twm + (size_t)r0 * N2;
        for (uint32_t e = threadIdx.x; e < P; e += kMrThreads) Y[e] = cmul(buf[at(e)], tm[e]);
    } else {
        // spectrum entry k1 + N1 k2 at row k1, position k2
        double *mags = a.mags + (size_t)sig * a.N + (size_t)r0 * N2;
        for (uint32_t e = threadIdx.x; e < P; e += kMrThreads) {
            const Cplx z = buf[at(e)];
            mags[e] = sqrt(z.re * z.re + z.im * z.im);
        }
    }
}

// Chirp-z only: columns j1 in [c0, c0 + C), the M1-point transforms over k1
// (output j2) and the magnitudes of the entries j = j1 + N2 j2 < N.
__global__ void __launch_bounds__(kMrThreads, 4) k_mr_final(MrArgs a)
{
    extern __shared__ Cplx buf[];
    const uint32_t per_sig = a.N2 / a.C;
    const uint32_t total = a.nsig * per_sig;
    const uint32_t v = mr_virtual(blockIdx.x, gridDim.x / 8);
    if (v >= total) return;
    const uint32_t sig = v / per_sig, c0 = (v % per_sig) * a.C;
    const uint32_t C = a.C, N1 = a.N1, N2 = a.N2, GS = lds_len_odd(N1);
    const Cplx *Y = a.Y + (size_t)sig * a.M;
    for (uint32_t e = threadIdx.x; e < C * N1; e += kMrThreads) {
        const uint32_t k1 = e / C, c = e - k1 * C;
        buf[c * GS + lp(k1)] = Y[(size_t)k1 * N2 + c0 + c];
    }
    __syncthreads();
    mr_fft(buf, C, a.p1, GS, a.tw1);
    double *mags = a.mags + (size_t)sig * a.N;
    for (uint32_t e = threadIdx.x; e < C * N1; e += kMrThreads) {
        const uint32_t j2 = e / C, c = e - j2 * C;
        const uint32_t j = c0 + c + j2 * N2;
        if (j < a.N) {
            const Cplx z = buf[c * GS + lp(j2)];
            mags[j] = sqrt(z.re * z.re + z.im * z.im);
        }
    }
}

template <bool NATURAL>
__global__ void __launch_bounds__(kMrThreads) k_mr_tune(MrArgs a)
{
    __shared__ double sv[kMrThreads / 64];
    __shared__ uint32_t si[kMrThreads / 64];
    const uint32_t sig = blockIdx.x;
    if (sig >= a.nsig) return;
    const uint32_t particle = a.p0 + sig / 2, plane = sig % 2;
    double *o = a.out + ((size_t)a.window * 2 + plane) * a.n + particle;
    if (a.flags[sig]) {
        if (threadIdx.x == 0) *o = __builtin_nan("");
        return;
    }
    const double *mags = a.mags + (size_t)sig * a.N;
    const uint32_t N1 = a.N1, N2 = a.N2;
    // first maximum in the spectrum's order: the lowest k among equal
    // magnitudes, whatever the storage order (k1 N2 + k2 for k = k1 + N1 k2)
    double v = -1.0;
    uint32_t i = 0xffffffffu;
    for (uint32_t p = threadIdx.x; p < a.N; p += kMrThreads) {
        const double mg = mags[p];
        uint32_t k = p;
        if constexpr (!NATURAL) {
            const uint32_t k1 = p / N2;
            k = k1 + N1 * (p - k1 * N2);
        }
        if (mg > v || (mg == v && k < i) || i == 0xffffffffu) { v = mg; i = k; }
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
        for (uint32_t q = 1; q < kMrThreads / 64; ++q)
            if (sv[q] > bv || (sv[q] == bv && si[q] < bi)) { bv = sv[q]; bi = si[q]; }
        if constexpr (NATURAL) *o = interp_tune_at([&](uint32_t k) { return mags[k]; }, a.N, bi);
        else *o = interp_tune_at([&](uint32_t k) { return mags[(size_t)(k % N1) * N2 + k / N1]; }, a.N, bi);
    }
}

// ---------------------------------------------------------------- whole-signal transforms
// A window whose transform length fits one workgroup's LDS (about 10^4 points
// in 160 KiB) runs in one kernel, one workgroup per signal: the samples are
// read once, every pass runs in LDS and the tune is found there, with no
// scratch in HBM and no second launch.
//   direct : lengths whose prime factors are 2, 3, 5, 7 (the radix passes above)
//            or other primes up to kWhPrimeMax (one prime-size pass each, below)
//   chirp-z: other lengths N, over a smooth M >= 2N - 1 that fits: the chirped
//            samples, FFT_M, times B = DFT_M(conj(c)), conjugated, FFT_M again;
//            |X_k| = |V_k| / M for k < N (the scale does not change the tune)
// The signal sits in LDS unpadded: the first pass takes an odd radix where the
// length has an odd factor, so its stores step by an odd number of 16-byte
// entries, and the later passes store runs of consecutive entries (a power of
// two, here only as a chirp-z length, pays some bank conflicts in its first
// pass).
constexpr uint32_t kWhPrimeMax = 257;
constexpr int kWhMaxPasses = 16;

// One pass: radix R after the radices of product Ns, B = L / R butterflies,
// twiddle stride step = L / (Ns R); m_* are the multipliers of udiv_m for Ns,
// H = (R - 1) / 2 and H + 1.
struct WhPass {
    uint32_t R, Ns, B, step, m_ns, m_h, m_h1;
};
struct WhPlan {
    uint32_t L, n_pass;
    WhPass pass[kWhMaxPasses];
};

// j / d for j < 2^16 and d <= 2^16 from m = floor(2^32 / d) + 1 (m = 0: d = 1):
// j m / 2^32 = j / d + j (m - 2^32 / d) / 2^32 exceeds j / d by less than
// 2^-16, and the fraction of j / d is at most 1 - 1 / d.
__device__ __forceinline__ uint32_t udiv_m(uint32_t j, uint32_t m) { return m ? __umulhi(j, m) : j; }

struct WhArgs {
    const double *hist;
    HistLayout lay;
    const double *means;  // [4][n] or NULL
    uint32_t n, p0, from, N, M;
    WhPlan plan;          // passes of the M-point transform
    const double *hann;   // [N]
    const Cplx *tw;       // W_M^q, q < M
    const Cplx *chirp;    // chirp-z: [N]
    const Cplx *bhat;     // chirp-z: [M], B in natural order
    double *out;          // [window][plane][n]
    uint32_t window;
};


// One radix-R Stockham pass (as mr_pass) over the whole signal of length L:
// butterfly j reads x[j + r L/R], twiddles by W_L^{k r L/(Ns R)} (k = j mod
// Ns), takes the R-point DFT and writes element r to (j - k) R + k + r Ns.
template <int R, uint32_t NT, uint32_t PMAX>
__device__ __forceinline__ void wh_pass(Cplx *buf, const WhPass &ps, const Cplx *__restrict__ tw)
{
    constexpr uint32_t PER = (PMAX / R + NT - 1) / NT;
    const uint32_t B = ps.B, step = ps.step, Ns = ps.Ns;
    Cplx v[PER][R];
    uint32_t base[PER];
#pragma unroll
    for (uint32_t q = 0; q < PER; ++q) {
        const uint32_t j = threadIdx.x + q * NT;
        base[q] = 0xffffffffu;
        if (j < B) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = buf[j + r * B];
            const uint32_t k = j - udiv_m(j, ps.m_ns) * Ns;
            if (k != 0) {
                const Cplx w1 = tw[k * step];
                Cplx w = w1;
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    v[q][r] = cmul(v[q][r], w);
                    if (r + 1 < R) w = cmul(w, w1);
                }
            }
            dft_any<R>(v[q]);
            base[q] = (j - k) * R + k;
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < PER; ++q)
        if (base[q] != 0xffffffffu)
#pragma unroll
            for (int r = 0; r < R; ++r) buf[base[q] + r * Ns] = v[q][r];
    __syncthreads();
}

// A pass of odd prime radix R (11 .. kWhPrimeMax), the same butterflies with
// the R-point DFT from the pairs (r, R - r) as in dft_prime:
//   A: each pair twiddled and replaced in place by a_r = x_r + x_{R-r} (at r)
//      and b_r = x_r - x_{R-r} (at R - r)
//   B: one thread per output pair (q, R - q) of a butterfly (q = 0: the sum),
//      Y_q = x_0 + sum_r (a_r cos t - i b_r sin t), Y_{R-q} with + i, t = 2 pi r q / R;
//      the lanes of a wavefront share the butterfly (the same a_r, b_r: LDS
//      broadcasts) and step cos t, sin t by the three-term recurrence
//      f(r+1) = 2 cos(2 pi q / R) f(r) - f(r-1) (two multiply-adds a step)
// R(R-1)/2 real multiply-adds per butterfly and plane pair instead of R^2
// complex products. The recurrence's rounding errors grow at most like
// r min(r, 1 / |sin(2 pi q / R)|) units in the last place: about 1e-12 of the
// unit roots at most over the (R-1)/2 <= 128 steps, far under what moves a
// tune (the tests hold the tunes to 1e-12 of the long-double DFT's).
template <uint32_t NT, uint32_t PMAX>
__device__ __forceinline__ void wh_pass_prime(Cplx *buf, const WhPass &ps, const Cplx *__restrict__ tw)
{
    // output pairs per thread: B (H + 1) <= L (R + 1) / (2R) <= 0.55 L for R >= 11
    constexpr uint32_t GP = ((PMAX * 11 / 20 + NT - 1) / NT + 1) / 2 * 2;
    const uint32_t R = ps.R, Ns = ps.Ns, B = ps.B, H = (R - 1) / 2, step = ps.step;
    for (uint32_t e = threadIdx.x; e < B * H; e += NT) {
        const uint32_t j = udiv_m(e, ps.m_h), r = e - j * H + 1, k = j - udiv_m(j, ps.m_ns) * Ns;
        const uint32_t iu = j + r * B, iv = j + (R - r) * B;
        Cplx u = buf[iu], v = buf[iv];
        if (k != 0) {
            // k r step < L: the exponents need no reduction
            u = cmul(u, tw[k * r * step]);
            v = cmul(v, tw[k * (R - r) * step]);
        }
        buf[iu] = cadd(u, v);
        buf[iv] = csub(u, v);
    }
    __syncthreads();
    // output positions (< 2^15) of Y_q in the low, of Y_{R-q} in the high half
    static_assert(PMAX < 0xffffu, "output positions are packed in 16 bits");
    static_assert(GP % 2 == 0, "output pairs are taken two at a time");
    Cplx y0[GP], y1[GP];
    uint32_t oo[GP];
    const uint32_t items = B * (H + 1);
#pragma unroll
    for (uint32_t g = 0; g < GP; g += 2) {
        // two output pairs at once (independent chains); q = 0 runs the same
        // loop with cos = 1, sin = 0, which leaves Y_0 = x_0 + sum a_r
        uint32_t ja[2], qa[2];
        Cplx pa[2], sb[2];
        double c[2], sn[2], cp[2], sp[2], t2[2];
#pragma unroll
        for (uint32_t h = 0; h < 2; ++h) {
            const uint32_t e = min(threadIdx.x + (g + h) * NT, items - 1);
            ja[h] = udiv_m(e, ps.m_h1);
            qa[h] = e - ja[h] * (H + 1);
            const Cplx wq = tw[qa[h] * B];  // W_L^{q L / R} = exp(-2 pi i q / R)
            c[h] = wq.re;
            sn[h] = -wq.im;
            cp[h] = 1.0;
            sp[h] = 0.0;
            t2[h] = 2.0 * wq.re;
            pa[h] = buf[ja[h]];
            sb[h] = Cplx{0.0, 0.0};
        }
        for (uint32_t r = 1; r <= H; ++r) {
            Cplx av[2], bv[2];
#pragma unroll
            for (uint32_t h = 0; h < 2; ++h) {
                av[h] = buf[ja[h] + r * B];
                bv[h] = buf[ja[h] + (R - r) * B];
            }
#pragma unroll
            for (uint32_t h = 0; h < 2; ++h) {
                pa[h].re = fma(av[h].re, c[h], pa[h].re);
                pa[h].im = fma(av[h].im, c[h], pa[h].im);
                sb[h].re = fma(bv[h].re, sn[h], sb[h].re);
                sb[h].im = fma(bv[h].im, sn[h], sb[h].im);
                // cos, sin of (r + 1) 2 pi q / R: f(r+1) = 2 cos(2 pi q / R) f(r) - f(r-1)
                const double cn = fma(t2[h], c[h], -cp[h]), sn2 = fma(t2[h], sn[h], -sp[h]);
                cp[h] = c[h];
                sp[h] = sn[h];
                c[h] = cn;
                sn[h] = sn2;
            }
        }
#pragma unroll
        for (uint32_t h = 0; h < 2; ++h) {
            const uint32_t e = threadIdx.x + (g + h) * NT;
            const uint32_t k = ja[h] - udiv_m(ja[h], ps.m_ns) * Ns, ob = (ja[h] - k) * R + k, q = qa[h];
            y0[g + h] = Cplx{pa[h].re + sb[h].im, pa[h].im - sb[h].re};  // Y_q
            y1[g + h] = Cplx{pa[h].re - sb[h].im, pa[h].im + sb[h].re};  // Y_{R-q}
            oo[g + h] = e >= items ? 0xffffffffu
                        : q == 0   ? (0xffff0000u | ob)
                                   : ((ob + q * Ns) | ((ob + (R - q) * Ns) << 16));
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t g = 0; g < GP; ++g) {
        const uint32_t lo = oo[g] & 0xffffu, hi = oo[g] >> 16;
        if (lo != 0xffffu) buf[lo] = y0[g];
        if (hi != 0xffffu) buf[hi] = y1[g];
    }
    __syncthreads();
}

// Forward DFT of the whole signal in LDS, in natural order. The plan lists its
// radices grouped (3s, 5s, 7s, the other primes, then 2s, 4s, 8s: an odd or
// small first radix keeps the first pass's strided stores off shared banks),
// so each radix gets a loop of its own: one switch over the inlined passes
// inside a single loop keeps more registers live than any pass needs.
template <uint32_t NT, uint32_t PMAX>
__device__ __forceinline__ void wh_fft(Cplx *buf, const WhPlan &p, const Cplx *__restrict__ tw)
{
    uint32_t s = 0;
    for (; s < p.n_pass && p.pass[s].R == 3; ++s) wh_pass<3, NT, PMAX>(buf, p.pass[s], tw);
    for (; s < p.n_pass && p.pass[s].R == 5; ++s) wh_pass<5, NT, PMAX>(buf, p.pass[s], tw);
    for (; s < p.n_pass && p.pass[s].R == 7; ++s) wh_pass<7, NT, PMAX>(buf, p.pass[s], tw);
    for (; s < p.n_pass && p.pass[s].R > 8; ++s) wh_pass_prime<NT, PMAX>(buf, p.pass[s], tw);
    for (; s < p.n_pass && p.pass[s].R == 2; ++s) wh_pass<2, NT, PMAX>(buf, p.pass[s], tw);
    for (; s < p.n_pass && p.pass[s].R == 4; ++s) wh_pass<4, NT, PMAX>(buf, p.pass[s], tw);
    for (; s < p.n_pass && p.pass[s].R == 8; ++s) wh_pass<8, NT, PMAX>(buf, p.pass[s], tw);
}

// One workgroup per signal (2 per particle: plane x, plane y): load, mean
// removal and Hanning, the transform(s), the first maximum of |X_k| over
// k < N and the three-point interpolation (dynamic_indicator.cpp:139-196).
template <bool CZ, uint32_t NT, uint32_t PMAX>
__global__ void __launch_bounds__(NT, 4) k_wh_tune(WhArgs a)
{
    extern __shared__ Cplx buf[];
    __shared__ double sv[NT / 64];
    __shared__ uint32_t si[NT / 64];
    const uint32_t sig = blockIdx.x;
    const uint32_t particle = a.p0 + sig / 2, plane = sig % 2;
    const size_t nn = a.n;
    double *o = a.out + ((size_t)a.window * 2 + plane) * nn + particle;
    const uint32_t N = a.N, M = a.M;
    double mc[4] = {0.0, 0.0, 0.0, 0.0};
    if (a.means)
        for (int c = 0; c < 4; ++c) mc[c] = a.means[c * nn + particle];
    if (any_nan(mc[0], mc[1], mc[2], mc[3])) {
        // a NaN mean makes every sample of its coordinate NaN: nothing to read
        if (threadIdx.x == 0) *o = __builtin_nan("");
        return;
    }
    // four finite means prove that no sample of the particle is NaN or
    // infinite: the other plane need not be read
    const bool scan_other =
        !a.means || !(isfinite(mc[0]) && isfinite(mc[1]) && isfinite(mc[2]) && isfinite(mc[3]));
    const int cx = 2 * plane, cp = cx + 1, co0 = 2 * (1 - plane), co1 = co0 + 1;
    int nan_seen = 0;
    // every sample load of the thread issued before the first use (a loop
    // that waits for each sample in turn leaves the workgroup idle)
    constexpr uint32_t LP = (CZ ? PMAX / 2 : PMAX + NT - 1) / NT + (CZ ? 1 : 0);
    double vx[LP], vp[LP], hh[LP];
#pragma unroll
    for (uint32_t u = 0; u < LP; ++u) {
        const uint32_t m = threadIdx.x + u * NT;
        vx[u] = vp[u] = hh[u] = 0.0;
        if (m < N) {
            const uint64_t sm = (uint64_t)a.from + m;
            vx[u] = a.hist[hat(a.lay, sm, cx, particle)];
            vp[u] = a.hist[hat(a.lay, sm, cp, particle)];
            hh[u] = a.hann[m];
        }
    }
    if (scan_other)
        for (uint32_t m = threadIdx.x; m < N; m += NT) {
            const uint64_t sm = (uint64_t)a.from + m;
            nan_seen |= __builtin_isnan(a.hist[hat(a.lay, sm, co0, particle)] - mc[co0]) ||
                        __builtin_isnan(a.hist[hat(a.lay, sm, co1, particle)] - mc[co1]);
        }
#pragma unroll
    for (uint32_t u = 0; u < LP; ++u) {
        const uint32_t m = threadIdx.x + u * NT;
        if (m < N) {
            const double x = vx[u] - mc[cx], xp = vp[u] - mc[cp];
            nan_seen |= __builtin_isnan(x) || __builtin_isnan(xp);
            Cplx z{x * hh[u], xp * hh[u]};
            if constexpr (CZ) z = cmul(z, a.chirp[m]);
            buf[m] = z;
        }
    }
    if constexpr (CZ)
        for (uint32_t m = N + threadIdx.x; m < M; m += NT) buf[m] = Cplx{0.0, 0.0};
    if (__syncthreads_or(nan_seen)) {
        if (threadIdx.x == 0) *o = __builtin_nan("");
        return;
    }
    wh_fft<NT, PMAX>(buf, a.plan, a.tw);
    if constexpr (CZ) {
        for (uint32_t k = threadIdx.x; k < M; k += NT) {
            const uint32_t l = k;
            const Cplx z = cmul(buf[l], a.bhat[k]);
            buf[l] = Cplx{z.re, -z.im};
        }
        __syncthreads();
        wh_fft<NT, PMAX>(buf, a.plan, a.tw);
    }
    auto mag = [&](uint32_t k) {
        const Cplx z = buf[k];
        return sqrt(z.re * z.re + z.im * z.im);
    };
    // first maximum: each thread scans increasing k, ties go to the lower k
    double v = -1.0;
    uint32_t i = 0xffffffffu;
    for (uint32_t k = threadIdx.x; k < N; k += NT) {
        const double mg = mag(k);
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
        for (uint32_t q = 1; q < NT / 64; ++q)
            if (sv[q] > bv || (sv[q] == bv && si[q] < bi)) { bv = sv[q]; bi = si[q]; }
        *o = interp_tune_at(mag, N, bi);
    }
}

__global__ void k_mr_zero(uint32_t *__restrict__ p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------- host plan
struct MrTables {
    uint32_t M = 0, N1 = 0, N2 = 0;  // M = 0: no plan for this length
    MrPlan p1{}, p2{};
    Cplx *tw1 = nullptr, *tw2 = nullptr, *twm = nullptr;
    Cplx *chirp = nullptr, *bhat = nullptr;  // chirp-z plans only
};

// Radix plan of a {2,3,5,7}-smooth length: radix 8 while it divides, then 4,
// 2, 3, 5, 7. False when L has another prime factor.
static bool radix_plan(uint32_t L, MrPlan &p)
{
    p = MrPlan{};
    p.L = L;
    uint32_t r = L;
    auto take = [&](uint32_t R) {
        while (r % R == 0 && p.n_pass < (uint32_t)kMrMaxPasses) {
            p.radices |= (uint64_t)R << (4 * p.n_pass++);
            r /= R;
        }
    };
    take(8); take(4); take(2); take(3); take(5); take(7);
    return r == 1;
}

// N = N1 x N2 with N1 <= N2 <= kMrPmax, N1 >= 2 and N1 as large as possible.
static bool split(uint32_t N, uint32_t &N1, uint32_t &N2)
{
    N1 = N2 = 0;
    for (uint32_t a = 2; (uint64_t)a * a <= N; ++a)
        if (N % a == 0 && N / a <= kMrPmax) { N1 = a; N2 = N / a; }
    return N1 != 0;
}

// Both factors of a split length have radix plans.
static bool plan_length(uint32_t M, MrTables &t)
{
    uint32_t r = M;
    for (uint32_t q : {2u, 3u, 5u, 7u})
        while (r % q == 0) r /= q;
    if (r != 1) return false;
    uint32_t N1, N2;
    if (!split(M, N1, N2) || !radix_plan(N1, t.p1) || !radix_plan(N2, t.p2)) return false;
    t.M = M; t.N1 = N1; t.N2 = N2;
    return true;
}

template <class T>
static T *upload(const std::vector<T> &v, hipStream_t s, const char *what)
{
    T *d = nullptr;
    hip_check(hipMalloc(&d, v.size() * sizeof(T)), what);
    hip_check(hipMemcpyAsync(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    return d;
}

using CplxL = std::complex<long double>;

// Forward DFT of a {2,3,5,7}-smooth length in long double (the chirp-z filter
// spectrum): recursive decimation in time by the smallest prime factor r of
// n = r m, X[k + m q] = sum_s W_n^{s (k + m q)} Y_s[k], Y_s the DFT of x[s + r j].
// root[e] = W_M^e; a sub-length n uses W_n^e = root[e M / n].
static void host_dft_rec(const CplxL *in, size_t stride, CplxL *out, size_t n, const std::vector<CplxL> &root,
                         size_t M, std::vector<CplxL> &tmp)
{
    if (n == 1) { out[0] = in[0]; return; }
    size_t r = 2;
    while (n % r) ++r;
    const size_t m = n / r, rs = M / n;
    for (size_t s = 0; s < r; ++s) host_dft_rec(in + s * stride, stride * r, out + s * m, m, root, M, tmp);
    CplxL *t = tmp.data();
    for (size_t k = 0; k < m; ++k)
        for (size_t q = 0; q < r; ++q) {
            const size_t kk = k + m * q;
            CplxL acc = out[k];
            for (size_t s = 1; s < r; ++s) acc += out[s * m + k] * root[((s * kk) % n) * rs];
            t[kk] = acc;
        }
    std::copy(t, t + n, out);
}

static std::vector<CplxL> host_dft(const std::vector<CplxL> &x)
{
    const size_t M = x.size();
    const long double two_pi = 6.283185307179586476925286766559005768L;
    std::vector<CplxL> root(M), out(M), tmp(M);
    for (size_t q = 0; q < M; ++q) {
        const long double ang = two_pi * (long double)q / (long double)M;
        root[q] = CplxL(cosl(ang), -sinl(ang));
    }
    host_dft_rec(x.data(), 1, out.data(), M, root, M, tmp);
    return out;
}

// Tables of one length (cached per device): the direct plan of a
// {2,3,5,7}-smooth N, else the chirp-z plan over the smallest smooth M >=
// 2N - 1 that splits; M = 0 when neither exists (M would exceed kMrPmax^2).
static const MrTables *mr_tables(uint32_t N, hipStream_t s)
{
    static std::mutex mu;
    static std::map<std::pair<int, uint32_t>, MrTables> cache;
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find({dev, N});
    if (it != cache.end()) return it->second.M ? &it->second : nullptr;
    MrTables t;
    bool cz = false;
    if (!plan_length(N, t)) {
        cz = true;
        const uint64_t lo = 2ull * N - 1, hi = (uint64_t)kMrPmax * kMrPmax;
        for (uint64_t M = lo; M <= hi; ++M)
            if (plan_length((uint32_t)M, t)) break;
        if (!t.M) {
            cache[{dev, N}] = MrTables{};
            return nullptr;
        }
    }
    const uint32_t M = t.M, N1 = t.N1, N2 = t.N2;
    const long double two_pi = 6.283185307179586476925286766559005768L, pi = two_pi / 2;
    auto roots = [&](uint32_t L) {
        std::vector<Cplx> v(L);
        for (uint32_t q = 0; q < L; ++q) {
            const long double ang = two_pi * (long double)q / (long double)L;
            v[q] = Cplx{(double)cosl(ang), (double)-sinl(ang)};
        }
        return v;
    };
    const std::vector<Cplx> r1 = roots(N1), r2 = roots(N2), rM = roots(M);
    // step twiddles W_M^{k1 n2}, reduced exactly mod M
    std::vector<Cplx> m((size_t)M);
    for (uint64_t k1 = 0; k1 < N1; ++k1)
        for (uint64_t n2 = 0; n2 < N2; ++n2) m[k1 * N2 + n2] = rM[(k1 * n2) % M];
    t.tw1 = upload(r1, s, "hipMalloc(mixed-radix twiddles)");
    t.tw2 = upload(r2, s, "hipMalloc(mixed-radix twiddles)");
    t.twm = upload(m, s, "hipMalloc(mixed-radix step twiddles)");
    if (cz) {
        // chirp c_m = exp(-i pi (m^2 mod 2N) / N) (m^2 exact in 64 bits), filter
        // b_j = conj(c_j) for |j| < N (j mod M), B = DFT_M(b) stored [k1][k2]
        std::vector<Cplx> chirp(N), bhat((size_t)M);
        std::vector<CplxL> b((size_t)M, CplxL(0, 0));
        for (uint64_t q = 0; q < N; ++q) {
            const long double ang = pi * (long double)((q * q) % (2ull * N)) / (long double)N;
            const long double c = cosl(ang), sn = sinl(ang);
            chirp[q] = Cplx{(double)c, (double)-sn};
            b[q] = CplxL(c, sn);
            if (q > 0) b[M - q] = b[q];
        }
        const std::vector<CplxL> B = host_dft(b);
        for (size_t k1 = 0; k1 < N1; ++k1)
            for (size_t k2 = 0; k2 < N2; ++k2) {
                const CplxL z = B[k1 + (size_t)N1 * k2];
                bhat[k1 * N2 + k2] = Cplx{(double)z.real(), (double)z.imag()};
            }
        t.chirp = upload(chirp, s, "hipMalloc(chirp)");
        t.bhat = upload(bhat, s, "hipMalloc(chirp-z filter)");
    }
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    cache[{dev, N}] = t;
    return &cache[{dev, N}];
}

static uint64_t mr_budget()
{
    const char *e = std::getenv("HM_FFT_ROW_BYTES");
    if (e && *e) {
        const unsigned long long v = std::strtoull(e, nullptr, 10);
        if (v > 0) return v;
    }
    return 8ull << 30;  // a few percent of the device memory: large batches, few launches
}

// largest divisor of v that is at most cap
static uint32_t divisor_at_most(uint32_t v, uint32_t cap)
{
    for (uint32_t d = std::min(v, cap); d > 1; --d)
        if (v % d == 0) return d;
    return 1;
}

template <class K>
static void set_lds(K kernel, size_t bytes)
{
    hip_check(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes),
              "hipFuncSetAttribute");
}

// ---------------------------------------------------------------- whole-signal plans
// The three shapes of k_wh_tune: up to 2048 points with 256 threads, up to
// 4096 with 512 (several workgroups per CU at four wavefronts per SIMD), up to
// kWhPmax with 1024 (one).
constexpr uint32_t kWhSmallPmax = 2048, kWhSmallThreads = 256;
constexpr uint32_t kWhMidPmax = 4096, kWhMidThreads = 512;
constexpr uint32_t kWhPmax = 10240, kWhThreads = 1024;

// Passes of the direct transform of L, grouped as wh_fft runs them: 3s, 5s,
// 7s, the other primes in increasing order (each at most kWhPrimeMax), then the
// power of two as 2, 4 and 8s (8 while it divides). `gen` = the sum of the
// other primes (their passes cost about 2R operations per point against a few
// tens for a small radix).
static uint32_t udiv_magic(uint32_t d) { return d == 1 ? 0u : (uint32_t)((1ull << 32) / d + 1); }

static bool wh_radix_plan(uint32_t L, WhPlan &p, uint32_t &gen)
{
    p = WhPlan{};
    p.L = L;
    gen = 0;
    uint32_t r = L, Ns = 1;
    auto push = [&](uint32_t R) {
        if (p.n_pass == (uint32_t)kWhMaxPasses) return false;
        WhPass &q = p.pass[p.n_pass++];
        q.R = R; q.Ns = Ns; q.B = L / R; q.step = L / (Ns * R);
        q.m_ns = udiv_magic(Ns); q.m_h = udiv_magic(std::max(1u, (R - 1) / 2)); q.m_h1 = udiv_magic((R - 1) / 2 + 1);
        Ns *= R;
        r /= R;
        return true;
    };
    for (uint32_t R : {3u, 5u, 7u})
        while (r % R == 0)
            if (!push(R)) return false;
    for (uint32_t q = 11; r > 1 && q <= kWhPrimeMax; q += 2)
        while (r % q == 0) {
            if (!push(q)) return false;
            gen += q;
        }
    uint32_t twos = 0;
    while (r % 2 == 0) { r /= 2; ++twos; }
    if (r != 1) return false;
    r = 1u << twos;  // the power of two, taken as 2^(twos mod 3 == 1), 4^(... == 2), then 8s
    if (twos % 3 == 1 && !push(2)) return false;
    if (twos % 3 == 2 && !push(4)) return false;
    while (r > 1)
        if (!push(8)) return false;
    return true;
}

struct WhTables {
    bool ok = false, cz = false;
    uint32_t M = 0, nt = 0;
    WhPlan plan{};
    Cplx *tw = nullptr, *chirp = nullptr, *bhat = nullptr;
};


// Whole-signal plan of a window length (cached per device): the direct
// transform when N is {2,3,5,7}-smooth or its other primes are cheap (sum at
// most kWhCzGen), else chirp-z over the smallest smooth M >= 2N - 1 that
// fits, else the direct transform with dearer prime passes; none when N does
// not fit or has a prime above kWhPrimeMax and no chirp-z length fits.
constexpr uint32_t kWhCzGen = 128;

static const WhTables *wh_tables(uint32_t N, hipStream_t s)
{
    static std::mutex mu;
    static std::map<std::pair<int, uint32_t>, WhTables> cache;
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find({dev, N});
    if (it != cache.end()) return it->second.ok ? &it->second : nullptr;
    int v = 0;
    hip_check(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev), "hipDeviceGetAttribute");
    // the kernels' own LDS (reductions) comes on top of the signal
    size_t fixed = 0;
    for (const void *k : {(const void *)k_wh_tune<false, kWhSmallThreads, kWhSmallPmax>,
                          (const void *)k_wh_tune<true, kWhSmallThreads, kWhSmallPmax>,
                          (const void *)k_wh_tune<false, kWhMidThreads, kWhMidPmax>,
                          (const void *)k_wh_tune<true, kWhMidThreads, kWhMidPmax>,
                          (const void *)k_wh_tune<false, kWhThreads, kWhPmax>,
                          (const void *)k_wh_tune<true, kWhThreads, kWhPmax>}) {
        hipFuncAttributes fa{};
        hip_check(hipFuncGetAttributes(&fa, k), "hipFuncGetAttributes");
        fixed = std::max(fixed, (size_t)fa.sharedSizeBytes);
    }
    const size_t lds = (size_t)v > fixed ? (size_t)v - fixed : 0;
    auto fits = [&](uint32_t M) { return M <= kWhPmax && (size_t)M * sizeof(Cplx) <= lds; };
    WhTables t;
    WhPlan pd{};
    uint32_t gen = 0;
    const bool direct = fits(N) && wh_radix_plan(N, pd, gen);
    uint32_t Mcz = 0;
    WhPlan pc{};
    if (!(direct && gen <= kWhCzGen)) {
        for (uint64_t M = 2ull * N - 1; M <= kWhPmax && fits((uint32_t)M); ++M) {
            uint32_t g = 0;
            if (wh_radix_plan((uint32_t)M, pc, g) && g == 0) { Mcz = (uint32_t)M; break; }
        }
    }
    if (direct && (gen <= kWhCzGen || !Mcz)) {
        t.ok = true; t.cz = false; t.M = N; t.plan = pd;
    } else if (Mcz) {
        t.ok = true; t.cz = true; t.M = Mcz; t.plan = pc;
    }
    if (!t.ok) {
        cache[{dev, N}] = WhTables{};
        return nullptr;
    }
    const uint32_t M = t.M;
    t.nt = M <= kWhSmallPmax ? kWhSmallThreads : M <= kWhMidPmax ? kWhMidThreads : kWhThreads;
    const long double two_pi = 6.283185307179586476925286766559005768L, pi = two_pi / 2;
    std::vector<Cplx> w(M);
    for (uint32_t q = 0; q < M; ++q) {
        const long double ang = two_pi * (long double)q / (long double)M;
        w[q] = Cplx{(double)cosl(ang), (double)-sinl(ang)};
    }
    t.tw = upload(w, s, "hipMalloc(whole-signal twiddles)");
    if (t.cz) {
        // chirp c_m = exp(-i pi (m^2 mod 2N) / N), filter b_j = conj(c_j) for
        // |j| < N (j mod M), B = DFT_M(b) in natural order
        std::vector<Cplx> chirp(N), bhat(M);
        std::vector<CplxL> b(M, CplxL(0, 0));
        for (uint64_t q = 0; q < N; ++q) {
            const long double ang = pi * (long double)((q * q) % (2ull * N)) / (long double)N;
            const long double c = cosl(ang), sn = sinl(ang);
            chirp[q] = Cplx{(double)c, (double)-sn};
            b[q] = CplxL(c, sn);
            if (q > 0) b[M - q] = b[q];
        }
        const std::vector<CplxL> B = host_dft(b);
        for (uint32_t k = 0; k < M; ++k) bhat[k] = Cplx{(double)B[k].real(), (double)B[k].imag()};
        t.chirp = upload(chirp, s, "hipMalloc(chirp)");
        t.bhat = upload(bhat, s, "hipMalloc(chirp-z filter)");
    }
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    cache[{dev, N}] = t;
    return &cache[{dev, N}];
}

template <bool CZ, uint32_t NT, uint32_t PMAX>
static void launch_wh(const WhArgs &a0, uint32_t n, size_t lds_bytes, hipStream_t s)
{
    auto k = k_wh_tune<CZ, NT, PMAX>;
    set_lds(k, lds_bytes);
    WhArgs a = a0;
    // 2 signals per particle, one workgroup each: grids below 2^31 workgroups
    const uint32_t chunk = 1u << 29;
    for (uint32_t p0 = 0; p0 < n; p0 += chunk) {
        const uint32_t nb = std::min(chunk, n - p0);
        a.p0 = p0;
        hipLaunchKernelGGL(k, dim3(2 * nb), dim3(NT), lds_bytes, s, a);
        launch_check("k_wh_tune");
    }
}

// The whole-signal transform of one window; false (nothing launched) when the
// length has no whole-signal plan.
static bool fft_tune_whole(const double *hist, HistLayout lay, uint32_t n, const double *means, uint32_t from,
                           uint32_t N, const double *hann, double *out, uint32_t window, hipStream_t s)
{
    const WhTables *t = wh_tables(N, s);
    if (!t) return false;
    if (n == 0) return true;
    WhArgs a{};
    a.hist = hist; a.lay = lay; a.means = means; a.n = n; a.from = from; a.N = N; a.M = t->M;
    a.plan = t->plan; a.hann = hann; a.tw = t->tw; a.chirp = t->chirp; a.bhat = t->bhat;
    a.out = out; a.window = window;
    const size_t lds_bytes = (size_t)t->M * sizeof(Cplx);
    LaunchTimer timer("k_wh_tune", s);
    if (t->nt == kWhSmallThreads) {
        if (t->cz) launch_wh<true, kWhSmallThreads, kWhSmallPmax>(a, n, lds_bytes, s);
        else launch_wh<false, kWhSmallThreads, kWhSmallPmax>(a, n, lds_bytes, s);
    } else if (t->nt == kWhMidThreads) {
        if (t->cz) launch_wh<true, kWhMidThreads, kWhMidPmax>(a, n, lds_bytes, s);
        else launch_wh<false, kWhMidThreads, kWhMidPmax>(a, n, lds_bytes, s);
    } else {
        if (t->cz) launch_wh<true, kWhThreads, kWhPmax>(a, n, lds_bytes, s);
        else launch_wh<false, kWhThreads, kWhPmax>(a, n, lds_bytes, s);
    }
    return true;
}

bool fft_tune_mixed(const double *hist, HistLayout lay, uint32_t n, const double *means, uint32_t from, uint32_t N,
                    const double *hann, double *out, uint32_t window, hipStream_t s)
{
    if (N < 2) return false;
    if (fft_tune_whole(hist, lay, n, means, from, N, hann, out, window, s)) return true;
    const MrTables *t = mr_tables(N, s);
    if (!t) return false;
    if (n == 0) return true;
    const bool cz = t->chirp != nullptr;
    MrArgs a{};
    a.hist = hist; a.lay = lay; a.means = means; a.n = n; a.from = from; a.N = N; a.M = t->M;
    a.N1 = t->N1; a.N2 = t->N2; a.p1 = t->p1; a.p2 = t->p2;
    a.C = divisor_at_most(a.N2, kMrPmax / a.N1);
    a.RR = divisor_at_most(a.N1, kMrPmax / a.N2);
    a.hann = hann; a.tw1 = t->tw1; a.tw2 = t->tw2; a.twm = t->twm; a.chirp = t->chirp; a.bhat = t->bhat;
    a.out = out; a.window = window;
    const size_t lds_cols = (size_t)a.C * lds_len_odd(a.N1) * sizeof(Cplx);
    const size_t lds_rows = (size_t)a.RR * lds_len(a.N2) * sizeof(Cplx);
    // scratch per particle (both planes): the M-point work arrays, magnitudes, flags
    const uint64_t per_particle = 2 * ((uint64_t)sizeof(Cplx) * a.M + 8ull * N) + 8;
    // every grid below 2^31 threads: 2 signals per particle, up to per_sig
    // workgroups of kMrThreads per signal
    const uint64_t per_sig = std::max<uint64_t>({1, a.N2 / a.C, a.N1 / a.RR});
    const uint64_t max_batch = ((1ull << 31) / kMrThreads) / (2 * per_sig);
    const uint32_t batch =
        (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>({(uint64_t)n, mr_budget() / per_particle, max_batch}));
    const size_t nsig_max = 2 * (size_t)batch;
    const size_t y_bytes = nsig_max * (size_t)a.M * sizeof(Cplx), mag_bytes = nsig_max * (size_t)N * 8;
    char *ws = static_cast<char *>(workspace(y_bytes + mag_bytes + nsig_max * 4 + 256, 3));
    a.Y = reinterpret_cast<Cplx *>(ws);
    a.mags = reinterpret_cast<double *>(ws + y_bytes);
    a.flags = reinterpret_cast<uint32_t *>(ws + y_bytes + mag_bytes);
    if (cz) {
        set_lds(k_mr_columns<true>, lds_cols);
        set_lds(k_mr_rows<true>, lds_rows);
        set_lds(k_mr_final, lds_cols);
    } else {
        set_lds(k_mr_columns<false>, lds_cols);
        set_lds(k_mr_rows<false>, lds_rows);
    }
    for (uint32_t p0 = 0; p0 < n; p0 += batch) {
        const uint32_t nb = std::min(batch, n - p0);
        a.p0 = p0;
        a.nsig = 2 * nb;
        hipLaunchKernelGGL(k_mr_zero, dim3((a.nsig + 255) / 256), dim3(256), 0, s, a.flags, a.nsig);
        launch_check("k_mr_zero");
        // grids of 8 x (workgroups per XCD); see mr_virtual
        const uint32_t cols = a.nsig * (a.N2 / a.C), rows = a.nsig * (a.N1 / a.RR);
        const dim3 gc((cols + 7) / 8 * 8), gr((rows + 7) / 8 * 8);
        {
            LaunchTimer timer("k_mr_columns", s);
            if (cz) hipLaunchKernelGGL(k_mr_columns<true>, gc, dim3(kMrThreads), lds_cols, s, a);
            else hipLaunchKernelGGL(k_mr_columns<false>, gc, dim3(kMrThreads), lds_cols, s, a);
            launch_check("k_mr_columns");
        }
        {
            LaunchTimer timer("k_mr_rows", s);
            if (cz) hipLaunchKernelGGL(k_mr_rows<true>, gr, dim3(kMrThreads), lds_rows, s, a);
            else hipLaunchKernelGGL(k_mr_rows<false>, gr, dim3(kMrThreads), lds_rows, s, a);
            launch_check("k_mr_rows");
        }
        if (cz) {
            LaunchTimer timer("k_mr_final", s);
            hipLaunchKernelGGL(k_mr_final, gc, dim3(kMrThreads), lds_cols, s, a);
            launch_check("k_mr_final");
        }
        LaunchTimer timer("k_mr_tune", s);
        if (cz) hipLaunchKernelGGL(k_mr_tune<true>, dim3(a.nsig), dim3(kMrThreads), 0, s, a);
        else hipLaunchKernelGGL(k_mr_tune<false>, dim3(a.nsig), dim3(kMrThreads), 0, s, a);
        launch_check("k_mr_tune");
    }
    return true;
}

}  // namespace hm
