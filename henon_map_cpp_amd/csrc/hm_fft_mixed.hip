// FFT tunes of windows whose length has only the prime factors 2, 3, 5 and 7
// (src/dynamic_indicator.cpp:32-61, 178-196, 255-331 for any N): a mixed-radix
// four-step transform of the length itself, N = N1 x N2 (index n = N2 n1 + n2
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

bool fft_tune_mixed(const double *hist, HistLayout lay, uint32_t n, const double *means, uint32_t from, uint32_t N,
                    const double *hann, double *out, uint32_t window, hipStream_t s)
{
    if (N < 2) return false;
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
