// Pieces shared by the tune kernels (hm_fft.hip: trajectory stores, Birkhoff
// and power-of-two FFT tunes; hm_chirpz.hip: FFT tunes of any length): the two
// history layouts, complex arithmetic, the small DFTs of the Stockham passes
// and the peak interpolation of src/dynamic_indicator.cpp:134-176.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

#include "hm_device.h"

namespace hm {

// Element (sample m, coordinate c, particle p) of a history at m*ms + c*cs + p*ps.
struct HistLayout {
    uint64_t ms, cs, ps;
};
__host__ __device__ inline HistLayout sample_major(uint64_t n) { return HistLayout{4 * n, n, 1}; }
__host__ __device__ inline HistLayout particle_rows(uint64_t L) { return HistLayout{1, L, 4 * L}; }
__device__ __forceinline__ size_t hat(const HistLayout &h, uint64_t m, uint64_t c, uint64_t p)
{
    return (size_t)(m * h.ms + c * h.cs + p * h.ps);
}

// 16-byte aligned, so that LDS and global accesses of one value are single
// 128-bit operations (ds_read_b128, not two 64-bit halves a lane-stride apart)
struct alignas(16) Cplx {
    double re, im;
};
__device__ __forceinline__ Cplx cadd(Cplx a, Cplx b) { return Cplx{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ Cplx csub(Cplx a, Cplx b) { return Cplx{a.re - b.re, a.im - b.im}; }
// complex product with fused multiply-adds (the tune transforms are held to a
// tolerance, not to the reference's rounding; the files build with contraction off)
__device__ __forceinline__ Cplx cmul(Cplx a, Cplx w)
{
    return Cplx{fma(a.re, w.re, -(a.im * w.im)), fma(a.re, w.im, a.im * w.re)};
}
__device__ __forceinline__ Cplx mul_mi(Cplx a) { return Cplx{a.im, -a.re}; }  // times -i

// forward DFT of sizes 2, 4, 8 (exp(-2 pi i m k / R))
__device__ __forceinline__ void dft2(Cplx *v)
{
    const Cplx a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}
__device__ __forceinline__ void dft4(Cplx &a, Cplx &b, Cplx &c, Cplx &d)
{
    const Cplx s0 = cadd(a, c), s1 = csub(a, c), s2 = cadd(b, d), s3 = mul_mi(csub(b, d));
    a = cadd(s0, s2);
    c = csub(s0, s2);
    b = cadd(s1, s3);
    d = csub(s1, s3);
}
__device__ __forceinline__ void dft8(Cplx *v)
{
    // even and odd halves, then X[k] = E[k] + w8^k O[k], X[k+4] = E[k] - w8^k O[k]
    Cplx e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    Cplx o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4(e0, e1, e2, e3);
    dft4(o0, o1, o2, o3);
    constexpr double c = 0.70710678118654752440;  // sqrt(2) / 2
    const Cplx t1 = Cplx{c * (o1.re + o1.im), c * (o1.im - o1.re)};   // (1 - i)/sqrt2
    const Cplx t2 = mul_mi(o2);                                        // -i
    const Cplx t3 = Cplx{c * (o3.im - o3.re), -c * (o3.re + o3.im)};  // (-1 - i)/sqrt2
    v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
    v[1] = cadd(e1, t1); v[5] = csub(e1, t1);
    v[2] = cadd(e2, t2); v[6] = csub(e2, t2);
    v[3] = cadd(e3, t3); v[7] = csub(e3, t3);
}
template <int R>
__device__ __forceinline__ void dft_r(Cplx *v)
{
    if constexpr (R == 8) dft8(v);
    else if constexpr (R == 4) dft4(v[0], v[1], v[2], v[3]);
    else dft2(v);
}

__device__ __forceinline__ double interp_A(double a, double b, double c)
{
    // dynamic_indicator.cpp:134-137
    return (-(a + b * c) * (a - b) + b * sqrt(c * c * ((a + b) * (a + b)) - 2 * a * b * (2 * (c * c) - c - 1))) /
           ((a * a) + (b * b) + 2 * a * b * c);
}

// dynamic_indicator.cpp:139-176 on magnitudes mag[0..N): first maximum wins
// `mag(k)` gives the magnitude of spectrum entry k.
template <class Mag>
__device__ double interp_tune_at(Mag mag, uint32_t N, uint32_t index)
{
    const double Nd = (double)N;
    if (index == 0) return 1.0;
    if (index == N - 1) return 0.0;
    uint32_t i1, i2;
    if (mag(index - 1) > mag(index + 1)) {
        i1 = index - 1; i2 = index; index = index - 1;
    } else {
        i1 = index; i2 = index + 1;
    }
    const double value = ((double)index / Nd) +
                         (1.0 / (2.0 * M_PI)) * asin(interp_A(mag(i1), mag(i2), cos(2.0 * M_PI / Nd)) *
                                                     sin(2.0 * M_PI / Nd));
    return fabs(1.0 - value);
}

// FFT tunes of one window of any length N >= 2 (hm_chirpz.hip): Bluestein's
// chirp-z transform over a power-of-two length M >= 2N - 1, computed as two
// LDS-resident passes of sub-transforms (M = M1 x M2) with HBM scratch between
// them; for the windows fft_tune_mixed does not take. hist/lay/means/from as the other tune kernels; hann: device [N];
// out: device [.. window][2][n] (tune x, tune y), window `window`.
void fft_tune_chirpz(const double *hist, HistLayout lay, uint32_t n, const double *means, uint32_t from,
                     uint32_t N, const double *hann, double *out, uint32_t window, hipStream_t s);

// The same by mixed-radix transforms (hm_fft_mixed.hip): windows that fit one
// workgroup's LDS (about 10^4 points) in one kernel per signal (the length
// itself, with prime passes up to 257, or chirp-z over a smooth length that
// fits); longer ones as four-step transforms of N itself when its prime
// factors are 2, 3, 5 and 7 only and it splits into two factors of at most
// 2048, else Bluestein's chirp-z over the smallest such length M >= 2N - 1.
// Returns false (nothing launched) when no such M exists
// (N > 2^21 roughly).
bool fft_tune_mixed(const double *hist, HistLayout lay, uint32_t n, const double *means, uint32_t from, uint32_t N,
                    const double *hann, double *out, uint32_t window, hipStream_t s);

}  // namespace hm
