// log10 for the per-turn MEGNO sums, by the method of fdlibm's e_log.c and
// e_log10.c (Sun Microsystems, 1993; freely distributable): x = 2^k m with m in
// [sqrt(2)/2, sqrt(2)), f = m - 1, s = f / (2 + f),
//   ln m = 2 atanh(s) = f - s (f - R(s^2))            (or, for m near sqrt(2),
//        = f - (f^2/2 - s (f^2/2 + R(s^2))))           the form with f^2/2)
// with R the degree-7 minimax polynomial of fdlibm's log (Lg1..Lg7), and
//   log10 x = k log10(2) + ln(m) / ln(10),
// log10(2) split so that k log10_2hi is exact. Branch-free: s from a
// reciprocal refined by Newton steps (2 + f lies in [1.7, 2.5]); zero,
// negative, infinite and NaN arguments selected at the end (-inf, NaN, +inf,
// NaN); subnormal arguments are not handled (the MEGNO norm ratio is never one).
// At most 2 units in the last place from the C library's log10, 88 % of
// arguments correctly rounded (tests/log10_check.cpp). The library's
// double-double log10 takes about 90 instructions on gfx950 and a branch, this
// one about 50 in the same straight block of the indicator kernel's turn.
// Compiles as host code too (no HIP types), so a host check can build it with
// g++ alone; device and host differ only in the reciprocal's first estimate.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

#if defined(__HIPCC__)
#define HM_LOG10_FN __host__ __device__ inline
#else
#define HM_LOG10_FN inline
#endif

namespace hm {

HM_LOG10_FN double log10_fdl(double x)
{
    constexpr double Lg1 = 6.666666666666735130e-01, Lg2 = 3.999999999940941908e-01,
                     Lg3 = 2.857142874366239149e-01, Lg4 = 2.222219843214978396e-01,
                     Lg5 = 1.818357216161805012e-01, Lg6 = 1.531383769920937332e-01,
                     Lg7 = 1.479819860511658591e-01;
    constexpr double ivln10 = 4.34294481903251816668e-01, log10_2hi = 3.01029995663611771306e-01,
                     log10_2lo = 3.69423907715893078616e-13;
    uint64_t u;
    std::memcpy(&u, &x, sizeof u);
    const uint32_t hi = (uint32_t)(u >> 32);
    const uint32_t hx = hi & 0x000fffffu;
    const uint32_t half = (hx + 0x95f64u) & 0x100000u;
    const uint64_t mu = ((uint64_t)(hx | (half ^ 0x3ff00000u)) << 32) | (u & 0xffffffffull);
    double m;
    std::memcpy(&m, &mu, sizeof m);
    const int k = (int)((hi >> 20) & 0x7ffu) - 1023 + (int)(half >> 20);
    const double f = m - 1.0;
    const double d = 2.0 + f;
#if defined(__HIP_DEVICE_COMPILE__)
    double y = __builtin_amdgcn_rcp(d);  // hardware estimate; two Newton steps: full precision
#else
    double y = (double)(1.0f / (float)d);  // about 24 bits; two Newton steps: full precision
#endif
    y = fma(fma(-d, y, 1.0), y, y);
    y = fma(fma(-d, y, 1.0), y, y);
    double s = f * y;
    s = fma(fma(-d, s, f), y, s);
    const double z = s * s, w = z * z;
    const double t1 = w * fma(w, fma(w, Lg6, Lg4), Lg2);
    const double t2 = z * fma(w, fma(w, fma(w, Lg7, Lg5), Lg3), Lg1);
    const double R = t2 + t1;
    const bool near_sqrt2 = (int32_t)((hx - 0x6147au) | (0x6b851u - hx)) > 0;
    const double hfsq = 0.5 * f * f;
    const double lnm = near_sqrt2 ? f - (hfsq - s * (hfsq + R)) : f - s * (f - R);
    const double dk = (double)k;
    const double r = fma(dk, log10_2hi, fma(dk, log10_2lo, lnm * ivln10));
    const double bad = x == 0.0 ? -HUGE_VAL : (x > 0.0 ? x : NAN);  // +inf, NaN stay themselves
    return (x > 0.0 && x < HUGE_VAL) ? r : bad;
}

}  // namespace hm
