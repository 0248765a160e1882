// Device-side tangent-map helpers shared by the indicator kernels (single
// tracker and batched jobs): the structured product M <- J M of one turn's
// Jacobian (src/henon.cu:175-222, :251-265) and squared Frobenius norms.
#pragma once
#include <hip/hip_runtime.h>

#include "hm_device.h"

namespace hm {

// M <- J M with J's structure: rows 0-1 touch M rows 0,1,2; rows 2-3 M rows 0,2,3.
template <int MODE>
__device__ __forceinline__ void mat_mul_structured(const Jac &j, const TabEntry &t, double *m)
{
    double j00, j02, j10, j12, j20, j22, j30, j32;
    if (MODE == kExact) {
        j00 = t.cx + (t.sx * j.a); j02 = t.sx * j.b;
        j10 = -t.sx + (t.cx * j.a); j12 = t.cx * j.b;
        j20 = t.sy * j.b; j22 = t.cy + (t.sy * j.an);
        j30 = t.cy * j.b; j32 = -t.sy + (t.cy * j.an);
    } else {
        j00 = fma(t.sx, j.a, t.cx); j02 = t.sx * j.b;
        j10 = fma(t.cx, j.a, -t.sx); j12 = t.cx * j.b;
        j20 = t.sy * j.b; j22 = fma(-t.sy, j.a, t.cy);
        j30 = t.cy * j.b; j32 = fma(-t.cy, j.a, -t.sy);
    }
    double r[16];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const double m0 = m[c], m1 = m[4 + c], m2 = m[8 + c], m3 = m[12 + c];
        if (MODE == kExact) {
            r[c] = ((j00 * m0) + (t.sx * m1)) + (j02 * m2);
            r[4 + c] = ((j10 * m0) + (t.cx * m1)) + (j12 * m2);
            r[8 + c] = ((j20 * m0) + (j22 * m2)) + (t.sy * m3);
            r[12 + c] = ((j30 * m0) + (j32 * m2)) + (t.cy * m3);
        } else {
            r[c] = fma(j00, m0, fma(t.sx, m1, j02 * m2));
            r[4 + c] = fma(j10, m0, fma(t.cx, m1, j12 * m2));
            r[8 + c] = fma(j20, m0, fma(j22, m2, t.sy * m3));
            r[12 + c] = fma(j30, m0, fma(j32, m2, t.cy * m3));
        }
    }
#pragma unroll
    for (int q = 0; q < 16; q++) m[q] = r[q];
}

// Squared Frobenius norm with fused multiply-adds: feeds only the MEGNO ratio
// (held to a tolerance); the renormalisation factor, which changes M itself,
// uses frob2<MODE>.
__device__ __forceinline__ double frob2_fused(const double *m)
{
    // four interleaved partial sums: a dependency depth of 6 instead of 16
    double s0 = m[0] * m[0], s1 = m[1] * m[1], s2 = m[2] * m[2], s3 = m[3] * m[3];
#pragma unroll
    for (int q = 4; q < 16; q += 4) {
        s0 = fma(m[q], m[q], s0);
        s1 = fma(m[q + 1], m[q + 1], s1);
        s2 = fma(m[q + 2], m[q + 2], s2);
        s3 = fma(m[q + 3], m[q + 3], s3);
    }
    return (s0 + s1) + (s2 + s3);
}

template <int MODE>
__device__ __forceinline__ double frob2(const double *m)
{
    double s = m[0] * m[0];
    if (MODE == kExact) {
#pragma unroll
        for (int q = 1; q < 16; q++) s = s + (m[q] * m[q]);
    } else {
#pragma unroll
        for (int q = 1; q < 16; q++) s = fma(m[q], m[q], s);
    }
    return s;
}

}  // namespace hm
