// Device-side primitives of the 4D modulated Henon map, shared by every kernel.
//
// Two arithmetic modes:
//  * exact: the expression trees of the reference's map (src/henon.cu:113-173,
//    420-542) and tangent matrix (:175-222), evaluated without FMA contraction
//    (the library is compiled with -ffp-contract=off). With the same inputs and
//    tables this reproduces the reference's no-FMA CPU build bit for bit.
//  * fast: the same map with hand-placed fused multiply-adds (explicit fma(),
//    which -ffp-contract=off does not touch): 25 VALU ops per turn instead of 40
//    for mu != 0 (16 instead of 26 for mu == 0), within rounding of the exact mode.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "hm_log10.h"

namespace hm {

// One turn of the modulation table: sin/cos of omega_x(n), omega_y(n).
// Interleaved so one turn is one 32-byte record (a single scalar load).
struct __attribute__((aligned(32))) TabEntry {
    double sx, cx, sy, cy;
};

// kExactOrbit (indicator kernels only): the map and loss test of kExact, the
// tangent products and norms of kFast.
enum Mode : int { kExact = 0, kFast = 1, kExactOrbit = 2 };
constexpr int map_mode(int m) { return m == kFast ? kFast : kExact; }
constexpr int tangent_mode(int m) { return m == kExact ? kExact : kFast; }

// Pointer into the constant address space: for data that no wavefront writes
// during the launch. Wave-uniform loads through it become scalar loads.
template <class T>
using ConstPtr = const __attribute__((address_space(4))) T *;

template <class T>
__device__ __forceinline__ ConstPtr<T> as_const(const T *p)
{
    return (ConstPtr<T>)p;
}

__device__ __forceinline__ TabEntry load_record(ConstPtr<TabEntry> r)
{
    return TabEntry{r->sx, r->cx, r->sy, r->cy};
}

// ---------------------------------------------------------------- forward map
// henon.cu:131-173 (sextupole, octupole, rotation), forward direction.
template <bool OCT>
__device__ __forceinline__ void fwd_exact(double &x, double &px, double &y, double &py, double mu,
                                          const TabEntry &t)
{
    px = px + ((x * x) - (y * y));
    py = py + ((-2.0 * x) * y);
    if (OCT) {
        const double tx = mu * (((x * x) * x) - (((3.0 * y) * y) * x));
        const double ty = mu * (((y * y) * y) - (((3.0 * x) * x) * y));
        px = px + tx;
        py = py + ty;
    }
    const double xt = x, pt = px, yt = y, qt = py;
    x = (xt * t.cx) + (pt * t.sx);
    px = ((-xt) * t.sx) + (pt * t.cx);
    y = (yt * t.cy) + (qt * t.sy);
    py = ((-yt) * t.sy) + (qt * t.cy);
}

template <bool OCT>
__device__ __forceinline__ void fwd_fast(double &x, double &px, double &y, double &py, double mu, double m3mu,
                                         const TabEntry &t)
{
    // octupole as x (mu x^2 - 3mu y^2) and y (mu y^2 - 3mu x^2): every product
    // lands in a fresh register, so no fmac needs a copied accumulator
    if (OCT) {
        const double xx = x * x, yy = y * y;
        const double tx = fma(m3mu, yy, mu * xx);
        const double ty = fma(m3mu, xx, mu * yy);
        px = fma(x, tx, fma(x, x, fma(-y, y, px)));
        py = fma(y, ty, fma(-2.0 * x, y, py));
    } else {
        px = fma(x, x, fma(-y, y, px));
        py = fma(-2.0 * x, y, py);
    }
    const double xt = x, yt = y;
    x = fma(xt, t.cx, px * t.sx);
    px = fma(-xt, t.sx, px * t.cx);
    y = fma(yt, t.cy, py * t.sy);
    py = fma(-yt, t.sy, py * t.cy);
}

// Exact forward turn that takes x*x, y*y of the incoming point and returns those
// of the outgoing one: the loss test of turn k and the sextupole of turn k+1
// multiply the same numbers, so carrying the products saves two multiplies per
// turn without changing a single rounding. Returns the screening radius
// (screen_radius2 below) of the outgoing point.
template <bool OCT>
__device__ __forceinline__ double fwd_exact_carry(double &x, double &px, double &y, double &py, double &xx,
                                                  double &yy, double mu, const TabEntry &t)
{
    px = px + (xx - yy);
    py = py + ((-2.0 * x) * y);
    if (OCT) {
        const double tx = mu * ((xx * x) - (((3.0 * y) * y) * x));
        const double ty = mu * ((yy * y) - (((3.0 * x) * x) * y));
        px = px + tx;
        py = py + ty;
    }
    const double xt = x, pt = px, yt = y, qt = py;
    x = (xt * t.cx) + (pt * t.sx);
    px = ((-xt) * t.sx) + (pt * t.cx);
    y = (yt * t.cy) + (qt * t.sy);
    py = ((-yt) * t.sy) + (qt * t.cy);
    xx = x * x;
    yy = y * y;
    return fma(px, px, xx) + fma(py, py, yy);
}

// Loss screening for the exact mode. The reference's radius
// ((x^2 + px^2) + (y^2 + py^2)) (henon.cu:113-116) and the fused form
// fma(px, px, x^2) + fma(py, py, y^2) each lie within 3 roundings (3 * 2^-53,
// relative) of the true sum of squares, so below b2 * (1 - 2^-44) the fused
// value proves the reference's test `r2 < b2` true. Only points at or beyond the
// screen (or NaN) need the reference's expression, evaluated by exact_radius2
// to decide the loss: the decision is bit-for-bit the reference's, at 3 VALU ops
// a turn instead of 5.
__device__ __forceinline__ double screen_bound(double b2) { return b2 * (1.0 - 0x1p-44); }

__device__ __forceinline__ double screen_radius2(double x, double px, double y, double py)
{
    return fma(px, px, x * x) + fma(py, py, y * y);
}

__device__ __forceinline__ double exact_radius2(double xx, double px, double yy, double py)
{
    return (xx + (px * px)) + (yy + (py * py));  // henon.cu:113-116, with xx = x*x, yy = y*y
}

// ---------------------------------------------------------------- inverse map
// henon.cu:453-469: inverse rotations, then the inverse kicks.
template <bool OCT>
__device__ __forceinline__ void inv_exact(double &x, double &px, double &y, double &py, double mu,
                                          const TabEntry &t)
{
    const double xt = x, pt = px, yt = y, qt = py;
    x = (xt * t.cx) - (pt * t.sx);
    px = (xt * t.sx) + (pt * t.cx);
    y = (yt * t.cy) - (qt * t.sy);
    py = (yt * t.sy) + (qt * t.cy);
    if (OCT) {
        const double tx = mu * (((x * x) * x) - (((3.0 * y) * y) * x));
        const double ty = mu * (((y * y) * y) - (((3.0 * x) * x) * y));
        px = px - tx;
        py = py - ty;
    }
    px = px - ((x * x) - (y * y));
    py = py - ((-2.0 * x) * y);
}

template <bool OCT>
__device__ __forceinline__ void inv_fast(double &x, double &px, double &y, double &py, double mu, double m3mu,
                                         const TabEntry &t)
{
    const double xt = x, yt = y;
    x = fma(xt, t.cx, -(px * t.sx));
    px = fma(xt, t.sx, px * t.cx);
    y = fma(yt, t.cy, -(py * t.sy));
    py = fma(yt, t.sy, py * t.cy);
    if (OCT) {
        const double xx = x * x, yy = y * y;
        const double tx = fma(m3mu, yy, mu * xx);
        const double ty = fma(m3mu, xx, mu * yy);
        px = fma(-x, tx, fma(-x, x, fma(y, y, px)));
        py = fma(-y, ty, fma(2.0 * x, y, py));
    } else {
        px = fma(-x, x, fma(y, y, px));
        py = fma(2.0 * x, y, py);
    }
}

// One turn. m3mu = -3 mu is used by the fast mode only.
template <int MODE, bool INV, bool OCT>
__device__ __forceinline__ void map_step(double &x, double &px, double &y, double &py, double mu, double m3mu,
                                         const TabEntry &t)
{
    if (MODE == kExact) {
        if (INV) inv_exact<OCT>(x, px, y, py, mu, t);
        else fwd_exact<OCT>(x, px, y, py, mu, t);
    } else {
        if (INV) inv_fast<OCT>(x, px, y, py, mu, m3mu, t);
        else fwd_fast<OCT>(x, px, y, py, mu, m3mu, t);
    }
}

// henon.cu:113-116: survive iff (x^2+px^2)+(y^2+py^2) < barrier^2
template <int MODE>
__device__ __forceinline__ double radius2(double x, double px, double y, double py)
{
    if (MODE == kExact) return ((x * x) + (px * px)) + ((y * y) + (py * py));
    return fma(x, x, fma(px, px, fma(y, y, py * py)));
}

// One forward turn and the loss test, with hm_track's decisions: exact mode
// carries x*x, y*y (xx, yy) between turns and screens the radius (b2s =
// screen_bound(b2)), confirming candidates with the reference's expression;
// fast mode tests its fused radius. Returns true when the point leaves the
// barrier (or is NaN).
template <int MODE, bool OCT>
__device__ __forceinline__ bool forward_turn_lost(double &x, double &px, double &y, double &py, double &xx, double &yy,
                                                  double mu, double m3mu, const TabEntry &t, double b2, double b2s)
{
    if (MODE == kExact) {
        const double r2 = fwd_exact_carry<OCT>(x, px, y, py, xx, yy, mu, t);
        if (r2 < b2s) return false;
        return !(exact_radius2(xx, px, yy, py) < b2);
    }
    map_step<kFast, false, OCT>(x, px, y, py, mu, m3mu, t);
    return !(radius2<kFast>(x, px, y, py) < b2);
}

// forward_turn_lost without branches: both radius tests are evaluated (exact
// mode: the screen, then the reference's expression), so a plain turn stays one
// basic block. Same decisions as forward_turn_lost.
template <int MODE, bool OCT>
__device__ __forceinline__ bool forward_turn_out(double &x, double &px, double &y, double &py, double &xx, double &yy,
                                                 double mu, double m3mu, const TabEntry &t, double b2, double b2s)
{
    if (MODE == kExact) {
        const double r2 = fwd_exact_carry<OCT>(x, px, y, py, xx, yy, mu, t);
        const bool screen = !(r2 < b2s);
        const bool out = !(exact_radius2(xx, px, yy, py) < b2);
        return screen & out;
    }
    map_step<kFast, false, OCT>(x, px, y, py, mu, m3mu, t);
    return !(radius2<kFast>(x, px, y, py) < b2);
}

__device__ __forceinline__ bool any_nan(double x, double px, double y, double py)
{
    return __builtin_isnan(x) || __builtin_isnan(px) || __builtin_isnan(y) || __builtin_isnan(py);
}

// ---------------------------------------------------------------- tangent map
// henon.cu:175-196: J of one forward turn at the pre-step (x, y), row-major.
// a = (2x + 3mu x^2) - 3mu y^2 and b = -2y - 6mu x y; the y-row term is -a.
struct Jac {
    double a, b, an;  // J = [[c+s a, s, s b, 0], [-s+c a, c, c b, 0],
                      //      [sy b, 0, cy+sy an, sy], [cy b, 0, -sy+cy an, cy]]
};

template <int MODE>
__device__ __forceinline__ Jac jac_coeffs(double x, double y, double mu3, double mu6)
{
    Jac j;
    if (MODE == kExact) {
        j.a = ((2.0 * x) + ((mu3 * x) * x)) - ((mu3 * y) * y);
        j.b = (-2.0 * y) - ((mu6 * x) * y);
        // the reference evaluates the y-row term separately; equal to -a up to
        // the sign of an exact zero
        j.an = ((-2.0 * x) - ((mu3 * x) * x)) + ((mu3 * y) * y);
    } else {
        j.a = fma(mu3 * x, x, fma(-mu3 * y, y, 2.0 * x));
        j.b = fma(-mu6 * x, y, -2.0 * y);
        j.an = -j.a;
    }
    return j;
}

// Dense J entries in the reference's exact expression order.
__device__ __forceinline__ void jac_dense_exact(const Jac &j, const TabEntry &t, double *m)
{
    const double an = j.an;
    m[0] = t.cx + (t.sx * j.a);  m[1] = t.sx;  m[2] = t.sx * j.b;  m[3] = 0.0;
    m[4] = -t.sx + (t.cx * j.a); m[5] = t.cx;  m[6] = t.cx * j.b;  m[7] = 0.0;
    m[8] = t.sy * j.b;  m[9] = 0.0;  m[10] = t.cy + (t.sy * an);  m[11] = t.sy;
    m[12] = t.cy * j.b; m[13] = 0.0; m[14] = -t.sy + (t.cy * an); m[15] = t.cy;
}

}  // namespace hm
