// Shadow ("ghost") particles: the displacement of each ghost from its base
// particle, and the renormalisation of that displacement.
//
// Layout (the reference's, src/henon.cu:737-817): the particle arrays hold the
// n_base base particles first, then ghost set 0 (one ghost per base particle,
// in base order), ghost set 1, ...; entry n_base * (1 + g) + b is ghost g of
// base b. The arrays are the same device SoA that tracking advances, so ghosts
// are tracked like any other particle.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

#include "hm_internal.h"

namespace hm {

// henon.cu:407-410: Euclidean distance in the 4D phase space, summed left to right.
__device__ __forceinline__ double displacement(double x1, double px1, double y1, double py1, double x2, double px2,
                                               double y2, double py2)
{
    return sqrt((((x1 - x2) * (x1 - x2) + (px1 - px2) * (px1 - px2)) + (y1 - y2) * (y1 - y2)) +
                (py1 - py2) * (py1 - py2));
}

// henon.cu:412-418: move a point along its displacement from a centre so that
// the displacement's length goes from initial to final (the ratio is formed per
// coordinate, as there).
__device__ __forceinline__ void realign(double &x, double &px, double &y, double &py, double xc, double pxc, double yc,
                                        double pyc, double initial, double final_)
{
    x = xc + (x - xc) * (final_ / initial);
    px = pxc + (px - pxc) * (final_ / initial);
    y = yc + (y - yc) * (final_ / initial);
    py = pyc + (py - pyc) * (final_ / initial);
}

// henon.cu:1257-1282 (CPU: 826-858): one thread per ghost; bases are read only.
__global__ void k_ghost_renormalize(double *__restrict__ x, double *__restrict__ px, double *__restrict__ y,
                                    double *__restrict__ py, uint64_t n_base, uint64_t n_ghost, double target)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_ghost) return;
    const uint64_t b = t % n_base, i = n_base + t;
    const double xb = x[b], pxb = px[b], yb = y[b], pyb = py[b];
    double gx = x[i], gpx = px[i], gy = y[i], gpy = py[i];
    const double module = displacement(gx, gpx, gy, gpy, xb, pxb, yb, pyb);
    realign(gx, gpx, gy, gpy, xb, pxb, yb, pyb, module, target);
    x[i] = gx; px[i] = gpx; y[i] = gy; py[i] = gpy;
}

// henon.cu:1284-1313 (module) and 1337-1389 (direction): NaN when the ghost or
// its base holds a NaN; the direction is the ghost realigned to distance 1 minus
// the base. Written in the layout the accessors return: module [n_base][sets],
// direction [4][n_base][sets].
__global__ void k_ghost_displacement(const double *__restrict__ x, const double *__restrict__ px,
                                     const double *__restrict__ y, const double *__restrict__ py, uint64_t n_base,
                                     uint64_t sets, double *__restrict__ module_out, double *__restrict__ dir_out)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_base * sets) return;
    const uint64_t g = t / n_base, b = t - g * n_base, i = n_base * (1 + g) + b;
    const uint64_t o = b * sets + g;
    const double xb = x[b], pxb = px[b], yb = y[b], pyb = py[b];
    double gx = x[i], gpx = px[i], gy = y[i], gpy = py[i];
    const bool nan_in = isnan(gx) || isnan(gpx) || isnan(gy) || isnan(gpy) || isnan(xb) || isnan(pxb) ||
                        isnan(yb) || isnan(pyb);
    const double q = __builtin_nan("");
    const double module = displacement(gx, gpx, gy, gpy, xb, pxb, yb, pyb);
    if (module_out) module_out[o] = nan_in ? q : module;
    if (dir_out) {
        realign(gx, gpx, gy, gpy, xb, pxb, yb, pyb, module, 1.0);
        const uint64_t plane = n_base * sets;
        dir_out[o] = nan_in ? q : gx - xb;
        dir_out[plane + o] = nan_in ? q : gpx - pxb;
        dir_out[2 * plane + o] = nan_in ? q : gy - yb;
        dir_out[3 * plane + o] = nan_in ? q : gpy - pyb;
    }
}

// henon_map_cpp/dynamic_indicators.py:154-366 (SALI, GALI): the deviation of
// ghost set k from base b normalised as there, d_k = (g - b) / |g - b|, with
// |.| = sqrt(((x^2 + px^2) + y^2) + py^2).
//   SALI (kind 0, sets 0-1): min(|d_0 + d_1|, |d_0 - d_1|)
//   GALI (kind 1, sets 0-3): |det [d_0 d_1 d_2 d_3]|, the product of the
//                            singular values the reference takes from an SVD
// NaN where any deviation holds a NaN.
__device__ __forceinline__ void deviation(const double *__restrict__ x, const double *__restrict__ px,
                                          const double *__restrict__ y, const double *__restrict__ py, uint64_t i,
                                          const double (&b)[4], double (&d)[4])
{
    d[0] = x[i] - b[0];
    d[1] = px[i] - b[1];
    d[2] = y[i] - b[2];
    d[3] = py[i] - b[3];
    const double norm = sqrt(((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]) + d[3] * d[3]);
    for (int c = 0; c < 4; c++) d[c] = d[c] / norm;
}

__global__ void k_ghost_alignment(const double *__restrict__ x, const double *__restrict__ px,
                                  const double *__restrict__ y, const double *__restrict__ py, uint64_t n_base,
                                  int kind, double *__restrict__ out)
{
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n_base) return;
    const double base[4] = {x[b], px[b], y[b], py[b]};
    const double q = __builtin_nan("");
    if (kind == 0) {
        double d0[4], d1[4];
        deviation(x, px, y, py, n_base + b, base, d0);
        deviation(x, px, y, py, 2 * n_base + b, base, d1);
        const double plus = sqrt((((d0[0] + d1[0]) * (d0[0] + d1[0]) + (d0[1] + d1[1]) * (d0[1] + d1[1])) +
                                  (d0[2] + d1[2]) * (d0[2] + d1[2])) +
                                 (d0[3] + d1[3]) * (d0[3] + d1[3]));
        const double minus = sqrt((((d0[0] - d1[0]) * (d0[0] - d1[0]) + (d0[1] - d1[1]) * (d0[1] - d1[1])) +
                                   (d0[2] - d1[2]) * (d0[2] - d1[2])) +
                                  (d0[3] - d1[3]) * (d0[3] - d1[3]));
        out[b] = (isnan(plus) || isnan(minus)) ? q : (plus < minus ? plus : minus);
        return;
    }
    // a[r][c]: component r of deviation c; |det| by elimination with partial
    // pivoting
    double a[4][4];
    bool nan_in = false;
    for (int c = 0; c < 4; c++) {
        double d[4];
        deviation(x, px, y, py, (uint64_t)(c + 1) * n_base + b, base, d);
        for (int r = 0; r < 4; r++) {
            a[r][c] = d[r];
            nan_in = nan_in || isnan(d[r]);
        }
    }
    double det = 1.0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        int p = k;
#pragma unroll
        for (int r = k + 1; r < 4; r++)
            if (fabs(a[r][k]) > fabs(a[p][k])) p = r;
        if (p != k) {
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const double t = a[k][c];
                a[k][c] = a[p][c];
                a[p][c] = t;
            }
        }
        det = det * a[k][k];
        if (a[k][k] != 0.0) {
#pragma unroll
            for (int r = k + 1; r < 4; r++) {
                const double f = a[r][k] / a[k][k];
#pragma unroll
                for (int c = k + 1; c < 4; c++) a[r][c] = a[r][c] - f * a[k][c];
            }
        }
    }
    out[b] = nan_in ? q : fabs(det);
}

}  // namespace hm

using namespace hm;

extern "C" int hm_ghost_renormalize(double *x, double *px, double *y, double *py, size_t n_base, size_t sets,
                                    double target, void *stream)
{
    return guarded([&] {
        if (n_base == 0 || sets == 0) return;
        if (!x || !px || !y || !py) throw Error(HM_ERR_ARG, "ghost_renormalize: NULL pointer");
        const uint64_t n_ghost = (uint64_t)n_base * sets;
        hipStream_t s = as_stream(stream);
        const uint64_t grid = (n_ghost + 255) / 256;
        if (grid > 0x7fffffffull) throw Error(HM_ERR_RANGE, "ghost_renormalize: too many ghosts");
        LaunchTimer timer("k_ghost_renormalize", s);
        hipLaunchKernelGGL(k_ghost_renormalize, dim3((uint32_t)grid), dim3(256), 0, s, x, px, y, py, (uint64_t)n_base,
                           n_ghost, target);
        launch_check("k_ghost_renormalize");
    });
}

extern "C" int hm_ghost_displacement(const double *x, const double *px, const double *y, const double *py,
                                     size_t n_base, size_t sets, double *module_out, double *direction_out,
                                     void *stream)
{
    return guarded([&] {
        if (n_base == 0 || sets == 0 || (!module_out && !direction_out)) return;
        if (!x || !px || !y || !py) throw Error(HM_ERR_ARG, "ghost_displacement: NULL pointer");
        const uint64_t n_ghost = (uint64_t)n_base * sets;
        hipStream_t s = as_stream(stream);
        const uint64_t grid = (n_ghost + 255) / 256;
        if (grid > 0x7fffffffull) throw Error(HM_ERR_RANGE, "ghost_displacement: too many ghosts");
        LaunchTimer timer("k_ghost_displacement", s);
        hipLaunchKernelGGL(k_ghost_displacement, dim3((uint32_t)grid), dim3(256), 0, s, x, px, y, py,
                           (uint64_t)n_base, (uint64_t)sets, module_out, direction_out);
        launch_check("k_ghost_displacement");
    });
}

extern "C" int hm_ghost_alignment(const double *x, const double *px, const double *y, const double *py, size_t n_base,
                                  size_t sets, int kind, double *out, void *stream)
{
    return guarded([&] {
        if (kind != HM_ALIGN_SALI && kind != HM_ALIGN_GALI) throw Error(HM_ERR_ARG, "alignment: unknown kind");
        if (sets < (kind == HM_ALIGN_SALI ? 2u : 4u))
            throw Error(HM_ERR_ARG, "alignment: SALI needs 2 ghost sets, GALI 4");
        if (n_base == 0) return;
        if (!x || !px || !y || !py || !out) throw Error(HM_ERR_ARG, "alignment: NULL pointer");
        hipStream_t s = as_stream(stream);
        const uint64_t grid = ((uint64_t)n_base + 255) / 256;
        if (grid > 0x7fffffffull) throw Error(HM_ERR_RANGE, "alignment: too many particles");
        LaunchTimer timer("k_ghost_alignment", s);
        hipLaunchKernelGGL(k_ghost_alignment, dim3((uint32_t)grid), dim3(256), 0, s, x, px, y, py, (uint64_t)n_base,
                           kind, out);
        launch_check("k_ghost_alignment");
    });
}
