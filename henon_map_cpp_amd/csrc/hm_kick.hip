// Tracking with the random 4D kick (src/henon.cu:362-405, 420-542): before the
// forward map (after the inverse map) every particle is moved by kick_module
// along a direction drawn from four standard normals. The reference draws them
// from per-thread cuRAND states seeded with the clock, so its kicks cannot be
// reproduced; here they come from a counter-based generator, Philox-4x32-10
// (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
// SC'11), keyed by a 64-bit seed with the counter (particle, draw): the same
// seed and draw give the same kicks, and calls that continue the draw index
// continue the stream. Two Philox blocks per kick give four 53-bit uniforms,
// turned into normals by the Box-Muller transform.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

#include "hm_device.h"
#include "hm_internal.h"

namespace hm {

struct Philox4 {
    uint32_t v[4];
};

// Philox-4x32 with 10 rounds (the published multipliers and Weyl key increments).
__device__ __forceinline__ Philox4 philox4x32_10(Philox4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.v[0], p1 = (uint64_t)0xCD9E8D57u * c.v[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0, hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c = Philox4{{hi1 ^ c.v[1] ^ k0, lo1, hi0 ^ c.v[3] ^ k1, lo0}};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// 53-bit uniform in [0, 1) from two 32-bit words
__device__ __forceinline__ double uniform53(uint32_t a, uint32_t b)
{
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

// Four standard normals for particle i's draw d.
__device__ __forceinline__ void normals4(uint64_t seed, uint64_t i, uint64_t d, double (&n)[4])
{
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
#pragma unroll
    for (uint32_t half = 0; half < 2; ++half) {
        const Philox4 r = philox4x32_10(
            Philox4{{(uint32_t)i, (uint32_t)(i >> 32), (uint32_t)d, ((uint32_t)(d >> 32) << 1) | half}}, k0, k1);
        const double u1 = 1.0 - uniform53(r.v[0], r.v[1]);  // (0, 1]
        const double u2 = uniform53(r.v[2], r.v[3]);
        const double rad = sqrt(-2.0 * log(u1));
        double s, c;
        sincos(2.0 * M_PI * u2, &s, &c);
        n[2 * half] = rad * c;
        n[2 * half + 1] = rad * s;
    }
}

// henon.cu:386-405 with the normals above.
__device__ __forceinline__ void random_kick(double &x, double &px, double &y, double &py, double module, uint64_t seed,
                                            uint64_t i, uint64_t d)
{
    double k[4];
    normals4(seed, i, d, k);
    const double mod = sqrt(((k[0] * k[0] + k[1] * k[1]) + k[2] * k[2]) + k[3] * k[3]);
    x += k[0] * module / mod;
    px += k[1] * module / mod;
    y += k[2] * module / mod;
    py += k[3] * module / mod;
}

struct KickArgs {
    double *x, *px, *y, *py;
    uint32_t *steps;
    uint64_t n;
    const TabEntry *tab;
    uint32_t tab_len, n_turns;
    double mu, b2, module;
    uint64_t seed, draw;
    uint32_t *status;
};

// One thread per particle, per turn as henon_step (henon.cu:482-542): stop at
// NaN, kick then map (inverse: map then kick), count the turn inside the
// barrier, else set NaN. A turn index outside the table stops the particle and
// raises status bit 0.
template <int MODE, bool INV, bool OCT>
__global__ void __launch_bounds__(256) k_track_kicked(KickArgs a)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double x = a.x[i], px = a.px[i], y = a.y[i], py = a.py[i];
    uint32_t s = a.steps[i];
    if (any_nan(x, px, y, py)) return;
    const double mu = a.mu, m3mu = -3.0 * mu;
    for (uint32_t k = 0; k < a.n_turns; ++k) {
        const uint32_t ti = INV ? s - 1 : s;
        if ((INV && s == 0) || ti >= a.tab_len) {
            atomicOr(a.status, 1u);
            break;
        }
        const TabEntry t = a.tab[ti];
        if (!INV) random_kick(x, px, y, py, a.module, a.seed, i, a.draw + k);
        map_step<MODE, INV, OCT>(x, px, y, py, mu, m3mu, t);
        if (INV) random_kick(x, px, y, py, a.module, a.seed, i, a.draw + k);
        if (radius2<MODE>(x, px, y, py) < a.b2) {
            s = INV ? s - 1 : s + 1;
        } else {
            x = px = y = py = __builtin_nan("");
            break;
        }
    }
    a.x[i] = x; a.px[i] = px; a.y[i] = y; a.py[i] = py;
    a.steps[i] = s;
}

}  // namespace hm

using namespace hm;

extern "C" int hm_track_kicked(double *x, double *px, double *y, double *py, uint32_t *steps, size_t n,
                               const double *table, size_t table_len, uint32_t n_turns, double mu, double barrier,
                               double kick_module, int inverse, int mode, uint64_t seed, uint64_t draw, void *stream)
{
    return guarded([&] {
        if (mode != HM_MODE_EXACT && mode != HM_MODE_FAST) throw Error(HM_ERR_ARG, "track_kicked: unknown mode");
        if (!std::isfinite(kick_module)) throw Error(HM_ERR_ARG, "track_kicked: kick_module must be finite");
        if (n == 0 || n_turns == 0) return;
        if (!x || !px || !y || !py || !steps || !table) throw Error(HM_ERR_ARG, "track_kicked: NULL pointer");
        if (table_len > UINT32_MAX) throw Error(HM_ERR_RANGE, "track_kicked: table too long");
        KickArgs a{x, px, y, py, steps, (uint64_t)n, reinterpret_cast<const TabEntry *>(table), (uint32_t)table_len,
                   n_turns, mu, barrier * barrier, kick_module, seed, draw, status_word()};
        hipStream_t s = as_stream(stream);
        const uint64_t grid = ((uint64_t)n + 255) / 256;
        if (grid > 0x7fffffffull) throw Error(HM_ERR_RANGE, "track_kicked: too many particles");
        const bool oct = mu != 0.0, inv = inverse != 0, exact = mode == HM_MODE_EXACT;
        using K = void (*)(KickArgs);
        K k;
        if (exact) {
            if (inv) k = oct ? k_track_kicked<kExact, true, true> : k_track_kicked<kExact, true, false>;
            else k = oct ? k_track_kicked<kExact, false, true> : k_track_kicked<kExact, false, false>;
        } else {
            if (inv) k = oct ? k_track_kicked<kFast, true, true> : k_track_kicked<kFast, true, false>;
            else k = oct ? k_track_kicked<kFast, false, true> : k_track_kicked<kFast, false, false>;
        }
        LaunchTimer timer("k_track_kicked", s);
        hipLaunchKernelGGL(k, dim3((uint32_t)grid), dim3(256), 0, s, a);
        launch_check("k_track_kicked");
    });
}
