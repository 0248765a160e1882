// Oracle for the 4D modulated Henon-map tracking path.
//
// TEST INFRASTRUCTURE ONLY. This file restates, on the CPU, the algorithms of
// the reference's CPU path so that the HIP product can be checked against it.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// it; the product library never links, loads or calls it.
//
// Every function cites the reference file:line it follows (paths relative to the
// reference checkout). It is compiled with `g++ -O3 -ffp-contract=off` and no
// `-march`, the same x86-64 SSE2, no-FMA host arithmetic the reference's CPU
// build uses (CMakeLists.txt:13-15), so with identical inputs the map and the
// tangent products reproduce the reference bit for bit. Expression trees are
// written with explicit parentheses in the order C++ evaluates the reference's
// expressions.
//
// Pinning (what checks each part; see tests/test_oracle_*.py):
//   modulation       : bit-exact against oracle/_ref (the reference's own
//                      src/modulation.cpp compiled from source) and closed forms
//   map / track      : golden x after one turn (tests/test_construct.ipynb:638),
//                      golden tunes of a 1e4-turn orbit (tests/new_tune.ipynb),
//                      symplecticity of the Jacobian, inverse(forward) = id,
//                      linear-rotation limit, NaN/barrier semantics
//   tangent matrix   : golden J (tests/test_construct.ipynb:101-104,134-137),
//                      central finite differences of the map, det J = 1
//   birkhoff weights : golden birkhoff_weights(10) (test_construct.ipynb:379-381)
//   birkhoff tune    : golden full-window tunes (new_tune.ipynb), pure rotation
//   fft tune         : golden full-window tunes (new_tune.ipynb), numpy.fft for
//                      the DFT, analytic single tones
//   megno (trace)    : closed-form recursions on linear maps
//   constructs       : golden lyapunov multi values (test_construct.ipynb:328-345)
//   shadow particles : golden ghost coordinates and random directions
//                      (tests/test_new_standard.ipynb cells 47-54), unit length
//                      after renormalize(1), closed-form axis directions
//   fused indicators : (new algorithms) closed forms on rotations, finite
//   and adjoint        differences of the oracle map, tangent-map products

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdlib>
#include <vector>
#include <thread>
#include <algorithm>
#include <numeric>
#include <complex>
#include <limits>
#include <random>
#include <array>

namespace {

unsigned resolve_threads(int n_threads)
{
    if (n_threads > 0) return (unsigned)n_threads;
    unsigned hc = std::thread::hardware_concurrency();
    return hc == 0 ? 1u : hc;
}

template <class F>
void strided_parallel(size_t n, unsigned nt, F fn)
{
    // henon.cu:2720-2745: thread t owns particles j = t, t+nt, ...
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++)
        th.emplace_back([&, t]() { for (size_t j = t; j < n; j += nt) fn(j, t); });
    for (auto &t : th) t.join();
}

// ---- map primitives: henon.cu:113-173 -------------------------------------

inline bool check_barrier(double x, double px, double y, double py, double b2)
{
    // henon.cu:113-116
    return ((x * x) + (px * px)) + ((y * y) + (py * py)) < b2;
}

inline bool check_nan(double x, double px, double y, double py)
{
    // henon.cu:118-121
    return std::isnan(x) || std::isnan(px) || std::isnan(y) || std::isnan(py);
}

inline void sextupole(double x, double &px, double y, double &py, bool reverse)
{
    // henon.cu:131-143
    if (!reverse) {
        px = px + ((x * x) - (y * y));
        py = py + ((-2.0 * x) * y);
    } else {
        px = px - ((x * x) - (y * y));
        py = py - ((-2.0 * x) * y);
    }
}

inline void octupole(double x, double &px, double y, double &py, double mu, bool reverse)
{
    // henon.cu:145-157
    double tx = mu * (((x * x) * x) - (((3.0 * y) * y) * x));
    double ty = mu * (((y * y) * y) - (((3.0 * x) * x) * y));
    if (!reverse) { px = px + tx; py = py + ty; }
    else { px = px - tx; py = py - ty; }
}

inline void rotation(double &x, double &px, double s, double c, bool reverse)
{
    // henon.cu:159-173
    double xt = x, pt = px;
    if (!reverse) {
        x = (xt * c) + (pt * s);
        px = ((-xt) * s) + (pt * c);
    } else {
        x = (xt * c) - (pt * s);
        px = (xt * s) + (pt * c);
    }
}

// henon.cu:420-480 (host henon_step), without the random kick (kick_module NaN)
inline void henon_step(double &x, double &px, double &y, double &py, unsigned &steps,
                       const double *sx, const double *cx, const double *sy, const double *cy,
                       double b2, double mu, bool reverse)
{
    if (check_nan(x, px, y, py)) return;
    if (!reverse) {
        sextupole(x, px, y, py, false);
        if (mu != 0.0) octupole(x, px, y, py, mu, false);
        rotation(x, px, sx[steps], cx[steps], false);
        rotation(y, py, sy[steps], cy[steps], false);
    } else {
        rotation(x, px, sx[steps - 1], cx[steps - 1], true);
        rotation(y, py, sy[steps - 1], cy[steps - 1], true);
        if (mu != 0.0) octupole(x, px, y, py, mu, true);
        sextupole(x, px, y, py, true);
    }
    if (check_barrier(x, px, y, py, b2)) {
        steps += (reverse ? -1 : 1);
    } else {
        const double n = std::numeric_limits<double>::quiet_NaN();
        x = n; px = n; y = n; py = n;
    }
}

// henon.cu:198-222 (host tangent_matrix), row-major 4x4 into out[16]
inline void tangent_matrix(double x, double px, double y, double py,
                           double sx, double cx, double sy, double cy, double mu, double *m)
{
    (void)px; (void)py;
    double a = ((2 * x) + (((mu * 3) * x) * x)) - (((mu * 3) * y) * y);
    double b = (-2 * y) - (((mu * 6) * x) * y);
    double a_neg = ((-2 * x) - (((mu * 3) * x) * x)) + (((mu * 3) * y) * y);
    m[0] = cx + (sx * a);   m[1] = sx;   m[2] = sx * b;          m[3] = 0.0;
    m[4] = -sx + (cx * a);  m[5] = cx;   m[6] = cx * b;          m[7] = 0.0;
    m[8] = sy * b;          m[9] = 0.0;  m[10] = cy + (sy * a_neg); m[11] = sy;
    m[12] = cy * b;         m[13] = 0.0; m[14] = -sy + (cy * a_neg); m[15] = cy;
}

// henon.cu:224-249 (inverse_tangent_matrix)
inline void inverse_tangent_matrix(double x, double px, double y, double py,
                                   double sx, double cx, double sy, double cy, double mu, double *m)
{
    double u = (cx * x) - (px * sx);
    double v = (cy * y) - (py * sy);
    double d = std::pow(u, 2) - std::pow(v, 2);
    m[0] = cx; m[1] = -sx; m[2] = 0; m[3] = 0;
    m[4] = ((((-3 * cx) * mu) * d) - ((2 * cx) * u)) + sx;
    m[5] = (cx + (((3 * mu) * sx) * d)) + ((2 * sx) * u);
    m[6] = ((2 * cy) * v) * (((3 * mu) * u) + 1);
    m[7] = ((2 * sy) * v) * (((-3 * mu) * u) - 1);
    m[8] = 0; m[9] = 0; m[10] = cy; m[11] = -sy;
    m[12] = ((2 * cx) * v) * (((3 * mu) * u) + 1);
    m[13] = ((2 * sx) * v) * (((-3 * mu) * u) - 1);
    m[14] = ((((3 * cy) * mu) * d) + ((2 * cy) * u)) + sy;
    m[15] = (cy - (((3 * mu) * sy) * d)) - ((2 * sy) * u);
}

// henon.cu:251-265 (multiply_matrices): out = a * b, i-j-k order, += from 0.0
inline void multiply_matrices(const double *a, const double *b, double *out)
{
    double r[16];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            double acc = 0.0;
            for (int k = 0; k < 4; k++) acc += a[i * 4 + k] * b[k * 4 + j];
            r[i * 4 + j] = acc;
        }
    std::memcpy(out, r, sizeof(r));
}

}  // namespace

extern "C" {

// ---- modulation: src/modulation.cpp ----------------------------------------

// modulation.cpp:3-12 (basic_modulation; index i, not i+start: quirk kept)
int or_basic_modulation(double tune, double omega, double epsilon, int start, int end, double *out)
{
    if (!(start < end)) return 1;
    for (int i = 0; i < end - start; i++)
        out[i] = (2 * M_PI * tune * (1 + epsilon * std::cos(2 * M_PI * omega * i)));
    return 0;
}

// Coefficients and line frequencies of the SPS kind: modulation.cpp:14-28
static const double kSpsCoef[7] = {1.000e-4, 0.218e-4, 0.708e-4, 0.254e-4, 0.100e-4, 0.078e-4, 0.218e-4};
static const double kSpsFreq[7] = {1 * (2 * M_PI / 868.12), 2 * (2 * M_PI / 868.12), 3 * (2 * M_PI / 868.12),
                                   6 * (2 * M_PI / 868.12), 7 * (2 * M_PI / 868.12), 10 * (2 * M_PI / 868.12),
                                   12 * (2 * M_PI / 868.12)};

// modulation.cpp:30-60 with caller-supplied (coef, freq); the loop body and the
// accumulation order are those of sps_modulation.
int or_lines_modulation(double tune, double epsilon, const double *coef, const double *freq, int n_lines,
                        int start, int end, double *out)
{
    if (!(start < end)) return 1;
    for (unsigned int i = 0; i < (unsigned)(end - start); i++) {
        auto sum = 0.0;
        for (int j = 0; j < n_lines; j++) sum += coef[j] * std::cos(freq[j] * (i + start));
        out[i] = (2 * M_PI * tune * (1 + epsilon * sum));
    }
    return 0;
}

int or_sps_modulation(double tune, double epsilon, int start, int end, double *out)
{
    return or_lines_modulation(tune, epsilon, kSpsCoef, kSpsFreq, 7, start, end, out);
}

void or_sps_lines(double *coef, double *freq)
{
    std::memcpy(coef, kSpsCoef, sizeof(kSpsCoef));
    std::memcpy(freq, kSpsFreq, sizeof(kSpsFreq));
}

// henon.cu:2670-2692: sin/cos tables of omega_x, omega_y
void or_sin_cos(const double *w, size_t n, double *s, double *c)
{
    for (size_t i = 0; i < n; i++) { s[i] = std::sin(w[i]); c[i] = std::cos(w[i]); }
}

// ---- tracking: henon.cu:2701-2752 -------------------------------------------
// Stateless: the caller owns global_steps and the "turns too large" checks.
// Returns 2 if a non-NaN kick is requested (random kick is not restated).
int or_track(double *x, double *px, double *y, double *py, unsigned *steps, size_t n,
             const double *sx, const double *cx, const double *sy, const double *cy,
             unsigned n_turns, double mu, double barrier, double kick_module, int inverse, int n_threads)
{
    if (!std::isnan(kick_module)) return 2;
    const double b2 = barrier * barrier;  // henon.cu:2717
    strided_parallel(n, resolve_threads(n_threads), [&](size_t j, unsigned) {
        for (unsigned k = 0; k < n_turns; k++)
            henon_step(x[j], px[j], y[j], py[j], steps[j], sx, cx, sy, cy, b2, mu, inverse != 0);
    });
    return 0;
}

// ---- tangent map: henon.cu:175-265, 3080-3112 -------------------------------

void or_tangent_matrix(double x, double px, double y, double py, double sx, double cx, double sy, double cy,
                       double mu, double *out)
{
    tangent_matrix(x, px, y, py, sx, cx, sy, cy, mu, out);
}

void or_inverse_tangent_matrix(double x, double px, double y, double py, double sx, double cx, double sy,
                               double cy, double mu, double *out)
{
    inverse_tangent_matrix(x, px, y, py, sx, cx, sy, cy, mu, out);
}

void or_multiply_matrices(const double *a, const double *b, double *out) { multiply_matrices(a, b, out); }

// henon.cu:3080-3112: per-particle J (or J^-1) at table index global_steps;
// invalid particles (valid[i]==0) give a NaN matrix.
void or_get_tangent_matrix(const double *x, const double *px, const double *y, const double *py,
                           const uint8_t *valid, size_t n, const double *sx, const double *cx,
                           const double *sy, const double *cy, unsigned global_steps, double mu,
                           int reverse, double *out)
{
    const unsigned s = global_steps;
    for (size_t i = 0; i < n; i++) {
        double *m = out + i * 16;
        if (valid[i]) {
            if (!reverse) tangent_matrix(x[i], px[i], y[i], py[i], sx[s], cx[s], sy[s], cy[s], mu, m);
            else inverse_tangent_matrix(x[i], px[i], y[i], py[i], sx[s], cx[s], sy[s], cy[s], mu, m);
        } else {
            for (int k = 0; k < 16; k++) m[k] = std::numeric_limits<double>::quiet_NaN();
        }
    }
}

// henon.cu:267-295 (get_matrices_and_multiply): M <- J(x_i; table[global_steps]) * M
void or_structured_multiply(double *mats, const double *x, const double *px, const double *y, const double *py,
                            size_t n, const double *sx, const double *cx, const double *sy, const double *cy,
                            unsigned global_steps, double mu)
{
    const unsigned s = global_steps;
    for (size_t i = 0; i < n; i++) {
        double j[16];
        tangent_matrix(x[i], px[i], y[i], py[i], sx[s], cx[s], sy[s], cy[s], mu, j);
        multiply_matrices(j, mats + i * 16, mats + i * 16);
    }
}

// ---- CPU MEGNO (trace form): henon.cu:2755-2887 ------------------------------
// out: n x n_samples. carry_off_diagonal=1 reproduces the reference's per-thread
// matrix reuse (only the diagonal is reset to 1 between the particles a thread
// owns, henon.cu:2794-2803), which makes results depend on n_threads;
// carry_off_diagonal=0 starts every particle from the identity.
int or_megno(double *x, double *px, double *y, double *py, unsigned *steps, size_t n,
             const double *sx, const double *cx, const double *sy, const double *cy,
             unsigned n_turns, double mu, double barrier, int inverse,
             const unsigned *samples, size_t n_samples, int n_threads, int carry_off_diagonal, double *out)
{
    for (size_t k = 0; k < n_samples; k++)
        if (samples[k] > n_turns) return 1;  // the reference would read out of bounds
    const double b2 = barrier * barrier;
    const unsigned nt = resolve_threads(n_threads);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++) {
        th.emplace_back([&, t]() {
            std::vector<double> megno_vals(n_turns + 1, 0.0);
            double full[16] = {0};
            for (size_t j = t; j < n; j += nt) {
                if (!carry_off_diagonal) for (int q = 0; q < 16; q++) full[q] = 0.0;
                full[0] = 1.0; full[5] = 1.0; full[10] = 1.0; full[15] = 1.0;
                double trace_1 = 4.0, trace_2 = 4.0;
                for (unsigned k = 0; k < n_turns; k++) {
                    double jm[16];
                    tangent_matrix(x[j], px[j], y[j], py[j], sx[k], cx[k], sy[k], cy[k], mu, jm);
                    multiply_matrices(jm, full, full);
                    trace_2 = trace_1;
                    trace_1 = ((full[0] + full[5]) + full[10]) + full[15];
                    megno_vals[k] = (k + 1) * std::log10(trace_1 / trace_2);
                    henon_step(x[j], px[j], y[j], py[j], steps[j], sx, cx, sy, cy, b2, mu, inverse != 0);
                }
                double sum = 0.0;
                for (unsigned k = 0; k < n_turns; k++) { sum += megno_vals[k]; megno_vals[k] = sum; }
                for (unsigned k = 0; k < n_turns; k++) megno_vals[k] /= (k + 1);
                sum = 0.0;
                for (unsigned k = 0; k < n_turns; k++) { sum += megno_vals[k]; megno_vals[k] = sum; }
                for (size_t k = 0; k < n_samples; k++)
                    out[j * n_samples + k] = megno_vals[samples[k]] / samples[k];
            }
        });
    }
    for (auto &tt : th) tt.join();
    return 0;
}

// ---- Birkhoff: dynamic_indicator.cpp:63-131, 198-253 -------------------------

// dynamic_indicator.cpp:63-79
void or_birkhoff_weights(unsigned n_weights, double *w)
{
    for (unsigned i = 0; i < n_weights; i++) {
        double t = (double)i / (double)(n_weights);
        w[i] = std::exp(-1 / (t * (1.0 - t)));
    }
    double sum = 0.0;
    for (unsigned i = 0; i < n_weights; i++) sum = sum + w[i];  // std::accumulate order
    for (unsigned i = 0; i < n_weights; i++) w[i] = w[i] / sum;
}

// dynamic_indicator.cpp:81-131
double or_birkhoff_tune(const double *x, const double *px, size_t n)
{
    std::vector<double> theta(n);
    for (size_t i = 0; i < n; i++) theta[i] = std::arg(std::complex<double>(x[i], px[i]));
    std::vector<double> w(n - 1);
    or_birkhoff_weights((unsigned)(n - 1), w.data());
    double sum = 0.0;
    for (size_t i = 0; i + 1 < n; i++) {
        double d = theta[i + 1] - theta[i];
        if (d < 0.0) d += 2.0 * M_PI;
        sum += w[i] * d;
    }
    return 1 - sum / (2 * M_PI);
}

static bool any_nan(const double *a, size_t from, size_t to_incl)
{
    for (size_t i = from; i <= to_incl; i++) if (std::isnan(a[i])) return true;
    return false;
}

// dynamic_indicator.cpp:198-253. out has 2*(nw+1) entries: per window (x, y),
// then the full signal (x, y). Returns 1 on an invalid window.
int or_birkhoff_tune_vec(const double *x, const double *px, const double *y, const double *py, size_t len,
                         const unsigned *from, const unsigned *to, size_t nw, double *out)
{
    const double qnan = std::numeric_limits<double>::quiet_NaN();
    for (size_t i = 0; i < nw; i++) {
        if (from[i] >= len || to[i] >= len) return 1;
        size_t a = from[i], b = to[i];
        if (any_nan(x, a, b) || any_nan(px, a, b) || any_nan(y, a, b) || any_nan(py, a, b)) {
            out[2 * i] = qnan; out[2 * i + 1] = qnan;
        } else {
            out[2 * i] = or_birkhoff_tune(x + a, px + a, b - a + 1);
            out[2 * i + 1] = or_birkhoff_tune(y + a, py + a, b - a + 1);
        }
    }
    out[2 * nw] = or_birkhoff_tune(x, px, len);
    out[2 * nw + 1] = or_birkhoff_tune(y, py, len);
    return 0;
}

// ---- FFT tune: dynamic_indicator.cpp:32-61, 134-196, 255-361 -----------------
// The forward DFT (FFTW_FORWARD, unnormalised, e^{-2 pi i k n / N}) is restated
// as a direct sum with exactly reduced twiddle angles (k*n mod N) evaluated in
// long double, accurate to a few ulp of the spectrum's scale.

// dynamic_indicator.cpp:32-61 (compute_fft with the Hanning window) -> |X_k|
void or_fft_magnitude(const double *re, const double *im, size_t N, int hanning, double *mag)
{
    std::vector<double> inr(N), ini(N);
    for (size_t i = 0; i < N; i++) { inr[i] = re[i]; ini[i] = im[i]; }
    if (hanning) {
        for (size_t i = 0; i < N; i++) {
            double h = 2 * std::pow((std::sin(M_PI * i / (N))), 2);
            inr[i] *= h; ini[i] *= h;
        }
    }
    std::vector<long double> cs(N), sn(N);
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (size_t q = 0; q < N; q++) {
        long double ang = two_pi * (long double)q / (long double)N;
        cs[q] = cosl(ang); sn[q] = sinl(ang);
    }
    for (size_t k = 0; k < N; k++) {
        long double ar = 0.0L, ai = 0.0L;
        size_t q = 0;  // (k*n) mod N
        for (size_t m = 0; m < N; m++) {
            // (a + ib) * (cos - i sin)
            ar += (long double)inr[m] * cs[q] + (long double)ini[m] * sn[q];
            ai += (long double)ini[m] * cs[q] - (long double)inr[m] * sn[q];
            q += k; if (q >= N) q -= N;
        }
        double xr = (double)ar, xi = (double)ai;
        mag[k] = std::sqrt(xr * xr + xi * xi);
    }
}

// dynamic_indicator.cpp:134-137
double or_interp_A(double a, double b, double c)
{
    return (-(a + b * c) * (a - b) + b * std::sqrt(std::pow(c, 2) * std::pow((a + b), 2) - 2 * a * b * (2 * std::pow(c, 2) - c - 1))) /
           (std::pow(a, 2) + std::pow(b, 2) + 2 * a * b * c);
}

// dynamic_indicator.cpp:139-176 (first maximum wins, like std::max_element)
double or_interpolation(const double *data, size_t n)
{
    size_t index = 0;
    for (size_t i = 1; i < n; i++) if (data[i] > data[index]) index = i;
    double N = double(n);
    if (index == 0) return 1.0;
    if (index == n - 1) return 0.0;
    size_t i1, i2;
    if (data[index - 1] > data[index + 1]) { i1 = index - 1; i2 = index; index = index - 1; }
    else { i1 = index; i2 = index + 1; }
    double value = ((double(index) / N) +
                    (1.0 / (2.0 * M_PI)) * std::asin(or_interp_A(data[i1], data[i2], std::cos(2.0 * M_PI / N)) *
                                                     std::sin(2.0 * M_PI / N)));
    return std::abs(1.0 - value);
}

// dynamic_indicator.cpp:178-196
double or_fft_tune(const double *x, const double *px, size_t n)
{
    std::vector<double> mag(n);
    or_fft_magnitude(x, px, n, 1, mag.data());
    return or_interpolation(mag.data(), n);
}

// dynamic_indicator.cpp:255-331, with the NaN check over [from, to] (the
// reference's iterator bound there is out of range) and the full signal taken
// over samples 1..len-1.
int or_fft_tune_vec(const double *x, const double *px, const double *y, const double *py, size_t len,
                    const unsigned *from, const unsigned *to, size_t nw, double *out)
{
    const double qnan = std::numeric_limits<double>::quiet_NaN();
    for (size_t i = 0; i < nw; i++) {
        if (from[i] >= len || to[i] >= len) return 1;
        size_t a = from[i], b = to[i];
        if (any_nan(x, a, b) || any_nan(px, a, b) || any_nan(y, a, b) || any_nan(py, a, b)) {
            out[2 * i] = qnan; out[2 * i + 1] = qnan;
            continue;
        }
        out[2 * i] = or_fft_tune(x + a, px + a, b - a + 1);
        out[2 * i + 1] = or_fft_tune(y + a, py + a, b - a + 1);
    }
    if (any_nan(x, 0, len - 1) || any_nan(px, 0, len - 1) || any_nan(y, 0, len - 1) || any_nan(py, 0, len - 1)) {
        out[2 * nw] = qnan; out[2 * nw + 1] = qnan;
    } else {
        out[2 * nw] = or_fft_tune(x + 1, px + 1, len - 1);
        out[2 * nw + 1] = or_fft_tune(y + 1, py + 1, len - 1);
    }
    return 0;
}

// dynamic_indicator.cpp:333-361: returns {fft, birkhoff}
void or_get_tunes(const double *x, const double *px, size_t n, double *out)
{
    if (any_nan(x, 0, n - 1) || any_nan(px, 0, n - 1)) {
        out[0] = out[1] = std::numeric_limits<double>::quiet_NaN();
        return;
    }
    out[1] = or_birkhoff_tune(x, px, n);
    out[0] = or_fft_tune(x, px, n);
}

// ---- trackers with history: henon.cu:2890-3078 -------------------------------

// henon.cu:2890-2967. out: n x 2*(nw+2) (windows, then the appended (0,n_turns)
// window, then the full signal).
int or_birkhoff_tunes(double *x, double *px, double *y, double *py, unsigned *steps, size_t n,
                      const double *sx, const double *cx, const double *sy, const double *cy,
                      unsigned n_turns, double mu, double barrier, int inverse,
                      const unsigned *from, const unsigned *to, size_t nw, int n_threads, double *out)
{
    std::vector<unsigned> f(from, from + nw), t(to, to + nw);
    f.push_back(0); t.push_back(n_turns);
    const size_t W = f.size();
    const double b2 = barrier * barrier;
    int bad = 0;
    const unsigned nt = resolve_threads(n_threads);
    std::vector<std::thread> th;
    for (unsigned tt = 0; tt < nt; tt++) {
        th.emplace_back([&, tt]() {
            std::vector<double> hx(n_turns + 1), hpx(n_turns + 1), hy(n_turns + 1), hpy(n_turns + 1);
            for (size_t j = tt; j < n; j += nt) {
                hx[0] = x[j]; hpx[0] = px[j]; hy[0] = y[j]; hpy[0] = py[j];
                for (unsigned k = 0; k < n_turns; k++) {
                    henon_step(x[j], px[j], y[j], py[j], steps[j], sx, cx, sy, cy, b2, mu, inverse != 0);
                    hx[k + 1] = x[j]; hpx[k + 1] = px[j]; hy[k + 1] = y[j]; hpy[k + 1] = py[j];
                }
                if (or_birkhoff_tune_vec(hx.data(), hpx.data(), hy.data(), hpy.data(), n_turns + 1, f.data(),
                                         t.data(), W, out + j * 2 * (W + 1)))
                    bad = 1;
            }
        });
    }
    for (auto &q : th) q.join();
    return bad;
}

// henon.cu:2969-3078: history, per-coordinate mean removal (std::accumulate
// order), Birkhoff and FFT tunes. out_birk, out_fft: n x 2*(nw+2).
int or_all_tunes(double *x, double *px, double *y, double *py, unsigned *steps, size_t n,
                 const double *sx, const double *cx, const double *sy, const double *cy,
                 unsigned n_turns, double mu, double barrier, int inverse,
                 const unsigned *from, const unsigned *to, size_t nw, int n_threads,
                 double *out_birk, double *out_fft)
{
    std::vector<unsigned> f(from, from + nw), t(to, to + nw);
    f.push_back(0); t.push_back(n_turns);
    const size_t W = f.size();
    const double b2 = barrier * barrier;
    int bad = 0;
    const unsigned nt = resolve_threads(n_threads);
    std::vector<std::thread> th;
    for (unsigned tt = 0; tt < nt; tt++) {
        th.emplace_back([&, tt]() {
            const size_t L = n_turns + 1;
            std::vector<double> h[4];
            for (auto &v : h) v.resize(L);
            for (size_t j = tt; j < n; j += nt) {
                h[0][0] = x[j]; h[1][0] = px[j]; h[2][0] = y[j]; h[3][0] = py[j];
                for (unsigned k = 0; k < n_turns; k++) {
                    henon_step(x[j], px[j], y[j], py[j], steps[j], sx, cx, sy, cy, b2, mu, inverse != 0);
                    h[0][k + 1] = x[j]; h[1][k + 1] = px[j]; h[2][k + 1] = y[j]; h[3][k + 1] = py[j];
                }
                for (auto &v : h) {
                    double s = 0.0;
                    for (size_t q = 0; q < L; q++) s = s + v[q];
                    double mean = s / L;
                    for (size_t q = 0; q < L; q++) v[q] = v[q] - mean;
                }
                int e1 = or_birkhoff_tune_vec(h[0].data(), h[1].data(), h[2].data(), h[3].data(), L, f.data(),
                                              t.data(), W, out_birk + j * 2 * (W + 1));
                int e2 = or_fft_tune_vec(h[0].data(), h[1].data(), h[2].data(), h[3].data(), L, f.data(),
                                         t.data(), W, out_fft + j * 2 * (W + 1));
                if (e1 || e2) bad = 1;
            }
        });
    }
    for (auto &q : th) q.join();
    return bad;
}

// ---- per-call construct formulas (GPU constructs of the reference) -----------

// henon.cu:2083-2098 (kernel_lyap_birk)
void or_lyap_birk_add(double *values, double *values_b, const double *vectors, const double *weights, size_t n,
                      size_t n_weights, size_t b_idx)
{
    for (size_t i = 0; i < n; i++) {
        double norm = 0.0;
        for (int k = 0; k < 4; k++) norm += vectors[i * 4 + k] * vectors[i * 4 + k];
        norm = std::sqrt(norm);
        values[i] += std::log(norm) / n_weights;
        values_b[i] += std::log(norm) * weights[b_idx];
    }
}

// henon.cu:2253-2284 (kernel_megno)
void or_megno_construct_add(double *l1, double *l2, double *l3, const double *a, const double *b, size_t n,
                            size_t idx)
{
    for (size_t i = 0; i < n; i++) {
        double ta = 0.0, tb = 0.0;
        for (int k = 0; k < 16; k++) ta += a[i * 16 + k] * a[i * 16 + k];
        ta = std::sqrt(ta);
        for (int k = 0; k < 16; k++) tb += b[i * 16 + k] * b[i * 16 + k];
        tb = std::sqrt(tb);
        l1[i] += std::log10(ta / tb) * idx;
        l2[i] += (1.0 / idx) * l1[i];
        l3[i] = (1.0 / idx) * l2[i];
    }
}

// henon.cu:2366-2402 (kernel_megno_birkhoff); weights: per sample j a pointer
// offset into a flat array (w_off[j]) of birkhoff_weights(turn_samples[j]).
void or_megno_birkhoff_add(double *l1, double *l2 /* ns x n */, const double *w_flat, const size_t *w_off,
                           const size_t *turn_samples, size_t ns, const double *a, const double *b, size_t n,
                           size_t idx)
{
    for (size_t i = 0; i < n; i++) {
        double ta = 0.0, tb = 0.0;
        for (int k = 0; k < 16; k++) ta += a[i * 16 + k] * a[i * 16 + k];
        ta = std::sqrt(ta);
        for (int k = 0; k < 16; k++) tb += b[i * 16 + k] * b[i * 16 + k];
        tb = std::sqrt(tb);
        l1[i] += std::log10(ta / tb) * idx;
        for (size_t j = 0; j < ns; j++)
            if (idx <= turn_samples[j]) l2[j * n + i] += l1[i] * w_flat[w_off[j] + idx - 1] / idx;
    }
}

// henon.cu:2533-2584 (kernel_tune_birkhoff_init / kernel_tune_birkhoff)
void or_tune_birkhoff_init(const double *x, const double *px, const double *y, const double *py, double *sx_,
                           double *sy_, size_t n)
{
    for (size_t i = 0; i < n; i++) { sx_[i] = std::atan2(px[i], x[i]); sy_[i] = std::atan2(py[i], y[i]); }
}

void or_tune_birkhoff_add(const double *x, const double *px, const double *y, const double *py, double *store_x,
                          double *store_y, double *t1x, double *t2x, double *t1y, double *t2y /* ns x n */,
                          const double *w_flat, const size_t *w_off, const size_t *turn_samples, size_t ns,
                          size_t n, size_t idx)
{
    for (size_t i = 0; i < n; i++) {
        double ax = std::atan2(px[i], x[i]);
        double ay = std::atan2(py[i], y[i]);
        double dx = ax - store_x[i];
        double dy = ay - store_y[i];
        dx = dx > 0 ? dx : dx + 2 * M_PI;
        dy = dy > 0 ? dy : dy + 2 * M_PI;
        store_x[i] = ax; store_y[i] = ay;
        for (size_t j = 0; j < ns; j++) {
            if (idx < turn_samples[j]) {
                t1x[j * n + i] += dx * w_flat[w_off[j] + idx];
                t1y[j * n + i] += dy * w_flat[w_off[j] + idx];
            } else if (idx < 2 * turn_samples[j]) {
                t2x[j * n + i] += dx * w_flat[w_off[j] + idx - turn_samples[j]];
                t2y[j * n + i] += dy * w_flat[w_off[j] + idx - turn_samples[j]];
            }
        }
    }
}

// henon.cu:316-338 (matrix_vector_multiply) and :341-359 (normalize_vectors)
void or_matvec(const double *m, double *v, size_t n)
{
    for (size_t i = 0; i < n; i++) {
        double t[4];
        for (int r = 0; r < 4; r++) {
            t[r] = 0.0;
            for (int c = 0; c < 4; c++) t[r] += m[i * 16 + r * 4 + c] * v[i * 4 + c];
        }
        for (int r = 0; r < 4; r++) v[i * 4 + r] = t[r];
    }
}

void or_normalize(double *v, size_t n)
{
    for (size_t i = 0; i < n; i++) {
        double norm = 0.0;
        for (int r = 0; r < 4; r++) norm += v[i * 4 + r] * v[i * 4 + r];
        norm = std::sqrt(norm);
        for (int r = 0; r < 4; r++) v[i * 4 + r] /= norm;
    }
}

// henon.cu:28-65 (nan_mean / nan_std)
double or_nan_mean(const double *v, size_t n)
{
    double sum = 0.0; int count = 0;
    for (size_t i = 0; i < n; i++) if (!std::isnan(v[i])) { sum += v[i]; count++; }
    if (count == 0) return std::numeric_limits<double>::quiet_NaN();
    return sum / count;
}

double or_nan_std(const double *v, size_t n)
{
    double mean = or_nan_mean(v, n);
    double sum = 0.0; int count = 0;
    for (size_t i = 0; i < n; i++) if (!std::isnan(v[i])) { sum += (v[i] - mean) * (v[i] - mean); count++; }
    if (count == 0) return std::numeric_limits<double>::quiet_NaN();
    return std::sqrt(sum / count);
}

// ---- shadow particles: henon.cu:362-418, 737-952 --------------------------------

static double displacement(double x1, double px1, double y1, double py1, double x2, double px2, double y2, double py2)
{
    // henon.cu:407-410
    return std::sqrt((x1 - x2) * (x1 - x2) + (px1 - px2) * (px1 - px2) + (y1 - y2) * (y1 - y2) +
                     (py1 - py2) * (py1 - py2));
}

static void realign(double &x, double &px, double &y, double &py, double xc, double pxc, double yc, double pyc,
                    double initial, double final_)
{
    // henon.cu:412-418
    x = xc + (x - xc) * (final_ / initial);
    px = pxc + (px - pxc) * (final_ / initial);
    y = yc + (y - yc) * (final_ / initial);
    py = pyc + (py - pyc) * (final_ / initial);
}

// henon.cu:362-384 (random_4d_kick) on the particle generator of henon.h:63 (a
// default-seeded std::mt19937_64 that add_ghost keeps using): the generator is
// replayed past the `skip` kicks drawn before, then kicks the n points in order.
void or_random_kicks(uint64_t skip, size_t n, double module, double *x, double *px, double *y, double *py)
{
    std::mt19937_64 rng;
    auto kick = [&](double &a, double &b, double &c, double &d) {
        std::normal_distribution<double> nd(0.0, 1.0);
        std::array<double, 4> k;
        for (auto &v : k) v = nd(rng);
        const double mod = std::sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2] + k[3] * k[3]);
        a += k[0] * module / mod;
        b += k[1] * module / mod;
        c += k[2] * module / mod;
        d += k[3] * module / mod;
    };
    double dummy[4] = {0.0, 0.0, 0.0, 0.0};
    for (uint64_t q = 0; q < skip; q++) kick(dummy[0], dummy[1], dummy[2], dummy[3]);
    for (size_t i = 0; i < n; i++) kick(x[i], px[i], y[i], py[i]);
}

// henon.cu:826-858: ghost i = n_base * (1 + g) + b realigned to distance target
// from base b.
void or_ghost_renormalize(double *x, double *px, double *y, double *py, size_t n_base, size_t sets, double target)
{
    for (size_t t = 0; t < n_base * sets; t++) {
        const size_t b = t % n_base, i = n_base + t;
        const double module = displacement(x[i], px[i], y[i], py[i], x[b], px[b], y[b], py[b]);
        realign(x[i], px[i], y[i], py[i], x[b], px[b], y[b], py[b], module, target);
    }
}

// henon.cu:860-952 (CPU accessors): module [n_base][sets], direction
// [4][n_base][sets] = the ghost realigned to distance 1, minus the base.
void or_ghost_displacement(const double *x, const double *px, const double *y, const double *py, size_t n_base,
                           size_t sets, double *module_out, double *dir_out)
{
    const size_t plane = n_base * sets;
    for (size_t t = 0; t < plane; t++) {
        const size_t g = t / n_base, b = t % n_base, i = n_base * (1 + g) + b, o = b * sets + g;
        const double module = displacement(x[i], px[i], y[i], py[i], x[b], px[b], y[b], py[b]);
        if (module_out) module_out[o] = module;
        if (dir_out) {
            double d[4] = {x[i], px[i], y[i], py[i]};
            realign(d[0], d[1], d[2], d[3], x[b], px[b], y[b], py[b], module, 1.0);
            dir_out[o] = d[0] - x[b];
            dir_out[plane + o] = d[1] - px[b];
            dir_out[2 * plane + o] = d[2] - y[b];
            dir_out[3 * plane + o] = d[3] - py[b];
        }
    }
}

// ---- random kick: henon.cu:362-405, 420-480 ---------------------------------------
// The reference draws the normals from std::mt19937_64 (CPU) or clock-seeded
// cuRAND (GPU); the product draws them from Philox-4x32-10 (Salmon et al.,
// SC'11) keyed by a seed with counter (particle, draw), Box-Muller normals. The
// generator is restated here from the paper and pinned by its published
// known-answer vectors; the kicked step is henon_step's.

void or_philox4x32_10(uint32_t *c, uint32_t k0, uint32_t k1)
{
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0, n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1, n3 = (uint32_t)p0;
        c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

static double uniform53(uint32_t a, uint32_t b)
{
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

void or_kick_normals(uint64_t seed, uint64_t i, uint64_t d, double *n)
{
    for (uint32_t half = 0; half < 2; half++) {
        uint32_t c[4] = {(uint32_t)i, (uint32_t)(i >> 32), (uint32_t)d, ((uint32_t)(d >> 32) << 1) | half};
        or_philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
        const double u1 = 1.0 - uniform53(c[0], c[1]), u2 = uniform53(c[2], c[3]);
        const double rad = std::sqrt(-2.0 * std::log(u1));
        n[2 * half] = rad * std::cos(2.0 * M_PI * u2);
        n[2 * half + 1] = rad * std::sin(2.0 * M_PI * u2);
    }
}

// henon.cu:420-480 (henon_step) with the kick above; per particle, particle
// index i, draw index draw + k at the call's k-th turn.
int or_track_kicked(double *x, double *px, double *y, double *py, unsigned *steps, size_t n, const double *sx,
                    const double *cx, const double *sy, const double *cy, size_t tab_len, unsigned n_turns, double mu,
                    double barrier, double module, int reverse, uint64_t seed, uint64_t draw)
{
    const double b2 = barrier * barrier;
    int bad = 0;
    for (size_t i = 0; i < n; i++) {
        for (unsigned k = 0; k < n_turns; k++) {
            if (check_nan(x[i], px[i], y[i], py[i])) break;
            const unsigned ti = reverse ? steps[i] - 1 : steps[i];
            if ((reverse && steps[i] == 0) || ti >= tab_len) { bad = 1; break; }
            double kk[4];
            if (!reverse) {
                or_kick_normals(seed, i, draw + k, kk);
                const double mod = std::sqrt(kk[0] * kk[0] + kk[1] * kk[1] + kk[2] * kk[2] + kk[3] * kk[3]);
                x[i] += kk[0] * module / mod; px[i] += kk[1] * module / mod;
                y[i] += kk[2] * module / mod; py[i] += kk[3] * module / mod;
                sextupole(x[i], px[i], y[i], py[i], false);
                if (mu != 0.0) octupole(x[i], px[i], y[i], py[i], mu, false);
                rotation(x[i], px[i], sx[ti], cx[ti], false);
                rotation(y[i], py[i], sy[ti], cy[ti], false);
            } else {
                rotation(x[i], px[i], sx[ti], cx[ti], true);
                rotation(y[i], py[i], sy[ti], cy[ti], true);
                if (mu != 0.0) octupole(x[i], px[i], y[i], py[i], mu, true);
                sextupole(x[i], px[i], y[i], py[i], true);
                or_kick_normals(seed, i, draw + k, kk);
                const double mod = std::sqrt(kk[0] * kk[0] + kk[1] * kk[1] + kk[2] * kk[2] + kk[3] * kk[3]);
                x[i] += kk[0] * module / mod; px[i] += kk[1] * module / mod;
                y[i] += kk[2] * module / mod; py[i] += kk[3] * module / mod;
            }
            if (check_barrier(x[i], px[i], y[i], py[i], b2)) {
                steps[i] += reverse ? -1 : 1;
            } else {
                x[i] = px[i] = y[i] = py[i] = std::numeric_limits<double>::quiet_NaN();
                break;
            }
        }
    }
    return bad;
}

}  // extern "C"

// ---- fused indicators (new algorithm: no reference counterpart) ---------------
// CPU restatement of the fused kernel's exact mode. Per turn k (table index
// steps+k): J at the pre-step point, M <- J M with J's zero entries skipped in
// a fixed order, map step and loss test, then the construct-form MEGNO of
// henon.cu:2253-2284 (L1 += log10(|M_k+1|_F / |M_k|_F) idx, L2 += L1 / idx),
// FLI = (log10 scale + log10 |M v0|) / t, scalar renormalisation every R turns,
// and online Birkhoff sums as dynamic_indicator.cpp:81-131.
// Pinned by: the reference construct formulas applied turn by turn to
// reference-order products (tests/test_oracle_indicators.py), closed forms on
// rotations, and renormalisation invariance.
static void mat_mul_structured(const double *j, const double *t, double *m)
{
    // j: a, b, an ; t: sx, cx, sy, cy
    const double sx = t[0], cx = t[1], sy = t[2], cy = t[3];
    const double j00 = cx + (sx * j[0]), j02 = sx * j[1];
    const double j10 = -sx + (cx * j[0]), j12 = cx * j[1];
    const double j20 = sy * j[1], j22 = cy + (sy * j[2]);
    const double j30 = cy * j[1], j32 = -sy + (cy * j[2]);
    double r[16];
    for (int c = 0; c < 4; c++) {
        const double m0 = m[c], m1 = m[4 + c], m2 = m[8 + c], m3 = m[12 + c];
        r[c] = ((j00 * m0) + (sx * m1)) + (j02 * m2);
        r[4 + c] = ((j10 * m0) + (cx * m1)) + (j12 * m2);
        r[8 + c] = ((j20 * m0) + (j22 * m2)) + (sy * m3);
        r[12 + c] = ((j30 * m0) + (j32 * m2)) + (cy * m3);
    }
    std::memcpy(m, r, sizeof(r));
}

static double frob2(const double *m)
{
    double s = m[0] * m[0];
    for (int q = 1; q < 16; q++) s = s + (m[q] * m[q]);
    return s;
}

extern "C" int or_indicators(double *x, double *px, double *y, double *py, unsigned *steps, size_t n,
                             const double *sx, const double *cx, const double *sy, const double *cy, size_t tab_len,
                             unsigned n_turns, double mu, double barrier, unsigned renorm_every,
                             const unsigned *samples, size_t n_samples, const double *fli_dir, double *megno,
                             double *fli, double *matrix, double *log_scale, const unsigned *birk_from,
                             const unsigned *birk_to, size_t n_birk, double *birk_out, int n_threads)
{
    const double b2 = barrier * barrier;
    const double mu3 = mu * 3, mu6 = mu * 6;
    std::vector<std::vector<double>> w(n_birk);
    for (size_t q = 0; q < n_birk; q++) {
        w[q].resize(birk_to[q] - birk_from[q]);
        or_birkhoff_weights(birk_to[q] - birk_from[q], w[q].data());
    }
    double v0[4];
    {
        const double nn = std::sqrt(((fli_dir[0] * fli_dir[0] + fli_dir[1] * fli_dir[1]) + fli_dir[2] * fli_dir[2]) +
                                    fli_dir[3] * fli_dir[3]);
        for (int q = 0; q < 4; q++) v0[q] = fli_dir[q] / nn;
    }
    int bad = 0;
    strided_parallel(n, resolve_threads(n_threads), [&](size_t i, unsigned) {
        double m[16];
        if (matrix) std::memcpy(m, matrix + i * 16, sizeof(m));
        else for (int q = 0; q < 16; q++) m[q] = (q % 5 == 0) ? 1.0 : 0.0;
        double lscale = log_scale ? log_scale[i] : 0.0;
        double fb = std::sqrt(frob2(m));
        double l1 = 0.0, l2 = 0.0;
        bool alive = !check_nan(x[i], px[i], y[i], py[i]);
        // a window holding a sample with any coordinate NaN is NaN in both
        // planes (dynamic_indicator.cpp:284-287); a particle lost inside it
        // is NaN in all four, one NaN from the start (not tracked) in some
        const bool start_nan = !alive;
        std::vector<double> bacc(2 * n_birk, 0.0);
        double thx = std::atan2(px[i], x[i]), thy = std::atan2(py[i], y[i]);
        size_t sq = 0;
        unsigned s = steps[i];
        for (unsigned k = 0; k < n_turns; k++) {
            const size_t ti = (size_t)s + k;
            if (ti >= tab_len) { bad = 1; break; }
            const double t[4] = {sx[ti], cx[ti], sy[ti], cy[ti]};
            const double xv = x[i], yv = y[i];
            const double jc[3] = {((2.0 * xv) + ((mu3 * xv) * xv)) - ((mu3 * yv) * yv),
                                  (-2.0 * yv) - ((mu6 * xv) * yv),
                                  ((-2.0 * xv) - ((mu3 * xv) * xv)) + ((mu3 * yv) * yv)};
            mat_mul_structured(jc, t, m);
            if (alive) {
                sextupole(x[i], px[i], y[i], py[i], false);
                if (mu != 0.0) octupole(x[i], px[i], y[i], py[i], mu, false);
                rotation(x[i], px[i], t[0], t[1], false);
                rotation(y[i], py[i], t[2], t[3], false);
                if (!check_barrier(x[i], px[i], y[i], py[i], b2)) {
                    const double qn = std::numeric_limits<double>::quiet_NaN();
                    x[i] = px[i] = y[i] = py[i] = qn;
                    for (int q = 0; q < 16; q++) m[q] = qn;
                    alive = false;
                    steps[i] = s + k;
                }
            }
            const unsigned idx = k + 1;
            const double fa = std::sqrt(frob2(m));
            l1 = l1 + std::log10(fa / fb) * (double)idx;
            l2 = l2 + (1.0 / (double)idx) * l1;
            fb = fa;
            if (sq < n_samples && samples[sq] == idx) {
                if (megno) megno[sq * n + i] = (1.0 / (double)idx) * l2;
                if (fli) {
                    double v[4];
                    for (int r = 0; r < 4; r++)
                        v[r] = (((m[r * 4] * v0[0]) + (m[r * 4 + 1] * v0[1])) + (m[r * 4 + 2] * v0[2])) +
                               (m[r * 4 + 3] * v0[3]);
                    const double nv = std::sqrt((((v[0] * v[0]) + (v[1] * v[1])) + (v[2] * v[2])) + (v[3] * v[3]));
                    fli[sq * n + i] = (lscale + std::log10(nv)) / (double)idx;
                }
                ++sq;
            }
            if (renorm_every && idx % renorm_every == 0 && alive) {
                const double inv = 1.0 / fa;
                for (int q = 0; q < 16; q++) m[q] = m[q] * inv;
                lscale = lscale + std::log10(fa);
                fb = std::sqrt(frob2(m));
            }
            if (n_birk) {
                const double nx = std::atan2(px[i], x[i]), ny = std::atan2(py[i], y[i]);
                double dx = nx - thx, dy = ny - thy;
                if (dx < 0.0) dx += 2.0 * M_PI;
                if (dy < 0.0) dy += 2.0 * M_PI;
                thx = nx; thy = ny;
                for (size_t q = 0; q < n_birk; q++) {
                    if (idx > birk_from[q] && idx <= birk_to[q]) {
                        const double wt = w[q][idx - 1 - birk_from[q]];
                        bacc[2 * q] = bacc[2 * q] + wt * dx;
                        bacc[2 * q + 1] = bacc[2 * q + 1] + wt * dy;
                    }
                }
            }
        }
        if (alive) steps[i] = s + n_turns;
        for (; sq < n_samples; ++sq) {
            if (megno) megno[sq * n + i] = std::numeric_limits<double>::quiet_NaN();
            if (fli) fli[sq * n + i] = std::numeric_limits<double>::quiet_NaN();
        }
        if (matrix) std::memcpy(matrix + i * 16, m, sizeof(m));
        if (log_scale) log_scale[i] = lscale;
        for (size_t q = 0; q < n_birk; q++) {
            const double qn = std::numeric_limits<double>::quiet_NaN();
            birk_out[(2 * q) * n + i] = start_nan ? qn : 1.0 - bacc[2 * q] / (2 * M_PI);
            birk_out[(2 * q + 1) * n + i] = start_nan ? qn : 1.0 - bacc[2 * q + 1] / (2 * M_PI);
        }
    });
    return bad;
}
