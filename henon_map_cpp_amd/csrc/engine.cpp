// henon_map_engine: the pybind11 module that replaces the reference's
// src/bind.cpp module of the same name. Same classes, method names, argument
// names and defaults; every computation goes through the C ABI of
// libhenon_hip.so (include/henon_hip.h), so every class is MI355X-backed:
// `particles_4d`/`henon_tracker` (the reference's CPU classes) and their `_gpu`
// subclasses share one implementation.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <unordered_set>
#include <thread>
#include <vector>

#include "henon_hip.h"

namespace py = pybind11;

namespace {

// ---------------------------------------------------------------- C ABI glue
struct NotImplemented : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void check(int rc)
{
    if (rc == HM_OK) return;
    std::string msg = hm_last_error();
    switch (rc) {
    case HM_ERR_ARG: throw std::invalid_argument(msg);
    case HM_ERR_UNSUPPORTED: throw NotImplemented(msg);
    case HM_ERR_ALLOC: throw std::bad_alloc();
    default: throw std::runtime_error(msg);
    }
}

void sync() { check(hm_stream_synchronize(nullptr)); }

void require_device()
{
    int c = 0;
    check(hm_device_count(&c));
    if (c < 1) throw std::runtime_error("henon_map_engine: no HIP device is visible (this engine runs on MI355X)");
}

// A device allocation owned through the C ABI; move-only (the reference's
// classes owned raw device pointers with implicit copies).
class DevBuf {
  public:
    DevBuf() = default;
    explicit DevBuf(size_t bytes) { alloc(bytes); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p_(o.p_), bytes_(o.bytes_) { o.p_ = nullptr; o.bytes_ = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept
    {
        if (this != &o) {
            release();
            p_ = o.p_; bytes_ = o.bytes_;
            o.p_ = nullptr; o.bytes_ = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }

    void alloc(size_t bytes)
    {
        release();
        check(hm_malloc(&p_, bytes ? bytes : 8));
        bytes_ = bytes;
    }
    template <class T> void upload(const std::vector<T> &v)
    {
        if (v.size() * sizeof(T) > bytes_) alloc(v.size() * sizeof(T));
        if (!v.empty()) check(hm_memcpy(p_, v.data(), v.size() * sizeof(T), 0, nullptr));
        sync();
    }
    template <class T> std::vector<T> download(size_t n) const
    {
        std::vector<T> v(n);
        if (n) check(hm_memcpy(v.data(), p_, n * sizeof(T), 1, nullptr));
        sync();
        return v;
    }
    template <class T> T *as() const { return static_cast<T *>(p_); }
    size_t bytes() const { return bytes_; }

  private:
    void release()
    {
        if (p_) (void)hm_free(p_);
        p_ = nullptr;
        bytes_ = 0;
    }
    void *p_ = nullptr;
    size_t bytes_ = 0;
};

double nan_mean(const std::vector<double> &v)
{
    // src/henon.cu:28-46
    double sum = 0.0;
    int count = 0;
    for (double a : v)
        if (!std::isnan(a)) { sum += a; count++; }
    if (count == 0) return std::numeric_limits<double>::quiet_NaN();
    return sum / count;
}

double nan_std(const std::vector<double> &v)
{
    // src/henon.cu:48-65
    double mean = nan_mean(v);
    double sum = 0.0;
    int count = 0;
    for (double a : v)
        if (!std::isnan(a)) { sum += (a - mean) * (a - mean); count++; }
    if (count == 0) return std::numeric_limits<double>::quiet_NaN();
    return std::sqrt(sum / count);
}

// ---------------------------------------------------------------- particles
uint64_t fresh_seed() { return ((uint64_t)std::random_device{}() << 32) ^ std::random_device{}(); }

// henon.cu:362-384: a kick of length module along a direction drawn from four
// standard normals (a fresh distribution per kick, as there); the reference's
// particle generator is a default-seeded std::mt19937_64, so the kicks repeat
// run to run.
void random_4d_kick(double &x, double &px, double &y, double &py, double module, std::mt19937_64 &rng)
{
    std::normal_distribution<double> nd(0.0, 1.0);
    double k[4];
    for (double &c : k) c = nd(rng);
    const double mod = std::sqrt(((k[0] * k[0] + k[1] * k[1]) + k[2] * k[2]) + k[3] * k[3]);
    x += k[0] * module / mod;
    px += k[1] * module / mod;
    y += k[2] * module / mod;
    py += k[3] * module / mod;
}

// particles_4d (src/henon.h:41-118, src/henon.cu:699-1151) backed by device SoA.
// With shadow particles the arrays hold the n_base base particles, then one set
// of n_base ghosts per add_ghost call (src/henon.cu:737-817); every accessor
// that returns per-particle arrays returns all entries, as the reference's GPU
// class does (src/henon.cu:1444-1690), and tracking advances all of them.
// One coordinate array as numpy hands it over (converted to contiguous float64
// if it is anything else, lists included): read with one copy, not element by
// element through the sequence protocol.
using Coords = py::array_t<double, py::array::c_style | py::array::forcecast>;
std::vector<double> coords_vector(const Coords &a)
{
    if (a.ndim() != 1) throw py::type_error("particles_4d: x, px, y, py must be one-dimensional");
    std::vector<double> v((size_t)a.shape(0));
    if (!v.empty()) std::memcpy(v.data(), a.data(), v.size() * sizeof(double));
    return v;
}

class Particles4D {
  public:
    Particles4D(const Coords &x, const Coords &px, const Coords &y, const Coords &py)
        : Particles4D(coords_vector(x), coords_vector(px), coords_vector(y), coords_vector(py))
    {
    }
    Particles4D(const std::vector<double> &x, const std::vector<double> &px, const std::vector<double> &y,
                const std::vector<double> &py)
    {
        if (px.size() != x.size() || y.size() != x.size() || py.size() != x.size())
            throw std::invalid_argument("particles_4d: x, px, y, py must have the same length");
        require_device();
        n_base = n_particles = x.size();
        x0 = x; px0 = px; y0 = y; py0 = py;
        upload_initial();
    }
    virtual ~Particles4D() = default;

    virtual void reset()
    {
        // src/henon.cu:727-735: global_steps is not reset; ghosts return to
        // the positions they were created at
        upload_initial();
    }
    virtual void add_ghost(const double &module, const std::string &direction)
    {
        // src/henon.cu:737-817: one ghost per base particle, displaced from the
        // particle's current state, with its current steps and valid flag.
        // (The reference's GPU class builds them from its host copy, which is
        // the initial state, and re-uploads every particle; here the device
        // state of the existing particles is kept.)
        if (direction != "x" && direction != "y" && direction != "px" && direction != "py" && direction != "random")
            throw std::runtime_error("Invalid direction.");
        const size_t nb = n_base, total = n_particles;
        std::vector<double> X = get_x(), PX = get_px(), Y = get_y(), PY = get_py();
        std::vector<unsigned int> S = get_steps();
        std::vector<uint8_t> V = get_valid();
        X.resize(total + nb); PX.resize(total + nb); Y.resize(total + nb); PY.resize(total + nb);
        S.resize(total + nb); V.resize(total + nb);
        x0.resize(total + nb); px0.resize(total + nb); y0.resize(total + nb); py0.resize(total + nb);
        for (size_t b = 0; b < nb; b++) {
            double gx = X[b], gpx = PX[b], gy = Y[b], gpy = PY[b];
            if (direction == "x") gx = gx + module;
            else if (direction == "y") gy = gy + module;
            else if (direction == "px") gpx = gpx + module;
            else if (direction == "py") gpy = gpy + module;
            else random_4d_kick(gx, gpx, gy, gpy, module, rng);
            const size_t g = total + b;
            X[g] = x0[g] = gx; PX[g] = px0[g] = gpx; Y[g] = y0[g] = gy; PY[g] = py0[g] = gpy;
            S[g] = S[b];
            V[g] = V[b];
        }
        dx.upload(X); dpx.upload(PX); dy.upload(Y); dpy.upload(PY);
        dsteps.upload(S); dvalid.upload(V);
        n_ghosts += 1;
        n_particles = total + nb;
    }
    virtual void renormalize(const double &module_target)
    {
        // src/henon.cu:1257-1282
        check(hm_ghost_renormalize(dx.as<double>(), dpx.as<double>(), dy.as<double>(), dpy.as<double>(), n_base,
                                   n_ghosts, module_target, nullptr));
        sync();
    }
    virtual std::vector<std::vector<double>> get_displacement_module() const
    {
        // src/henon.cu:1315-1335: [n_particles][n_ghosts_per_particle]
        std::vector<std::vector<double>> out(n_base, std::vector<double>(n_ghosts));
        if (n_ghosts == 0 || n_base == 0) return out;
        DevBuf d(n_base * n_ghosts * sizeof(double));
        check(hm_ghost_displacement(dx.as<double>(), dpx.as<double>(), dy.as<double>(), dpy.as<double>(), n_base,
                                    n_ghosts, d.as<double>(), nullptr, nullptr));
        const std::vector<double> h = d.download<double>(n_base * n_ghosts);
        for (size_t b = 0; b < n_base; b++)
            std::copy(h.begin() + b * n_ghosts, h.begin() + (b + 1) * n_ghosts, out[b].begin());
        return out;
    }
    virtual std::vector<std::vector<std::vector<double>>> get_displacement_direction() const
    {
        // src/henon.cu:1391-1442: [4][n_particles][n_ghosts_per_particle]
        std::vector<std::vector<std::vector<double>>> out(
            4, std::vector<std::vector<double>>(n_base, std::vector<double>(n_ghosts)));
        if (n_ghosts == 0 || n_base == 0) return out;
        const size_t plane = n_base * n_ghosts;
        DevBuf d(4 * plane * sizeof(double));
        check(hm_ghost_displacement(dx.as<double>(), dpx.as<double>(), dy.as<double>(), dpy.as<double>(), n_base,
                                    n_ghosts, nullptr, d.as<double>(), nullptr));
        const std::vector<double> h = d.download<double>(4 * plane);
        for (size_t c = 0; c < 4; c++)
            for (size_t b = 0; b < n_base; b++)
                std::copy(h.begin() + c * plane + b * n_ghosts, h.begin() + c * plane + (b + 1) * n_ghosts,
                          out[c][b].begin());
        return out;
    }

    // Alignment indices of the first ghost sets at the current state
    // (henon_map_cpp/dynamic_indicators.py:154-366): no reference method.
    std::vector<double> alignment_index(int kind) const
    {
        DevBuf d(n_base * sizeof(double));
        check(hm_ghost_alignment(dx.as<double>(), dpx.as<double>(), dy.as<double>(), dpy.as<double>(), n_base,
                                 n_ghosts, kind, d.as<double>(), nullptr));
        return d.download<double>(n_base);
    }

    virtual std::vector<double> get_x() const { return dx.download<double>(n_particles); }
    virtual std::vector<double> get_px() const { return dpx.download<double>(n_particles); }
    virtual std::vector<double> get_y() const { return dy.download<double>(n_particles); }
    virtual std::vector<double> get_py() const { return dpy.download<double>(n_particles); }
    virtual std::vector<unsigned int> get_steps() const { return dsteps.download<unsigned int>(n_particles); }

    std::vector<double> observable(int kind) const
    {
        DevBuf out(n_particles * sizeof(double));
        check(hm_observable(kind, dx.as<double>(), dpx.as<double>(), dy.as<double>(), dpy.as<double>(), n_particles,
                            out.as<double>(), nullptr));
        return out.download<double>(n_particles);
    }
    virtual std::vector<double> get_radius() const { return observable(HM_OBS_RADIUS); }
    virtual double get_radius_mean() const { return nan_mean(get_radius()); }
    virtual double get_radius_std() const { return nan_std(get_radius()); }
    virtual std::vector<double> get_action() const { return observable(HM_OBS_ACTION); }
    virtual double get_action_mean() const { return nan_mean(get_action()); }
    virtual double get_action_std() const { return nan_std(get_action()); }
    virtual std::vector<double> get_action_x() const { return observable(HM_OBS_ACTION_X); }
    virtual double get_action_x_mean() const { return nan_mean(get_action_x()); }
    virtual double get_action_x_std() const { return nan_std(get_action_x()); }
    virtual std::vector<double> get_action_y() const { return observable(HM_OBS_ACTION_Y); }
    virtual double get_action_y_mean() const { return nan_mean(get_action_y()); }
    virtual double get_action_y_std() const { return nan_std(get_action_y()); }
    virtual std::vector<double> get_angles_x() const { return observable(HM_OBS_ANGLE_X); }
    virtual double get_angles_x_mean() const { return nan_mean(get_angles_x()); }
    virtual double get_angles_x_std() const { return nan_std(get_angles_x()); }
    virtual std::vector<double> get_angles_y() const { return observable(HM_OBS_ANGLE_Y); }
    virtual double get_angles_y_mean() const { return nan_mean(get_angles_y()); }
    virtual double get_angles_y_std() const { return nan_std(get_angles_y()); }

    virtual std::vector<uint8_t> get_valid() const { return dvalid.download<uint8_t>(n_particles); }
    virtual std::vector<uint8_t> get_ghost() const
    {
        std::vector<uint8_t> v(n_particles, 1);
        std::fill(v.begin(), v.begin() + n_base, 0);
        return v;
    }
    virtual std::vector<size_t> get_idx() const
    {
        // base particles: their index; ghosts: their set (src/henon.cu:755)
        std::vector<size_t> v(n_particles);
        for (size_t i = 0; i < n_particles; i++) v[i] = i < n_base ? i : i / n_base - 1;
        return v;
    }
    virtual std::vector<size_t> get_idx_base() const
    {
        std::vector<size_t> v(n_particles);
        for (size_t i = 0; i < n_particles; i++) v[i] = n_base ? i % n_base : i;
        return v;
    }
    virtual size_t get_n_particles() const { return n_base; }
    virtual size_t get_n_ghosts_per_particle() const { return n_ghosts; }

    // Kick generator (hm_track_kicked): a fresh seed per object, as the
    // reference seeds its states from the clock; set_kick_seed fixes it.
    void set_kick_seed(uint64_t seed)
    {
        kick_seed = seed;
        kick_draws = 0;
    }
    uint64_t kick_seed = fresh_seed();
    uint64_t kick_draws = 0;

    size_t n_base = 0;       // the reference's n_particles: particles without their ghosts
    size_t n_ghosts = 0;     // ghost sets
    size_t n_particles = 0;  // entries of the device arrays: n_base * (1 + n_ghosts)
    unsigned int global_steps = 0;
    std::vector<double> x0, px0, y0, py0;
    DevBuf dx, dpx, dy, dpy, dsteps, dvalid;

  private:
    void upload_initial()
    {
        dx.upload(x0);
        dpx.upload(px0);
        dy.upload(y0);
        dpy.upload(py0);
        dsteps.upload(std::vector<unsigned int>(n_particles, 0u));
        dvalid.upload(std::vector<uint8_t>(n_particles, 1));
    }
    std::mt19937_64 rng;  // src/henon.h:63, default-seeded
};

class Particles4DGpu : public Particles4D {
  public:
    using Particles4D::Particles4D;
};

// ---------------------------------------------------------------- modulation
std::vector<double> basic_modulation(double tune, double omega, double epsilon, int start, int end)
{
    if (!(start < end)) throw std::invalid_argument("basic_modulation: start must be smaller than end");
    std::vector<double> v(end - start);
    check(hm_basic_modulation(tune, omega, epsilon, start, end, v.data()));
    return v;
}

std::vector<double> sps_modulation(double tune, double epsilon, int start, int end)
{
    if (!(start < end)) throw std::invalid_argument("sps_modulation: start must be smaller than end");
    std::vector<double> v(end - start);
    check(hm_sps_modulation(tune, epsilon, start, end, v.data()));
    return v;
}

std::vector<double> lines_modulation(double tune, double epsilon, const std::vector<double> &coef,
                                     const std::vector<double> &freq, int start, int end)
{
    if (!(start < end)) throw std::invalid_argument("modulation: start must be smaller than end");
    if (coef.size() != freq.size()) throw std::invalid_argument("modulation: coefficients and frequencies differ in length");
    std::vector<double> v(end - start);
    check(hm_lines_modulation(tune, epsilon, coef.data(), freq.data(), (int)coef.size(), start, end, v.data()));
    return v;
}


std::vector<double> gaussian_modulation(double tune, double sigma, int start, int end, py::object seed)
{
    if (!(start < end)) throw std::invalid_argument("gaussian_modulation: start must be smaller than end");
    std::vector<double> v(end - start);
    check(hm_gaussian_modulation(tune, sigma, start, end, seed.is_none() ? fresh_seed() : seed.cast<uint64_t>(),
                                 v.data()));
    return v;
}

std::vector<double> uniform_modulation(double from, double to, int start, int end, py::object seed)
{
    if (!(start < end)) throw std::invalid_argument("uniform_modulation: start must be smaller than end");
    std::vector<double> v(end - start);
    check(hm_uniform_modulation(from, to, start, end, seed.is_none() ? fresh_seed() : seed.cast<uint64_t>(),
                                v.data()));
    return v;
}

std::vector<double> birkhoff_weights(unsigned int n)
{
    std::vector<double> v(n);
    check(hm_birkhoff_weights(n, v.data()));
    return v;
}

int mode_of(const std::string &m)
{
    if (m == "exact") return HM_MODE_EXACT;
    if (m == "fast") return HM_MODE_FAST;
    if (m == "exact_orbit") return HM_MODE_EXACT_ORBIT;  // tangent_indicators only
    throw std::invalid_argument("mode must be 'exact' or 'fast' (or 'exact_orbit' for tangent_indicators)");
}

// ---------------------------------------------------------------- tracker
// henon_tracker (src/henon.h:196-232, src/henon.cu:2657-3112) with device tables.
class HenonTracker {
  public:
    HenonTracker(unsigned int N, double omega_x, double omega_y, std::string modulation_kind, double omega_0,
                 double epsilon, unsigned int offset, std::vector<double> coefficients = {},
                 std::vector<double> frequencies = {})
    {
        require_device();
        compute(N, omega_x, omega_y, modulation_kind, omega_0, epsilon, offset, coefficients, frequencies);
    }
    virtual ~HenonTracker() = default;

    virtual void compute_a_modulation(unsigned int N, double omega_x, double omega_y, std::string modulation_kind,
                                      double omega_0, double epsilon, unsigned int offset)
    {
        compute(N, omega_x, omega_y, modulation_kind, omega_0, epsilon, offset, coef_, freq_);
    }

    void check_turns(const Particles4D &p, unsigned int n_turns, bool inverse) const
    {
        // src/henon.cu:2706-2715
        if (inverse) {
            if (n_turns > p.global_steps) throw std::runtime_error("The number of turns is too large.");
            // particles tracked forward by a tracker with a longer table: the
            // inverse turns would index past this tracker's table (the
            // reference reads out of bounds there)
            if (p.global_steps > allowed_steps) throw std::runtime_error("The number of turns is too large.");
        } else {
            if ((uint64_t)n_turns + p.global_steps > allowed_steps) throw std::runtime_error("The number of turns is too large.");
        }
    }

    void track(Particles4D &p, unsigned int n_turns, double mu, double barrier, double kick_module, bool inverse,
               const std::string &mode)
    {
        check_turns(p, n_turns, inverse);
        const int m = mode_of(mode);
        {
            py::gil_scoped_release nogil;
            if (std::isnan(kick_module)) {
                check(hm_track(p.dx.as<double>(), p.dpx.as<double>(), p.dy.as<double>(), p.dpy.as<double>(),
                               p.dsteps.as<uint32_t>(), p.n_particles, dtable.as<double>(), allowed_steps, n_turns,
                               mu, barrier, kick_module, inverse ? 1 : 0, m, nullptr, nullptr));
            } else {
                // src/henon.cu:3155-3185 with the kick: the particle object's
                // generator stream (seed, draws so far) continues call to call
                check(hm_track_kicked(p.dx.as<double>(), p.dpx.as<double>(), p.dy.as<double>(), p.dpy.as<double>(),
                                      p.dsteps.as<uint32_t>(), p.n_particles, dtable.as<double>(), allowed_steps,
                                      n_turns, mu, barrier, kick_module, inverse ? 1 : 0, m, p.kick_seed,
                                      p.kick_draws, nullptr));
                p.kick_draws += n_turns;
            }
            uint32_t flags = 0;
            check(hm_take_status(&flags, nullptr));
            if (flags & 1u)
                throw std::runtime_error("track: a particle's turn index left the modulation table; it was stopped "
                                         "there (the reference would read out of bounds)");
        }
        if (!inverse) p.global_steps += n_turns;
        else p.global_steps -= n_turns;
        if (PyErr_CheckSignals() != 0) throw py::error_already_set();
    }

    // Launch-configuration autotuner (hm_autotune_track) on this tracker's
    // table and the given particles; the particles are not modified.
    py::dict autotune(const Particles4D &p, double mu, double barrier, const std::string &mode,
                      unsigned int warm_turns, unsigned int trial_turns)
    {
        const int m = mode_of(mode);
        hm_launch_cfg best{};
        double ms = 0.0;
        {
            py::gil_scoped_release nogil;
            check(hm_autotune_track(p.dx.as<double>(), p.dpx.as<double>(), p.dy.as<double>(), p.dpy.as<double>(),
                                    p.dsteps.as<uint32_t>(), p.n_particles, dtable.as<double>(), allowed_steps,
                                    warm_turns, trial_turns, mu, barrier, m, &best, &ms, nullptr));
        }
        py::dict d;
        d["block"] = best.block;
        d["compact_every"] = best.compact_every;
        d["particles_per_lane"] = best.particles_per_lane;
        d["trial_ms"] = ms;
        return d;
    }

    // One job of a batched launch (track_batch): this tracker's table, the
    // particles' device arrays, the same turn-range check as track().
    hm_job batch_job(Particles4D &p, unsigned int n_turns, double mu, double barrier) const
    {
        check_turns(p, n_turns, false);
        hm_job j{};
        j.x = p.dx.as<double>(); j.px = p.dpx.as<double>(); j.y = p.dy.as<double>(); j.py = p.dpy.as<double>();
        j.steps = p.dsteps.as<uint32_t>();
        j.n = p.n_particles;
        j.table = dtable.as<double>();
        j.table_len = allowed_steps;
        j.n_turns = n_turns;
        j.mu = mu;
        j.barrier = barrier;
        j.kick_module = std::nan("");
        j.megno = nullptr;
        j.renorm_every = 0;
        return j;
    }

    py::array_t<double> get_tangent_matrix(const Particles4D &p, double mu, bool reverse) const
    {
        // src/henon.cu:3080-3112 / :3187-3219, table index global_steps
        if (p.global_steps >= allowed_steps)
            throw std::runtime_error("get_tangent_matrix: global_steps is past the end of the modulation table");
        DevBuf out(p.n_particles * 16 * sizeof(double));
        check(hm_tangent_matrix(p.dx.as<double>(), p.dpx.as<double>(), p.dy.as<double>(), p.dpy.as<double>(),
                                p.dvalid.as<uint8_t>(), p.n_particles, dtable.as<double>(), allowed_steps,
                                p.global_steps, mu, reverse ? 1 : 0, out.as<double>(), nullptr));
        py::array_t<double> a({(py::ssize_t)p.n_particles, (py::ssize_t)4, (py::ssize_t)4});
        if (p.n_particles) check(hm_memcpy(a.mutable_data(), out.as<double>(), out.bytes(), 1, nullptr));
        sync();
        return a;
    }


    // henon_tracker::megno (src/henon.cu:2755-2887) on the GPU: trace form,
    // per-particle rows of turn samples, J at table index k of the call and the
    // map (forward, or inverse with inverse=True) at the particle's own index.
    // The reference's threads each take the particles j = t, t + n_threads, ...
    // in turn and carry the tangent matrix from one to the next with only its
    // diagonal reset to 1 (henon.cu:2794-2803), so a particle's result depends
    // on the particle before it in its thread (a lost one makes every later one
    // NaN). With carry_off_diagonal (the default) that chain is reproduced: a
    // first pass gives each particle's product P of its tangent matrices from the
    // identity, the host chains M_start(next) = P M_start(prev) with the
    // diagonal set to 1 (the reference's running product up to rounding), and
    // the MEGNO pass starts every particle from its M_start. n_threads = -1 is
    // std::thread::hardware_concurrency(), as in the reference.
    py::array_t<double> megno(Particles4D &p, unsigned int n_turns, double mu, double barrier, double kick_module,
                              bool inverse, std::vector<unsigned int> turn_samples, int n_threads,
                              bool carry_off_diagonal)
    {
        if (!std::isnan(kick_module)) throw NotImplemented("megno: the random kick is not available");
        if (n_threads == 0 || n_threads < -1) throw std::invalid_argument("megno: n_threads must be positive or -1");
        check_turns(p, n_turns, inverse);
        if (turn_samples.empty()) turn_samples.push_back(n_turns);
        for (unsigned int t : turn_samples)
            if (t > n_turns) throw std::invalid_argument("megno: turn sample beyond n_turns");
        const size_t n = p.n_particles;
        const size_t nt = n_threads == -1 ? std::max(1u, std::thread::hardware_concurrency()) : (size_t)n_threads;
        std::vector<uint32_t> ts(turn_samples.begin(), turn_samples.end());
        py::array_t<double> out({(py::ssize_t)n, (py::ssize_t)ts.size()});
        if (n == 0) {
            p.global_steps = inverse ? p.global_steps - n_turns : p.global_steps + n_turns;
            return out;
        }
        // every pass restarts from the same state: keep a copy
        DevBuf sx(n * 8), spx(n * 8), sy(n * 8), spy(n * 8), ss(n * 4);
        snapshot(p, sx, spx, sy, spy, ss);
        const bool chain = carry_off_diagonal && nt < n;
        DevBuf m_start(chain ? n * 128 : 8);
        const int inv = inverse ? 1 : 0;
        if (chain) {
            DevBuf prod(n * 128), scratch(n * 8);
            {
                py::gil_scoped_release nogil;
                check(hm_megno_trace(p.dx.as<double>(), p.dpx.as<double>(), p.dy.as<double>(), p.dpy.as<double>(),
                                     p.dsteps.as<uint32_t>(), n, dtable.as<double>(), allowed_steps, n_turns, mu,
                                     barrier, inv, nullptr, 0, nullptr, prod.as<double>(), scratch.as<double>(),
                                     nullptr));
                sync();
            }
            restore(p, sx, spx, sy, spy, ss);
            const std::vector<double> P = prod.download<double>(n * 16);
            std::vector<double> M(n * 16);
            for (size_t c = 0; c < nt; ++c) {
                double cur[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
                for (size_t j = c; j < n; j += nt) {
                    cur[0] = cur[5] = cur[10] = cur[15] = 1.0;
                    std::copy(cur, cur + 16, M.begin() + 16 * j);
                    // multiply_matrices' order (henon.cu:251-265): += from 0.0 over k
                    double next[16];
                    for (int r = 0; r < 4; ++r)
                        for (int q = 0; q < 4; ++q) {
                            double acc = 0.0;
                            for (int k = 0; k < 4; ++k) acc += P[16 * j + 4 * r + k] * cur[4 * k + q];
                            next[4 * r + q] = acc;
                        }
                    std::copy(next, next + 16, cur);
                }
            }
            m_start.upload(M);
        }
        for (size_t q0 = 0; q0 < ts.size(); q0 += HM_MAX_SAMPLES) {
            const size_t nq = std::min<size_t>(HM_MAX_SAMPLES, ts.size() - q0);
            DevBuf o(n * nq * sizeof(double));
            {
                py::gil_scoped_release nogil;
                check(hm_megno_trace(p.dx.as<double>(), p.dpx.as<double>(), p.dy.as<double>(), p.dpy.as<double>(),
                                     p.dsteps.as<uint32_t>(), n, dtable.as<double>(), allowed_steps, n_turns, mu,
                                     barrier, inv, ts.data() + q0, nq, chain ? m_start.as<double>() : nullptr,
                                     nullptr, o.as<double>(), nullptr));
                sync();
            }
            auto v = o.download<double>(n * nq);
            auto r = out.mutable_unchecked<2>();
            for (size_t i = 0; i < n; i++)
                for (size_t q = 0; q < nq; q++) r(i, q0 + q) = v[i * nq + q];
            if (q0 + nq < ts.size()) restore(p, sx, spx, sy, spy, ss);
        }
        uint32_t flags = 0;
        check(hm_take_status(&flags, nullptr));
        if (flags & 1u)
            throw std::runtime_error("megno: a particle's turn index left the modulation table; it was stopped there "
                                     "(the reference would read out of bounds)");
        if (!inverse) p.global_steps += n_turns;
        else p.global_steps -= n_turns;
        return out;
    }

    // History-based tunes (src/henon.cu:2890-3078): windows plus the appended
    // (0, n_turns) window, then the full signal; all_tunes removes the
    // per-coordinate means first. Particles are processed in chunks whose
    // trajectories fit the history budget.
    py::object history_tunes(Particles4D &p, unsigned int n_turns, double mu, double barrier, double kick_module,
                              bool inverse, std::vector<unsigned int> from_idx, std::vector<unsigned int> to_idx,
                              bool all)
    {
        if (!std::isnan(kick_module)) throw NotImplemented("tunes: the random kick is not available");
        if (from_idx.size() != to_idx.size())
            throw std::runtime_error("compute_tune: from and to must be of the same size");
        check_turns(p, n_turns, inverse);
        from_idx.push_back(0);
        to_idx.push_back(n_turns);
        const size_t W = from_idx.size();
        const uint32_t L = n_turns + 1;
        for (size_t w = 0; w < W; w++)
            if (from_idx[w] >= L || to_idx[w] >= L) throw std::runtime_error("compute_tune: from and to must be valid indices");
        // Birkhoff windows: the given ones (with the appended), then the full signal
        std::vector<uint32_t> bf(from_idx.begin(), from_idx.end()), bt(to_idx.begin(), to_idx.end());
        bf.push_back(0);
        bt.push_back(L - 1);
        // FFT windows: [from, to] (length to-from+1), then samples 1..n (length n)
        std::vector<uint32_t> ff(bf), fl;
        for (size_t w = 0; w < W; w++) fl.push_back(to_idx[w] - from_idx[w] + 1);
        ff.back() = 1;
        fl.push_back(L - 1);
        const size_t NW = bf.size();
        // the trajectory as particle rows (row_len a multiple of 16 samples):
        // the FFT reads each particle's samples in place
        const size_t row_len = ((size_t)L + 15) / 16 * 16;
        const size_t per = row_len * 32;
        // as many particles per chunk as half the free device memory holds
        // (at least 8 GiB): tracking the rows needs many wavefronts
        size_t free_b = 0, total_b = 0;
        check(hm_mem_info(&free_b, &total_b));
        const size_t budget = std::max<size_t>((size_t)8 << 30, free_b / 2);
        const size_t chunk = std::max<size_t>(1, std::min<size_t>(p.n_particles, budget / per));
        std::vector<double> birk(p.n_particles * 2 * NW), fft(all ? p.n_particles * 2 * NW : 0);
        DevBuf hist(chunk * per), means(all ? chunk * 32 : 8), ob(chunk * 2 * NW * 8), of(chunk * 2 * NW * 8);
        {
            py::gil_scoped_release nogil;
            for (size_t c0 = 0; c0 < p.n_particles; c0 += chunk) {
                const size_t nc = std::min(chunk, p.n_particles - c0);
                // all_tunes: the per-coordinate means come out of the same pass
                check(hm_track_rows(p.dx.as<double>() + c0, p.dpx.as<double>() + c0, p.dy.as<double>() + c0,
                                    p.dpy.as<double>() + c0, p.dsteps.as<uint32_t>() + c0, nc, dtable.as<double>(),
                                    allowed_steps, n_turns, mu, barrier, inverse ? 1 : 0, hist.as<double>(), row_len,
                                    all ? means.as<double>() : nullptr, nullptr));
                const double *mp = all ? means.as<double>() : nullptr;
                // the last two Birkhoff windows, (0, n_turns) and the full signal
                // 0..L-1 = 0..n_turns, are the same: computed once, copied below
                const size_t last = NW - 1;
                if (all) {
                    // the given windows and (0, n_turns): both tunes from one read
                    // of each window; the full signal's FFT takes samples 1..n_turns
                    check(hm_tunes_rows(hist.as<double>(), nc, row_len, L, mp, bf.data(), bt.data(), last,
                                        ob.as<double>(), of.as<double>(), nullptr));
                    check(hm_fft_tunes_rows(hist.as<double>(), nc, row_len, L, mp, ff.data() + last, fl.data() + last,
                                            1, of.as<double>() + last * 2 * nc, nullptr));
                } else {
                    check(hm_birkhoff_tunes_rows(hist.as<double>(), nc, row_len, L, mp, bf.data(), bt.data(), last,
                                                 ob.as<double>(), nullptr));
                }
                auto vb = ob.download<double>(nc * 2 * last);
                for (size_t w = 0; w < NW; w++) {
                    const size_t src = std::min(w, last - 1);
                    for (int pl = 0; pl < 2; pl++)
                        for (size_t i = 0; i < nc; i++)
                            birk[(c0 + i) * 2 * NW + 2 * w + pl] = vb[(src * 2 + pl) * nc + i];
                }
                if (all) {
                    auto vf = of.download<double>(nc * 2 * NW);
                    for (size_t w = 0; w < NW; w++)
                        for (int pl = 0; pl < 2; pl++)
                            for (size_t i = 0; i < nc; i++) fft[(c0 + i) * 2 * NW + 2 * w + pl] = vf[(w * 2 + pl) * nc + i];
                }
            }
            uint32_t flags = 0;
            check(hm_take_status(&flags, nullptr));
            // a stopped lane's frozen tail would be stored as samples: the
            // tunes are not valid then
            if (flags & 1u)
                throw std::runtime_error("tunes: a particle's turn index left the modulation table; it was stopped "
                                         "there (the reference would read out of bounds)");
        }
        if (!inverse) p.global_steps += n_turns;
        else p.global_steps -= n_turns;
        py::array_t<double> rb({(py::ssize_t)p.n_particles, (py::ssize_t)(2 * NW)});
        std::memcpy(rb.mutable_data(), birk.data(), birk.size() * sizeof(double));
        if (!all) return rb;
        py::array_t<double> rf({(py::ssize_t)p.n_particles, (py::ssize_t)(2 * NW)});
        std::memcpy(rf.mutable_data(), fft.data(), fft.size() * sizeof(double));
        return py::make_tuple(rb, rf);
    }

    // Fused tracking + tangent map + indicators (no reference counterpart; the
    // reference builds these from per-turn constructs driven from Python).
    py::dict tangent_indicators(Particles4D &p, unsigned int n_turns, double mu, double barrier,
                                unsigned int renorm_every, std::vector<unsigned int> samples,
                                std::vector<double> fli_direction, std::vector<std::pair<unsigned, unsigned>> birkhoff,
                                std::pair<unsigned, unsigned> history, const std::string &mode,
                                std::vector<unsigned int> megno_birkhoff, std::vector<unsigned int> tune_birkhoff)
    {
        check_turns(p, n_turns, false);
        if (fli_direction.size() != 4) throw std::invalid_argument("tangent_indicators: fli_direction needs 4 entries");
        const size_t n = p.n_particles, S = samples.size(), B = birkhoff.size();
        const size_t MB = megno_birkhoff.size(), TB = tune_birkhoff.size();
        // The kernel takes 16 samples and 4 windows of each kind per launch: more
        // run as further passes over the same turns from the same start
        auto chunks = [](size_t v, size_t per) { return (v + per - 1) / per; };
        const size_t passes = std::max<size_t>(
            {1, chunks(S, HM_MAX_SAMPLES), chunks(B, HM_MAX_BIRKHOFF), chunks(MB, HM_MAX_BIRKHOFF),
             chunks(TB, HM_MAX_BIRKHOFF)});
        DevBuf mg(n * S * 8), fl(n * S * 8), mat(n * 128), ls(n * 8), bo(n * 2 * B * 8),
            hi(n * 32 * (size_t)history.second), mbo(n * MB * 8), tbo(n * 4 * TB * 8);
        std::vector<double> eye(16 * n, 0.0);
        for (size_t i = 0; i < n; i++) eye[16 * i] = eye[16 * i + 5] = eye[16 * i + 10] = eye[16 * i + 15] = 1.0;
        DevBuf sx(passes > 1 ? n * 8 : 8), spx(passes > 1 ? n * 8 : 8), sy(passes > 1 ? n * 8 : 8),
            spy(passes > 1 ? n * 8 : 8), ss(passes > 1 ? n * 4 : 8);
        if (passes > 1) snapshot(p, sx, spx, sy, spy, ss);
        for (size_t pass = 0; pass < passes; ++pass) {
            if (pass) restore(p, sx, spx, sy, spy, ss);
            mat.upload(eye);
            ls.upload(std::vector<double>(n, 0.0));
            hm_indicator_spec spec{};
            spec.renorm_every = renorm_every;
            for (int q = 0; q < 4; q++) spec.fli_dir[q] = fli_direction[q];
            spec.matrix = mat.as<double>();
            spec.log_scale = ls.as<double>();
            const size_t s0 = std::min(S, pass * HM_MAX_SAMPLES), s1 = std::min(S, s0 + HM_MAX_SAMPLES);
            spec.n_samples = (uint32_t)(s1 - s0);
            for (size_t q = s0; q < s1; q++) spec.samples[q - s0] = samples[q];
            if (s1 > s0) { spec.megno = mg.as<double>() + s0 * n; spec.fli = fl.as<double>() + s0 * n; }
            const size_t b0 = std::min(B, pass * HM_MAX_BIRKHOFF), b1 = std::min(B, b0 + HM_MAX_BIRKHOFF);
            spec.n_birkhoff = (uint32_t)(b1 - b0);
            for (size_t w = b0; w < b1; w++) {
                spec.birk_from[w - b0] = birkhoff[w].first;
                spec.birk_to[w - b0] = birkhoff[w].second;
            }
            if (b1 > b0) spec.birk_out = bo.as<double>() + b0 * 2 * n;
            const size_t m0 = std::min(MB, pass * HM_MAX_BIRKHOFF), m1 = std::min(MB, m0 + HM_MAX_BIRKHOFF);
            spec.n_megno_birkhoff = (uint32_t)(m1 - m0);
            for (size_t w = m0; w < m1; w++) spec.megno_birkhoff_samples[w - m0] = megno_birkhoff[w];
            if (m1 > m0) spec.megno_birkhoff_out = mbo.as<double>() + m0 * n;
            const size_t t0 = std::min(TB, pass * HM_MAX_BIRKHOFF), t1 = std::min(TB, t0 + HM_MAX_BIRKHOFF);
            spec.n_tune_birkhoff = (uint32_t)(t1 - t0);
            for (size_t w = t0; w < t1; w++) spec.tune_birkhoff_samples[w - t0] = tune_birkhoff[w];
            if (t1 > t0) spec.tune_birkhoff_out = tbo.as<double>() + t0 * 4 * n;
            if (pass == 0 && history.second) {
                spec.hist_from = history.first;
                spec.hist_len = history.second;
                spec.hist = hi.as<double>();
            }
            py::gil_scoped_release nogil;
            check(hm_track_indicators(p.dx.as<double>(), p.dpx.as<double>(), p.dy.as<double>(), p.dpy.as<double>(),
                                      p.dsteps.as<uint32_t>(), n, dtable.as<double>(), allowed_steps, n_turns, mu,
                                      barrier, mode_of(mode), &spec, nullptr));
            sync();
        }
        uint32_t flags = 0;
        check(hm_take_status(&flags, nullptr));
        if (flags & 1u)
            throw std::runtime_error("tangent_indicators: a particle's turn index left the modulation table; it was "
                                     "stopped there (the reference would read out of bounds)");
        p.global_steps += n_turns;
        py::dict d;
        auto arr = [](const DevBuf &b, std::vector<py::ssize_t> shape) {
            size_t cnt = 1;
            for (auto v : shape) cnt *= (size_t)v;
            py::array_t<double> a(shape);
            if (cnt) check(hm_memcpy(a.mutable_data(), b.as<double>(), cnt * 8, 1, nullptr));
            sync();
            return a;
        };
        d["megno"] = arr(mg, {(py::ssize_t)S, (py::ssize_t)n});
        d["fli"] = arr(fl, {(py::ssize_t)S, (py::ssize_t)n});
        d["matrix"] = arr(mat, {(py::ssize_t)n, 4, 4});
        d["log_scale"] = arr(ls, {(py::ssize_t)n});
        d["birkhoff"] = arr(bo, {(py::ssize_t)B, 2, (py::ssize_t)n});
        d["megno_birkhoff"] = arr(mbo, {(py::ssize_t)MB, (py::ssize_t)n});
        d["tune_birkhoff"] = arr(tbo, {(py::ssize_t)TB, 4, (py::ssize_t)n});
        if (history.second) d["history"] = arr(hi, {(py::ssize_t)history.second, 4, (py::ssize_t)n});
        return d;
    }

    // Reversibility error (henon_map_cpp/dynamic_indicators.py:136-151): for each t,
    // |x0 - B^t F^t x0| over the four coordinates, from one forward pass with
    // snapshots at the requested turns and one inverse pass per snapshot.
    py::array_t<double> reversibility_error(Particles4D &p, std::vector<unsigned int> t_list, double mu,
                                            double barrier, const std::string &mode)
    {
        std::vector<unsigned int> ts(t_list);
        std::sort(ts.begin(), ts.end());
        if (ts.empty()) return py::array_t<double>(std::vector<py::ssize_t>{0, (py::ssize_t)p.n_particles});
        check_turns(p, ts.back(), false);
        const size_t n = p.n_particles;
        const int m = mode_of(mode);
        auto x0 = p.get_x(), px0 = p.get_px(), y0 = p.get_y(), py0 = p.get_py();
        std::vector<std::vector<double>> err(ts.size());
        DevBuf sx(n * 8), spx(n * 8), sy(n * 8), spy(n * 8), ss(n * 4);
        DevBuf kx(n * 8), kpx(n * 8), ky(n * 8), kpy(n * 8), ks(n * 4);
        unsigned int done = 0;
        for (size_t q = 0; q < ts.size(); q++) {
            {
                py::gil_scoped_release nogil;
                if (ts[q] > done)
                    check(hm_track(p.dx.as<double>(), p.dpx.as<double>(), p.dy.as<double>(), p.dpy.as<double>(),
                                   p.dsteps.as<uint32_t>(), n, dtable.as<double>(), allowed_steps, ts[q] - done, mu,
                                   barrier, std::numeric_limits<double>::quiet_NaN(), 0, m, nullptr, nullptr));
            }
            done = ts[q];
            snapshot(p, kx, kpx, ky, kpy, ks);  // forward state at t
            snapshot(p, sx, spx, sy, spy, ss);
            {
                py::gil_scoped_release nogil;
                check(hm_track(sx.as<double>(), spx.as<double>(), sy.as<double>(), spy.as<double>(), ss.as<uint32_t>(),
                               n, dtable.as<double>(), allowed_steps, ts[q], mu, barrier,
                               std::numeric_limits<double>::quiet_NaN(), 1, m, nullptr, nullptr));
                sync();
            }
            auto bx = sx.download<double>(n), bpx = spx.download<double>(n), by = sy.download<double>(n),
                 bpy = spy.download<double>(n);
            err[q].resize(n);
            for (size_t i = 0; i < n; i++)
                err[q][i] = std::sqrt((bx[i] - x0[i]) * (bx[i] - x0[i]) + (bpx[i] - px0[i]) * (bpx[i] - px0[i]) +
                                      (by[i] - y0[i]) * (by[i] - y0[i]) + (bpy[i] - py0[i]) * (bpy[i] - py0[i]));
        }
        p.global_steps += done;
        py::array_t<double> out({(py::ssize_t)ts.size(), (py::ssize_t)n});
        auto r = out.mutable_unchecked<2>();
        for (size_t q = 0; q < ts.size(); q++)
            for (size_t i = 0; i < n; i++) r(q, i) = err[q][i];
        return out;
    }

    static void snapshot(const Particles4D &p, DevBuf &x, DevBuf &px, DevBuf &y, DevBuf &py, DevBuf &st)
    {
        const size_t n = p.n_particles;
        check(hm_memcpy(x.as<void>(), p.dx.as<void>(), n * 8, 2, nullptr));
        check(hm_memcpy(px.as<void>(), p.dpx.as<void>(), n * 8, 2, nullptr));
        check(hm_memcpy(y.as<void>(), p.dy.as<void>(), n * 8, 2, nullptr));
        check(hm_memcpy(py.as<void>(), p.dpy.as<void>(), n * 8, 2, nullptr));
        check(hm_memcpy(st.as<void>(), p.dsteps.as<void>(), n * 4, 2, nullptr));
        sync();
    }
    static void restore(Particles4D &p, DevBuf &x, DevBuf &px, DevBuf &y, DevBuf &py, DevBuf &st)
    {
        const size_t n = p.n_particles;
        check(hm_memcpy(p.dx.as<void>(), x.as<void>(), n * 8, 2, nullptr));
        check(hm_memcpy(p.dpx.as<void>(), px.as<void>(), n * 8, 2, nullptr));
        check(hm_memcpy(p.dy.as<void>(), y.as<void>(), n * 8, 2, nullptr));
        check(hm_memcpy(p.dpy.as<void>(), py.as<void>(), n * 8, 2, nullptr));
        check(hm_memcpy(p.dsteps.as<void>(), st.as<void>(), n * 4, 2, nullptr));
        sync();
    }

    std::vector<double> get_omega_x() const { return wx; }
    std::vector<double> get_omega_y() const { return wy; }

    size_t allowed_steps = 0;
    std::vector<double> wx, wy;
    DevBuf dtable;

  protected:
    void compute(unsigned int N, double omega_x, double omega_y, const std::string &kind, double omega_0,
                 double epsilon, unsigned int offset, const std::vector<double> &coef, const std::vector<double> &freq)
    {
        // src/modulation.cpp:98-134 (pick_a_modulation), then the sin/cos table
        const int a = (int)offset, b = (int)(offset + N);
        if (N == 0) throw std::invalid_argument("henon_tracker: the number of turns must be positive");
        if (kind == "sps") {
            wx = sps_modulation(omega_x, epsilon, a, b);
            wy = sps_modulation(omega_y, epsilon, a, b);
        } else if (kind == "basic") {
            if (std::isnan(omega_0)) throw std::invalid_argument("basic modulation needs omega_0");
            wx = basic_modulation(omega_x, omega_0, epsilon, a, b);
            wy = basic_modulation(omega_y, omega_0, epsilon, a, b);
        } else if (kind == "none") {
            // the reference fills empty vectors (undefined tables); here: constant 2 pi tune
            wx.assign(N, 2 * M_PI * omega_x);
            wy.assign(N, 2 * M_PI * omega_y);
        } else if (kind == "gaussian") {
            wx = gaussian_modulation(omega_x, epsilon, a, b, py::none());
            wy = gaussian_modulation(omega_y, epsilon, a, b, py::none());
        } else if (kind == "uniform") {
            // src/modulation.cpp:124-128 passes (omega, epsilon) as (from, to)
            wx = uniform_modulation(omega_x, epsilon, a, b, py::none());
            wy = uniform_modulation(omega_y, epsilon, a, b, py::none());
        } else if (kind == "custom") {
            if (coef.empty()) throw std::invalid_argument("custom modulation needs coefficients and frequencies");
            wx = lines_modulation(omega_x, epsilon, coef, freq, a, b);
            wy = lines_modulation(omega_y, epsilon, coef, freq, a, b);
        } else {
            throw std::runtime_error("Unknown modulation kind");
        }
        coef_ = coef;
        freq_ = freq;
        std::vector<double> tab(4 * (size_t)N);
        check(hm_build_table(wx.data(), wy.data(), N, tab.data()));
        dtable.upload(tab);
        allowed_steps = N;
    }
    std::vector<double> coef_, freq_;
};

class HenonTrackerGpu : public HenonTracker {
  public:
    using HenonTracker::HenonTracker;
};


// ---------------------------------------------------------------- constructs
// An (n, 4) numpy array of 4-vectors, read with one copy.
std::vector<double> flatten_n4(const Coords &a, size_t n)
{
    if (a.ndim() != 2 || (size_t)a.shape(0) != n) throw std::invalid_argument("expected one 4-vector per particle");
    if (a.shape(1) != 4) throw std::invalid_argument("expected 4-vectors");
    std::vector<double> f(4 * n);
    if (n) std::memcpy(f.data(), a.data(), f.size() * sizeof(double));
    return f;
}

std::vector<double> flatten_n4(const std::vector<std::vector<double>> &v, size_t n)
{
    if (v.size() != n) throw std::invalid_argument("expected one 4-vector per particle");
    std::vector<double> f(4 * n);
    for (size_t i = 0; i < n; i++) {
        if (v[i].size() != 4) throw std::invalid_argument("expected 4-vectors");
        for (int q = 0; q < 4; q++) f[4 * i + q] = v[i][q];
    }
    return f;
}

py::array_t<double> to_array(const DevBuf &b, std::vector<py::ssize_t> shape)
{
    size_t cnt = 1;
    for (auto v : shape) cnt *= (size_t)v;
    py::array_t<double> a(shape);
    if (cnt) check(hm_memcpy(a.mutable_data(), b.as<double>(), cnt * 8, 1, nullptr));
    sync();
    return a;
}

// matrix_4d_vector (src/henon.cu:1694-1787) and matrix_4d_vector_gpu
// (:1789-1946): N tangent matrices on the device, identity at construction.
class Matrix4DVector {
  public:
    explicit Matrix4DVector(size_t N) : N(N), d(N * 16 * sizeof(double)) { require_device(); reset(); }
    void reset() { check(hm_matrix_identity(d.as<double>(), N, nullptr)); sync(); }
    void multiply(py::array_t<double, py::array::c_style | py::array::forcecast> l)
    {
        if ((size_t)l.size() != N * 16) throw std::invalid_argument("multiply: expected an (N, 4, 4) array");
        DevBuf lb(N * 128);
        check(hm_memcpy(lb.as<void>(), l.data(), N * 128, 0, nullptr));
        check(hm_matrix_left_multiply(d.as<double>(), lb.as<double>(), N, nullptr));
        sync();
    }
    // matrix_4d_vector::structured_multiply (henon.cu:1743-1755): M <- J M, J
    // (or J^-1 when reverse) of the tracker at the particles' global_steps
    void structured_multiply_rev(const HenonTracker &t, const Particles4D &p, double mu, bool reverse)
    {
        if (p.n_particles != N) throw std::invalid_argument("structured_multiply: size mismatch");
        if (!reverse) return structured_multiply(t, p, mu);
        if (p.global_steps >= t.allowed_steps) throw std::runtime_error("structured_multiply: past the table end");
        DevBuf j(N * 128);
        check(hm_tangent_matrix(p.dx.as<double>(), p.dpx.as<double>(), p.dy.as<double>(), p.dpy.as<double>(),
                                p.dvalid.as<uint8_t>(), N, t.dtable.as<double>(), t.allowed_steps, p.global_steps, mu,
                                1, j.as<double>(), nullptr));
        check(hm_matrix_left_multiply(d.as<double>(), j.as<double>(), N, nullptr));
        sync();
    }
    // get_matrices_and_multiply (henon.cu:267-295, launched at :1827-1835)
    void structured_multiply(const HenonTracker &t, const Particles4D &p, double mu)
    {
        if (p.n_particles != N) throw std::invalid_argument("structured_multiply: size mismatch");
        check(hm_tangent_apply(d.as<double>(), p.dx.as<double>(), p.dpx.as<double>(), p.dy.as<double>(),
                               p.dpy.as<double>(), N, t.dtable.as<double>(), t.allowed_steps, p.global_steps, mu, 1,
                               nullptr));
        sync();
    }
    // get_matrices_and_set (henon.cu:297-313, :1837-1845)
    void set_with_tracker(const HenonTracker &t, const Particles4D &p, double mu)
    {
        if (p.n_particles != N) throw std::invalid_argument("set_with_tracker: size mismatch");
        check(hm_tangent_apply(d.as<double>(), p.dx.as<double>(), p.dpx.as<double>(), p.dy.as<double>(),
                               p.dpy.as<double>(), N, t.dtable.as<double>(), t.allowed_steps, p.global_steps, mu, 0,
                               nullptr));
        sync();
    }
    void explicit_copy(const Matrix4DVector &o)
    {
        if (o.N != N) throw std::invalid_argument("explicit_copy: size mismatch");
        check(hm_memcpy(d.as<void>(), o.d.as<void>(), N * 128, 2, nullptr));
        sync();
    }
    py::array_t<double> get_matrix() const { return to_array(d, {(py::ssize_t)N, 4, 4}); }
    // get_vector (henon.cu:1758-1787 / 1848-1903): M rv, one 4-vector per particle
    py::array_t<double> get_vector(const std::vector<std::vector<double>> &rv) const
    {
        DevBuf v(N * 32), o(N * 32);
        v.upload(flatten_n4(rv, N));
        check(hm_matvec(d.as<double>(), v.as<double>(), o.as<double>(), N, 0, nullptr));
        return to_array(o, {(py::ssize_t)N, 4});
    }
    size_t N;
    DevBuf d;
};

class Matrix4DVectorGpu : public Matrix4DVector {
  public:
    using Matrix4DVector::Matrix4DVector;
};

// vector_4d_gpu (src/henon.cu:1948-2039)
class Vector4D {
  public:
    explicit Vector4D(const std::vector<std::vector<double>> &rv) : N(rv.size()), d(rv.size() * 32)
    {
        require_device();
        set_vectors(rv);
    }
    explicit Vector4D(const Coords &a) : N(a.ndim() == 2 ? (size_t)a.shape(0) : 0), d(N * 32)
    {
        if (a.ndim() != 2) throw py::type_error("vector_4d_gpu: expected an array of 4-vectors");
        require_device();
        set_array(a);
    }
    void set_vectors(const std::vector<std::vector<double>> &rv) { d.upload(flatten_n4(rv, N)); }
    // one 4-vector per particle, or one 4-vector for all of them
    void set_array(const Coords &a)
    {
        if (a.ndim() == 1) set_vector(coords_vector(a));
        else d.upload(flatten_n4(a, N));
    }
    void set_vector(const std::vector<double> &v)
    {
        if (v.size() != 4) throw std::invalid_argument("set_vectors: expected 4 entries");
        std::vector<double> f(4 * N);
        for (size_t i = 0; i < N; i++)
            for (int q = 0; q < 4; q++) f[4 * i + q] = v[q];
        d.upload(f);
    }
    void multiply(const Matrix4DVector &m)
    {
        if (m.N != N) throw std::invalid_argument("multiply: size mismatch");
        check(hm_matvec(m.d.as<double>(), d.as<double>(), d.as<double>(), N, 1, nullptr));
        sync();
    }
    void normalize() { check(hm_normalize(d.as<double>(), N, nullptr)); sync(); }
    py::array_t<double> get_vectors() const { return to_array(d, {(py::ssize_t)N, 4}); }
    size_t N;
    DevBuf d;
};

std::vector<double> weights_of(size_t n)
{
    std::vector<double> w(n);
    if (n) check(hm_birkhoff_weights((uint32_t)n, w.data()));
    return w;
}

// lyapunov_birkhoff_construct (src/henon.cu:2042-2127)
class LyapunovConstruct {
  public:
    LyapunovConstruct(size_t N, size_t n_weights) : N(N), v(N * 8), vb(N * 8) { require_device(); change_weights(n_weights); reset(); }
    void reset()
    {
        idx = 0;
        v.upload(std::vector<double>(N, 0.0));
        vb.upload(std::vector<double>(N, 0.0));
    }
    void change_weights(size_t n) { n_weights = n; w.upload(weights_of(n)); }
    void add(const Vector4D &vec)
    {
        if (vec.N != N) throw std::invalid_argument("add: size mismatch");
        check(hm_lyapunov_add(v.as<double>(), vb.as<double>(), vec.d.as<double>(), w.as<double>(), N, n_weights, idx,
                              nullptr));
        sync();
        idx++;
    }
    py::array_t<double> get_weights() const { return to_array(w, {(py::ssize_t)n_weights}); }
    py::array_t<double> get_values_raw() const { return to_array(v, {(py::ssize_t)N}); }
    py::array_t<double> get_values_b() const { return to_array(vb, {(py::ssize_t)N}); }
    size_t N, n_weights = 0, idx = 0;
    DevBuf v, vb, w;
};

// lyapunov_birkhoff_construct_multi (src/henon.cu:2129-2220)
class LyapunovConstructMulti {
  public:
    LyapunovConstructMulti(size_t N, std::vector<size_t> nw) : N(N), n_weights(nw)
    {
        for (auto n : nw) {
            parts.emplace_back(std::make_unique<LyapunovConstruct>(N, n));
        }
    }
    void reset()
    {
        idx = 0;
        for (auto &p : parts) p->reset();
    }
    void add(const Vector4D &vec)
    {
        for (size_t i = 0; i < parts.size(); i++) {
            if (idx < n_weights[i]) {  // henon.cu:2173-2184
                parts[i]->idx = idx;
                parts[i]->add(vec);
            }
        }
        idx++;
    }
    py::list get_values_raw() const
    {
        py::list l;
        for (auto &p : parts) l.append(p->get_values_raw());
        return l;
    }
    py::list get_values_b() const
    {
        py::list l;
        for (auto &p : parts) l.append(p->get_values_b());
        return l;
    }
    size_t N, idx = 0;
    std::vector<size_t> n_weights;
    std::vector<std::unique_ptr<LyapunovConstruct>> parts;
};

// megno_construct (src/henon.cu:2223-2299)
class MegnoConstruct {
  public:
    explicit MegnoConstruct(size_t N) : N(N), l1(N * 8), l2(N * 8), l3(N * 8) { require_device(); reset(); }
    void reset()
    {
        idx = 1;
        for (DevBuf *b : {&l1, &l2, &l3}) b->upload(std::vector<double>(N, 0.0));
    }
    void add(const Matrix4DVector &a, const Matrix4DVector &b)
    {
        if (a.N != N || b.N != N) throw std::invalid_argument("add: size mismatch");
        check(hm_megno_add(l1.as<double>(), l2.as<double>(), l3.as<double>(), a.d.as<double>(), b.d.as<double>(), N, idx,
                           nullptr));
        sync();
        idx++;
    }
    py::array_t<double> get_values() const { return to_array(l3, {(py::ssize_t)N}); }
    size_t N, idx = 1;
    DevBuf l1, l2, l3;
};

// Flat device weight tables for per-sample Birkhoff constructs.
struct SampleWeights {
    explicit SampleWeights(const std::vector<size_t> &ts)
    {
        std::vector<double> flat;
        std::vector<uint64_t> off, smp;
        for (auto t : ts) {
            off.push_back(flat.size());
            smp.push_back(t);
            auto w = weights_of(t);
            flat.insert(flat.end(), w.begin(), w.end());
        }
        if (flat.empty()) flat.push_back(0.0);
        w.upload(flat);
        o.upload(off.empty() ? std::vector<uint64_t>{0} : off);
        s.upload(smp.empty() ? std::vector<uint64_t>{0} : smp);
    }
    DevBuf w, o, s;
};

// megno_birkhoff_construct (src/henon.cu:2302-2426)
class MegnoBirkhoffConstruct {
  public:
    MegnoBirkhoffConstruct(size_t N, std::vector<size_t> ts) : N(N), ts(ts), sw(ts), l1(N * 8), l2(N * 8 * std::max<size_t>(ts.size(), 1))
    {
        require_device();
        reset();
    }
    void reset()
    {
        idx = 1;
        l1.upload(std::vector<double>(N, 0.0));
        l2.upload(std::vector<double>(N * std::max<size_t>(ts.size(), 1), 0.0));
    }
    void add(const Matrix4DVector &a, const Matrix4DVector &b)
    {
        check(hm_megno_birkhoff_add(l1.as<double>(), l2.as<double>(), a.d.as<double>(), b.d.as<double>(), N, idx,
                                    sw.w.as<double>(), sw.o.as<uint64_t>(), sw.s.as<uint64_t>(), (uint32_t)ts.size(),
                                    nullptr));
        sync();
        idx++;
    }
    py::array_t<double> get_values() const { return to_array(l2, {(py::ssize_t)ts.size(), (py::ssize_t)N}); }
    size_t N, idx = 1;
    std::vector<size_t> ts;
    SampleWeights sw;
    DevBuf l1, l2;
};

// tune_birkhoff_construct (src/henon.cu:2428-2655)
class TuneBirkhoffConstruct {
  public:
    TuneBirkhoffConstruct(size_t N, std::vector<size_t> ts) : N(N), ts(ts), sw(ts), sx(N * 8), sy(N * 8)
    {
        require_device();
        const size_t m = N * 8 * std::max<size_t>(ts.size(), 1);
        t1x.alloc(m); t2x.alloc(m); t1y.alloc(m); t2y.alloc(m);
        reset();
    }
    void reset()
    {
        idx = 1;
        const std::vector<double> z(N * std::max<size_t>(ts.size(), 1), 0.0);
        for (DevBuf *b : {&t1x, &t2x, &t1y, &t2y}) b->upload(z);
    }
    void first_add(const Particles4D &p) { call(p, 1); }
    void add(const Particles4D &p)
    {
        call(p, 0);
        idx++;
    }
    py::array_t<double> get(const DevBuf &b) const { return to_array(b, {(py::ssize_t)ts.size(), (py::ssize_t)N}); }
    size_t N, idx = 1;
    std::vector<size_t> ts;
    SampleWeights sw;
    DevBuf sx, sy, t1x, t2x, t1y, t2y;

  private:
    void call(const Particles4D &p, int init)
    {
        if (p.n_particles != N) throw std::invalid_argument("tune_birkhoff_construct: size mismatch");
        check(hm_tune_birkhoff_add(p.dx.as<double>(), p.dpx.as<double>(), p.dy.as<double>(), p.dpy.as<double>(),
                                   sx.as<double>(), sy.as<double>(), t1x.as<double>(), t2x.as<double>(), t1y.as<double>(),
                                   t2y.as<double>(), sw.w.as<double>(), sw.o.as<uint64_t>(), sw.s.as<uint64_t>(),
                                   (uint32_t)ts.size(), N, idx, init, nullptr));
        sync();
    }
};

// storage_4d (src/henon.cu:3221-3348): a growing trajectory store on the device
// (samples x 4 x N), with FFT and Birkhoff tunes of windows plus the full signal.
class Storage4D {
  public:
    explicit Storage4D(size_t N) : N(N) { require_device(); }
    void store(const Particles4D &p)
    {
        if (p.n_particles != N) throw std::invalid_argument("store: size mismatch");
        if ((T + 1) * 32 * N > h.bytes()) {
            DevBuf nh(std::max<size_t>(2 * h.bytes(), 64 * 32 * N));
            if (T) check(hm_memcpy(nh.as<void>(), h.as<void>(), T * 32 * N, 2, nullptr));
            sync();
            h = std::move(nh);
        }
        double *base = h.as<double>() + T * 4 * N;
        check(hm_memcpy(base, p.dx.as<void>(), N * 8, 2, nullptr));
        check(hm_memcpy(base + N, p.dpx.as<void>(), N * 8, 2, nullptr));
        check(hm_memcpy(base + 2 * N, p.dy.as<void>(), N * 8, 2, nullptr));
        check(hm_memcpy(base + 3 * N, p.dpy.as<void>(), N * 8, 2, nullptr));
        sync();
        T++;
    }
    py::array_t<double> tunes(std::vector<unsigned int> from, std::vector<unsigned int> to, bool fft) const
    {
        if (from.size() != to.size()) throw std::runtime_error("compute_tune: from and to must be of the same size");
        for (size_t i = 0; i < from.size(); i++)
            if (from[i] > to[i]) throw std::invalid_argument("cpu_henon::fft_tunes: from[i] > to[i]");
        if (T < 2) throw std::runtime_error("tunes: at least two stored samples are needed");
        std::vector<uint32_t> f(from.begin(), from.end()), t(to.begin(), to.end()), len;
        f.push_back(fft ? 1 : 0);  // full signal: samples 1..T-1 for the FFT, all for Birkhoff
        t.push_back((uint32_t)T - 1);
        for (size_t i = 0; i < f.size(); i++) len.push_back(t[i] - f[i] + 1);
        for (size_t i = 0; i + 1 < f.size(); i++)
            if (f[i] >= T || t[i] >= T) throw std::runtime_error("compute_tune: from and to must be valid indices");
        const size_t W = f.size();
        DevBuf o(N * 2 * W * 8);
        if (fft) check(hm_fft_tunes(h.as<double>(), N, (uint32_t)T, nullptr, f.data(), len.data(), W, o.as<double>(), nullptr));
        else check(hm_birkhoff_tunes_history(h.as<double>(), N, (uint32_t)T, nullptr, f.data(), t.data(), W, o.as<double>(), nullptr));
        auto v = o.download<double>(N * 2 * W);
        py::array_t<double> r({(py::ssize_t)N, (py::ssize_t)(2 * W)});
        auto rr = r.mutable_unchecked<2>();
        for (size_t w = 0; w < W; w++)
            for (int pl = 0; pl < 2; pl++)
                for (size_t i = 0; i < N; i++) rr(i, 2 * w + pl) = v[(w * 2 + pl) * N + i];
        return r;
    }
    // (N, T): particle-major like the reference's vector<vector> per particle
    py::array_t<double> coord(int c) const
    {
        std::vector<double> all = h.download<double>(T * 4 * N);
        py::array_t<double> r({(py::ssize_t)N, (py::ssize_t)T});
        auto rr = r.mutable_unchecked<2>();
        for (size_t m = 0; m < T; m++)
            for (size_t i = 0; i < N; i++) rr(i, m) = all[m * 4 * N + c * N + i];
        return r;
    }
    size_t N, T = 0;
    DevBuf h;
};

// storage_4d_gpu (src/henon.cu:3351-3463): fixed capacity, (batch, N) readback
class Storage4DGpu {
  public:
    Storage4DGpu(size_t N, size_t batch) : N(N), batch(batch), h(batch * 32 * N) { require_device(); reset(); }
    void store(const Particles4D &p)
    {
        if (p.n_particles != N) throw std::invalid_argument("store: size mismatch");
        if (idx >= batch) throw std::runtime_error("storage_4d_gpu: the batch is full");
        const double *src[4] = {p.dx.as<double>(), p.dpx.as<double>(), p.dy.as<double>(), p.dpy.as<double>()};
        for (int c = 0; c < 4; c++) check(hm_memcpy(h.as<double>() + (c * batch + idx) * N, src[c], N * 8, 2, nullptr));
        sync();
        idx++;
    }
    void reset()
    {
        idx = 0;
        check(hm_memset(h.as<void>(), 0, batch * 32 * N, nullptr));
        sync();
    }
    py::array_t<double> coord(int c) const
    {
        DevBuf tmp(batch * N * 8);
        check(hm_memcpy(tmp.as<void>(), h.as<double>() + c * batch * N, batch * N * 8, 2, nullptr));
        return to_array(tmp, {(py::ssize_t)batch, (py::ssize_t)N});
    }
    size_t N, batch, idx = 0;
    DevBuf h;
};

// get_tunes (src/dynamic_indicator.cpp:333-361): {fft, birkhoff} of one signal
// Many independent (tracker, particles, n_turns, mu, barrier) jobs in one
// batched launch sequence (hm_track_batch); each job's particles advance as
// with tracker.track(particles, n_turns, mu, barrier, inverse=False).
static std::vector<hm_job> batch_jobs(const py::list &jobs, std::vector<std::pair<Particles4D *, unsigned int>> &adv)
{
    std::vector<hm_job> js;
    std::unordered_set<const Particles4D *> seen;
    for (const py::handle &h : jobs) {
        const py::tuple t = h.cast<py::tuple>();
        if (t.size() != 5) throw std::invalid_argument("track_batch: jobs are (tracker, particles, n_turns, mu, barrier)");
        HenonTracker &tr = t[0].cast<HenonTracker &>();
        Particles4D &p = t[1].cast<Particles4D &>();
        const unsigned int n_turns = t[2].cast<unsigned int>();
        if (!seen.insert(&p).second) throw std::invalid_argument("track_batch: a particle set appears in two jobs");
        js.push_back(tr.batch_job(p, n_turns, t[3].cast<double>(), t[4].cast<double>()));
        adv.emplace_back(&p, n_turns);
    }
    return js;
}

// Batch autotuner (hm_autotune_batch): times launch configurations on copies of
// the jobs' particles, each job cut to trial_turns turns, and makes later
// track_batch calls of the same job-mix class use the fastest. The particles
// are not advanced.
py::dict autotune_batch(const py::list &jobs, const std::string &mode, unsigned int trial_turns)
{
    std::vector<std::pair<Particles4D *, unsigned int>> adv;
    const std::vector<hm_job> js = batch_jobs(jobs, adv);
    hm_batch_cfg best{};
    double ms = 0.0;
    {
        py::gil_scoped_release nogil;
        check(hm_autotune_batch(js.data(), js.size(), mode_of(mode), trial_turns, &best, &ms, nullptr));
    }
    py::dict d;
    d["block"] = best.block;
    d["compact_every"] = best.compact_every;
    d["waves"] = best.waves;
    d["trial_ms"] = ms;
    return d;
}

void track_batch(const py::list &jobs, const std::string &mode)
{
    std::vector<std::pair<Particles4D *, unsigned int>> adv;
    const std::vector<hm_job> js = batch_jobs(jobs, adv);
    const int m = mode_of(mode);
    {
        py::gil_scoped_release nogil;
        check(hm_track_batch(js.data(), js.size(), m, nullptr, nullptr));
        uint32_t flags = 0;
        check(hm_take_status(&flags, nullptr));
        if (flags & 1u) throw std::runtime_error("track_batch: a particle's turn index left its modulation table");
    }
    for (auto &a : adv) a.first->global_steps += a.second;
}

std::vector<double> get_tunes(const std::vector<double> &x, const std::vector<double> &px)
{
    if (x.size() != px.size()) throw std::runtime_error("get_tunes: x and px must be of the same size");
    const double qn = std::numeric_limits<double>::quiet_NaN();
    for (size_t i = 0; i < x.size(); i++)
        if (std::isnan(x[i]) || std::isnan(px[i])) return {qn, qn};
    require_device();
    const size_t L = x.size();
    if (L < 2) throw std::runtime_error("get_tunes: at least two samples are needed");
    std::vector<double> hist(L * 4);
    for (size_t m = 0; m < L; m++) {
        hist[4 * m] = x[m]; hist[4 * m + 1] = px[m]; hist[4 * m + 2] = x[m]; hist[4 * m + 3] = px[m];
    }
    DevBuf h(L * 32), o(16), b(16);
    h.upload(hist);
    uint32_t f = 0, len = (uint32_t)L, t = (uint32_t)L - 1;
    check(hm_fft_tunes(h.as<double>(), 1, (uint32_t)L, nullptr, &f, &len, 1, o.as<double>(), nullptr));
    check(hm_birkhoff_tunes_history(h.as<double>(), 1, (uint32_t)L, nullptr, &f, &t, 1, b.as<double>(), nullptr));
    return {o.download<double>(1)[0], b.download<double>(1)[0]};
}

// ---------------------------------------------------------------- trampolines
// src/bind.cpp:43-222: particles and trackers stay subclassable from Python.
struct PyParticles4D : Particles4D {
    using Particles4D::Particles4D;
    void reset() override { PYBIND11_OVERRIDE(void, Particles4D, reset, ); }
};

struct PyHenonTracker : HenonTracker {
    using HenonTracker::HenonTracker;
    void compute_a_modulation(unsigned int N, double omega_x, double omega_y, std::string modulation_kind,
                              double omega_0, double epsilon, unsigned int offset) override
    {
        PYBIND11_OVERRIDE(void, HenonTracker, compute_a_modulation, N, omega_x, omega_y, modulation_kind, omega_0,
                          epsilon, offset);
    }
};

// The arrays the reference's `py::array out = py::cast(vector)` gives
// (src/bind.cpp:256-275): float64 for double vectors, int64 for the integer ones
// (numpy's type for a list of Python ints), float64 when empty; filled directly
// instead of through a Python list.
template <class T> py::array as_array(const std::vector<T> &v)
{
    using Out = std::conditional_t<std::is_floating_point<T>::value, double, int64_t>;
    if (v.empty()) return py::array_t<double>(0);
    py::array_t<Out> a((py::ssize_t)v.size());
    Out *o = a.mutable_data();
    if (std::is_same<T, Out>::value) std::memcpy(o, v.data(), v.size() * sizeof(T));
    else for (size_t i = 0; i < v.size(); ++i) o[i] = static_cast<Out>(v[i]);
    return a;
}

}  // namespace

PYBIND11_MODULE(henon_map_engine, m)
{
    m.doc() = "Henon map engine (MI355X / gfx950, HIP)";
    py::register_exception<NotImplemented>(m, "NotAvailableError", PyExc_NotImplementedError);

    m.def("is_cuda_device_available", [] {
        int c = 0;
        return hm_device_count(&c) == HM_OK && c > 0;
    });
    m.def("is_hip_device_available", [] {
        int c = 0;
        return hm_device_count(&c) == HM_OK && c > 0;
    });
    m.def("basic_modulation", &basic_modulation, "Basic modulation", py::arg("tune"), py::arg("omega"),
          py::arg("epsilon"), py::arg("start"), py::arg("end"));
    m.def("sps_modulation", &sps_modulation, "SPS modulation", py::arg("tune"), py::arg("epsilon"), py::arg("start"),
          py::arg("end"));
    m.def("gaussian_modulation", &gaussian_modulation, "Gaussian modulation", py::arg("tune"), py::arg("sigma"),
          py::arg("start"), py::arg("end"), py::arg("seed") = py::none());
    m.def("uniform_modulation", &uniform_modulation, "Uniform modulation", py::arg("from"), py::arg("to"),
          py::arg("start"), py::arg("end"), py::arg("seed") = py::none());
    m.def("lines_modulation", &lines_modulation, "Line-spectrum modulation (the sps loop with given lines)",
          py::arg("tune"), py::arg("epsilon"), py::arg("coefficients"), py::arg("frequencies"), py::arg("start"),
          py::arg("end"));
    m.def("birkhoff_weights", &birkhoff_weights, "Birkhoff weights", py::arg("n_weights"));

    py::class_<Particles4D, PyParticles4D>(m, "particles_4d")
        .def(py::init<const Coords &, const Coords &, const Coords &, const Coords &>())
        .def(py::init<const std::vector<double> &, const std::vector<double> &, const std::vector<double> &,
                      const std::vector<double> &>())
        .def("reset", &Particles4D::reset)
        .def("add_ghost", &Particles4D::add_ghost, py::arg("displacement_module"), py::arg("direction"))
        .def("renormalize", &Particles4D::renormalize, py::arg("module_target"))
        .def("get_displacement_module", &Particles4D::get_displacement_module)
        .def("get_displacement_direction", &Particles4D::get_displacement_direction)
        .def("set_kick_seed", &Particles4D::set_kick_seed, py::arg("seed"))
        .def("get_sali", [](Particles4D &s) { return as_array(s.alignment_index(HM_ALIGN_SALI)); })
        .def("get_gali", [](Particles4D &s) { return as_array(s.alignment_index(HM_ALIGN_GALI)); })
        .def("get_x", [](Particles4D &s) { return as_array(s.get_x()); })
        .def("get_px", [](Particles4D &s) { return as_array(s.get_px()); })
        .def("get_y", [](Particles4D &s) { return as_array(s.get_y()); })
        .def("get_py", [](Particles4D &s) { return as_array(s.get_py()); })
        .def("get_steps", [](Particles4D &s) { return as_array(s.get_steps()); })
        .def("get_radius", [](Particles4D &s) { return as_array(s.get_radius()); })
        .def("get_radius_mean", [](Particles4D &s) { return py::array(py::cast(s.get_radius_mean())); })
        .def("get_radius_std", [](Particles4D &s) { return py::array(py::cast(s.get_radius_std())); })
        .def("get_action", [](Particles4D &s) { return as_array(s.get_action()); })
        .def("get_action_mean", [](Particles4D &s) { return py::array(py::cast(s.get_action_mean())); })
        .def("get_action_std", [](Particles4D &s) { return py::array(py::cast(s.get_action_std())); })
        .def("get_action_x", [](Particles4D &s) { return as_array(s.get_action_x()); })
        .def("get_action_x_mean", [](Particles4D &s) { return py::array(py::cast(s.get_action_x_mean())); })
        .def("get_action_x_std", [](Particles4D &s) { return py::array(py::cast(s.get_action_x_std())); })
        .def("get_action_y", [](Particles4D &s) { return as_array(s.get_action_y()); })
        .def("get_action_y_mean", [](Particles4D &s) { return py::array(py::cast(s.get_action_y_mean())); })
        .def("get_action_y_std", [](Particles4D &s) { return py::array(py::cast(s.get_action_y_std())); })
        .def("get_angles_x", [](Particles4D &s) { return as_array(s.get_angles_x()); })
        .def("get_angles_x_mean", [](Particles4D &s) { return py::array(py::cast(s.get_angles_x_mean())); })
        .def("get_angles_x_std", [](Particles4D &s) { return py::array(py::cast(s.get_angles_x_std())); })
        .def("get_angles_y", [](Particles4D &s) { return as_array(s.get_angles_y()); })
        .def("get_angles_y_mean", [](Particles4D &s) { return py::array(py::cast(s.get_angles_y_mean())); })
        .def("get_angles_y_std", [](Particles4D &s) { return py::array(py::cast(s.get_angles_y_std())); })
        .def("get_valid", [](Particles4D &s) { return as_array(s.get_valid()); })
        .def("get_ghost", [](Particles4D &s) { return as_array(s.get_ghost()); })
        .def("get_idx", [](Particles4D &s) { return as_array(s.get_idx()); })
        .def("get_idx_base", [](Particles4D &s) { return as_array(s.get_idx_base()); })
        .def("get_n_particles", [](Particles4D &s) { return py::array(py::cast(s.get_n_particles())); })
        .def("get_n_ghosts_per_particle", [](Particles4D &s) { return py::array(py::cast(s.get_n_ghosts_per_particle())); })
        .def_readonly("global_steps", &Particles4D::global_steps);

    py::class_<Particles4DGpu, Particles4D>(m, "particles_4d_gpu")
        .def(py::init<const Coords &, const Coords &, const Coords &, const Coords &>())
        .def(py::init<const std::vector<double> &, const std::vector<double> &, const std::vector<double> &,
                      const std::vector<double> &>());

    auto tracker = [&](auto &cls) {
        using C = typename std::decay_t<decltype(cls)>::type;
        cls.def("compute_a_modulation", &C::compute_a_modulation, "Compute a modulation", py::arg("N"),
                py::arg("omega_x"), py::arg("omega_y"), py::arg("modulation_kind"), py::arg("omega_0"),
                py::arg("epsilon"), py::arg("offset"))
            .def("autotune", &C::autotune, "Pick the fastest launch configuration for these particles",
                 py::arg("particles"), py::arg("mu"), py::arg("barrier") = 100.0, py::arg("mode") = "exact",
                 py::arg("warm_turns") = 2000, py::arg("trial_turns") = 4000)
            .def("track", &C::track, "Track particles", py::arg("particles"), py::arg("n_turns"), py::arg("mu"),
                 py::arg("barrier") = 100.0, py::arg("kick_module") = std::numeric_limits<double>::quiet_NaN(),
                 py::arg("inverse") = false, py::arg("mode") = "exact")
            .def("get_tangent_matrix", &C::get_tangent_matrix, "Get tangent matrix", py::arg("particles"),
                 py::arg("mu"), py::arg("reverse"))
            .def("get_omega_x", [](C &s) { return as_array(s.get_omega_x()); })
            .def("get_omega_y", [](C &s) { return as_array(s.get_omega_y()); })
            .def_readonly("allowed_steps", &C::allowed_steps);
    };

    py::class_<HenonTracker, PyHenonTracker> ht(m, "henon_tracker");
    ht.def(py::init<unsigned int, double, double, std::string, double, double, unsigned int, std::vector<double>,
                    std::vector<double>>(),
           py::arg("N"), py::arg("omega_x"), py::arg("omega_y"), py::arg("modulation_kind"), py::arg("omega_0"),
           py::arg("epsilon"), py::arg("offset"), py::arg("coefficients") = std::vector<double>(),
           py::arg("frequencies") = std::vector<double>());
    tracker(ht);

    py::class_<HenonTrackerGpu, HenonTracker> htg(m, "henon_tracker_gpu");
    htg.def(py::init<unsigned int, double, double, std::string, double, double, unsigned int, std::vector<double>,
                     std::vector<double>>(),
            py::arg("N"), py::arg("omega_x"), py::arg("omega_y"), py::arg("modulation_kind"), py::arg("omega_0"),
            py::arg("epsilon"), py::arg("offset"), py::arg("coefficients") = std::vector<double>(),
            py::arg("frequencies") = std::vector<double>());

    m.def("get_tunes", &get_tunes, "Get tunes", py::arg("x"), py::arg("px"));
    m.def("autotune_batch", &autotune_batch, "Pick the fastest batch launch configuration for this job mix",
          py::arg("jobs"), py::arg("mode") = "exact", py::arg("trial_turns") = 4000);
    m.def("track_batch", &track_batch, "Track many (tracker, particles, n_turns, mu, barrier) jobs in one batch",
          py::arg("jobs"), py::arg("mode") = "exact");

    auto tracker_more = [&](auto &cls) {
        using C = typename std::decay_t<decltype(cls)>::type;
        cls.def("megno", &C::megno, "Compute megno", py::arg("particles"), py::arg("n_turns"), py::arg("mu"),
                py::arg("barrier") = 100.0, py::arg("kick_module") = std::numeric_limits<double>::quiet_NaN(),
                py::arg("inverse") = false, py::arg("turn_samples") = std::vector<unsigned int>(),
                py::arg("n_threads") = -1, py::arg("carry_off_diagonal") = true)
            .def("birkhoff_tunes",
                 [](C &s, Particles4D &p, unsigned n, double mu, double barrier, double kick, bool inverse,
                    std::vector<unsigned> f, std::vector<unsigned> t) {
                     return s.history_tunes(p, n, mu, barrier, kick, inverse, f, t, false);
                 },
                 "Compute birkhoff tunes", py::arg("particles"), py::arg("n_turns"), py::arg("mu"),
                 py::arg("barrier") = 100.0, py::arg("kick_module") = std::numeric_limits<double>::quiet_NaN(),
                 py::arg("inverse") = false, py::arg("from_idx") = std::vector<unsigned int>(),
                 py::arg("to_idx") = std::vector<unsigned int>())
            .def("all_tunes",
                 [](C &s, Particles4D &p, unsigned n, double mu, double barrier, double kick, bool inverse,
                    std::vector<unsigned> f, std::vector<unsigned> t) {
                     return s.history_tunes(p, n, mu, barrier, kick, inverse, f, t, true);
                 },
                 "Compute birkhoff and FFT tunes", py::arg("particles"), py::arg("n_turns"), py::arg("mu"),
                 py::arg("barrier") = 100.0, py::arg("kick_module") = std::numeric_limits<double>::quiet_NaN(),
                 py::arg("inverse") = false, py::arg("from_idx") = std::vector<unsigned int>(),
                 py::arg("to_idx") = std::vector<unsigned int>())
            .def("tangent_indicators", &C::tangent_indicators, "Fused tangent map, MEGNO, FLI, Birkhoff tunes",
                 py::arg("particles"), py::arg("n_turns"), py::arg("mu"), py::arg("barrier") = 100.0,
                 py::arg("renorm_every") = 100u, py::arg("samples") = std::vector<unsigned int>(),
                 py::arg("fli_direction") = std::vector<double>{1.0, 0.0, 0.0, 0.0},
                 py::arg("birkhoff") = std::vector<std::pair<unsigned, unsigned>>(),
                 py::arg("history") = std::pair<unsigned, unsigned>(0u, 0u), py::arg("mode") = "exact",
                 py::arg("megno_birkhoff") = std::vector<unsigned int>(),
                 py::arg("tune_birkhoff") = std::vector<unsigned int>())
            .def("reversibility_error", &C::reversibility_error, "Reversibility error", py::arg("particles"),
                 py::arg("t_list"), py::arg("mu"), py::arg("barrier") = 100.0, py::arg("mode") = "exact");
    };
    tracker_more(ht);

    py::class_<Matrix4DVector>(m, "matrix_4d_vector")
        .def(py::init<size_t>())
        .def("reset", &Matrix4DVector::reset)
        .def("multiply", &Matrix4DVector::multiply, py::arg("matrix"))
        .def("structured_multiply", &Matrix4DVector::structured_multiply_rev, py::arg("tracker"), py::arg("particles"),
             py::arg("mu"), py::arg("reverse"))
        .def("get_matrix", &Matrix4DVector::get_matrix)
        .def("get_vector", &Matrix4DVector::get_vector, py::arg("vector"));

    py::class_<Matrix4DVectorGpu, Matrix4DVector>(m, "matrix_4d_vector_gpu")
        .def(py::init<size_t>())
        .def("reset", &Matrix4DVector::reset)
        .def("structured_multiply", &Matrix4DVector::structured_multiply, py::arg("tracker"), py::arg("particles"),
             py::arg("mu"))
        .def("set_with_tracker", &Matrix4DVector::set_with_tracker, py::arg("tracker"), py::arg("particles"),
             py::arg("mu"))
        .def("explicit_copy", &Matrix4DVector::explicit_copy, py::arg("other"))
        .def("get_matrix", &Matrix4DVector::get_matrix)
        .def("get_vector", &Matrix4DVector::get_vector, py::arg("vector"));

    py::class_<Vector4D>(m, "vector_4d_gpu")
        .def(py::init<const Coords &>())
        .def(py::init<const std::vector<std::vector<double>> &>())
        .def("set_vectors", &Vector4D::set_array, py::arg("vectors"))
        .def("set_vectors", &Vector4D::set_vectors, py::arg("vectors"))
        .def("set_vectors", &Vector4D::set_vector, py::arg("vectors"))
        .def("multiply", &Vector4D::multiply, py::arg("matrix"))
        .def("normalize", &Vector4D::normalize)
        .def("get_vectors", &Vector4D::get_vectors);

    py::class_<LyapunovConstruct>(m, "lyapunov_birkhoff_construct")
        .def(py::init<size_t, size_t>())
        .def("reset", &LyapunovConstruct::reset)
        .def("change_weights", &LyapunovConstruct::change_weights, py::arg("n_weights"))
        .def("add", &LyapunovConstruct::add, py::arg("vectors"))
        .def("get_weights", &LyapunovConstruct::get_weights)
        .def("get_values_raw", &LyapunovConstruct::get_values_raw)
        .def("get_values_b", &LyapunovConstruct::get_values_b);

    py::class_<LyapunovConstructMulti>(m, "lyapunov_birkhoff_construct_multi")
        .def(py::init<size_t, std::vector<size_t>>())
        .def("reset", &LyapunovConstructMulti::reset)
        .def("add", &LyapunovConstructMulti::add, py::arg("vectors"))
        .def("get_values_raw", [](LyapunovConstructMulti &s) { return py::array(s.get_values_raw()); })
        .def("get_values_b", [](LyapunovConstructMulti &s) { return py::array(s.get_values_b()); });

    py::class_<MegnoConstruct>(m, "megno_construct")
        .def(py::init<size_t>())
        .def("reset", &MegnoConstruct::reset)
        .def("add", &MegnoConstruct::add, py::arg("matrix_a"), py::arg("matrix_b"))
        .def("get_values", &MegnoConstruct::get_values);

    py::class_<MegnoBirkhoffConstruct>(m, "megno_birkhoff_construct")
        .def(py::init<size_t, std::vector<size_t>>())
        .def("reset", &MegnoBirkhoffConstruct::reset)
        .def("add", &MegnoBirkhoffConstruct::add, py::arg("matrix_a"), py::arg("matrix_b"))
        .def("get_values", &MegnoBirkhoffConstruct::get_values);

    py::class_<TuneBirkhoffConstruct>(m, "tune_birkhoff_construct")
        .def(py::init<size_t, std::vector<size_t>>())
        .def("reset", &TuneBirkhoffConstruct::reset)
        .def("first_add", &TuneBirkhoffConstruct::first_add, py::arg("particles"))
        .def("add", &TuneBirkhoffConstruct::add, py::arg("particles"))
        .def("get_tune1_x", [](TuneBirkhoffConstruct &s) { return s.get(s.t1x); })
        .def("get_tune1_y", [](TuneBirkhoffConstruct &s) { return s.get(s.t1y); })
        .def("get_tune2_x", [](TuneBirkhoffConstruct &s) { return s.get(s.t2x); })
        .def("get_tune2_y", [](TuneBirkhoffConstruct &s) { return s.get(s.t2y); });

    py::class_<Storage4D>(m, "storage_4d")
        .def(py::init<size_t>())
        .def("store", &Storage4D::store, "Store particles", py::arg("particles"))
        .def("tune_fft", [](Storage4D &s, std::vector<unsigned> f, std::vector<unsigned> t) { return s.tunes(f, t, true); },
             "Tune FFT", py::arg("from_idx"), py::arg("to_idx"))
        .def("tune_birkhoff",
             [](Storage4D &s, std::vector<unsigned> f, std::vector<unsigned> t) { return s.tunes(f, t, false); },
             "Tune birkhoff", py::arg("from_idx"), py::arg("to_idx"))
        .def("get_x", [](Storage4D &s) { return s.coord(0); })
        .def("get_px", [](Storage4D &s) { return s.coord(1); })
        .def("get_y", [](Storage4D &s) { return s.coord(2); })
        .def("get_py", [](Storage4D &s) { return s.coord(3); });

    py::class_<Storage4DGpu>(m, "storage_4d_gpu")
        .def(py::init<size_t, size_t>())
        .def("store", &Storage4DGpu::store, "Store particles", py::arg("particles"))
        .def("reset", &Storage4DGpu::reset, "Reset storage")
        .def("get_x", [](Storage4DGpu &s) { return s.coord(0); })
        .def("get_px", [](Storage4DGpu &s) { return s.coord(1); })
        .def("get_y", [](Storage4DGpu &s) { return s.coord(2); })
        .def("get_py", [](Storage4DGpu &s) { return s.coord(3); });
}
