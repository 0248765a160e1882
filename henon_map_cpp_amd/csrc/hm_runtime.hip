// Host runtime of the library: error state, memory plumbing, instrumentation,
// and the modulation tables (generated on the host with glibc libm, so that the
// tables are bit-identical to the reference's src/modulation.cpp output).
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <mutex>
#include <map>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "hm_internal.h"

namespace hm {

static thread_local std::string g_last_error = "";

void set_error(const std::string &msg) { g_last_error = msg; }

// ---------------------------------------------------------------- profiling
namespace {
struct PendingLaunch {
    std::string name;
    hipEvent_t start, stop;
};
struct ProfileTotals {
    double ms = 0.0;
    uint64_t launches = 0;
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::vector<PendingLaunch> g_pending;
std::map<std::string, ProfileTotals> g_totals;

void drain_pending_locked()
{
    for (auto &p : g_pending) {
        float ms = 0.f;
        if (hipEventSynchronize(p.stop) == hipSuccess && hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
            auto &t = g_totals[p.name];
            t.ms += ms;
            t.launches += 1;
        }
        (void)hipEventDestroy(p.start);
        (void)hipEventDestroy(p.stop);
    }
    g_pending.clear();
}
}  // namespace

LaunchTimer::LaunchTimer(const char *name, hipStream_t s) : name_(name), stream_(s)
{
    bool on;
    {
        std::lock_guard<std::mutex> lk(g_prof_mu);
        on = g_prof_on;
    }
    if (on) {
        hip_check(hipEventCreate(&start_), "hipEventCreate");
        hip_check(hipEventRecord(start_, stream_), "hipEventRecord");
    }
}

LaunchTimer::~LaunchTimer()
{
    if (!start_) return;
    hipEvent_t stop;
    if (hipEventCreate(&stop) != hipSuccess) return;
    (void)hipEventRecord(stop, stream_);
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_pending.push_back({name_, start_, stop});
}

// ---------------------------------------------------------------- device words
namespace {
std::mutex g_ws_mu;
struct DeviceScratch {
    void *slots[kWorkspaceSlots] = {};
    size_t sizes[kWorkspaceSlots] = {};
};
std::map<int, DeviceScratch> g_scratch;

// The status flags word lives in static device memory: each device holds its
// own zero-initialised copy from module load, so no call allocates for it.
// Host code uses the address hipGetSymbolAddress gives on each device.
__device__ __align__(256) uint32_t g_status_word[64];
constexpr int kMaxDevices = 64;
std::mutex g_status_mu;
uint32_t *g_status_ptr[kMaxDevices] = {};
}  // namespace

uint32_t *status_word()
{
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) throw Error(HM_ERR_ARG, "device index beyond the library's limit of 64");
    std::lock_guard<std::mutex> lk(g_status_mu);
    if (!g_status_ptr[dev]) {
        void *p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(g_status_word)), "hipGetSymbolAddress(status)");
        g_status_ptr[dev] = static_cast<uint32_t *>(p);
    }
    return g_status_ptr[dev];
}

uint32_t device_simds()
{
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    static std::mutex mu;
    static std::map<int, uint32_t> simds;
    std::lock_guard<std::mutex> lk(mu);
    auto it = simds.find(dev);
    if (it != simds.end()) return it->second;
    int cus = 0;
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    return simds[dev] = (uint32_t)cus * 4u;
}

void *workspace(size_t bytes, int slot)
{
    if (slot < 0 || slot >= kWorkspaceSlots) throw Error(HM_ERR_ARG, "workspace slot out of range");
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lk(g_ws_mu);
    auto &d = g_scratch[dev];
    if (d.sizes[slot] < bytes) {
        if (d.slots[slot]) {
            hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
            hip_check(hipFree(d.slots[slot]), "hipFree(workspace)");
            d.slots[slot] = nullptr;
            d.sizes[slot] = 0;
        }
        size_t sz = bytes < 4096 ? 4096 : bytes;
        hip_check(hipMalloc(&d.slots[slot], sz), "hipMalloc(workspace)");
        d.sizes[slot] = sz;
    }
    return d.slots[slot];
}

// ---------------------------------------------------------------- modulation
// src/modulation.cpp:14-28
static const double kSpsCoef[7] = {1.000e-4, 0.218e-4, 0.708e-4, 0.254e-4, 0.100e-4, 0.078e-4, 0.218e-4};
static const double kSpsFreq[7] = {1 * (2 * M_PI / 868.12), 2 * (2 * M_PI / 868.12), 3 * (2 * M_PI / 868.12),
                                   6 * (2 * M_PI / 868.12), 7 * (2 * M_PI / 868.12), 10 * (2 * M_PI / 868.12),
                                   12 * (2 * M_PI / 868.12)};

// Splits [0, n) over the host's threads, strided like the reference's
// modulation and table loops (src/modulation.cpp:34-58, src/henon.cu:2670-2692).
template <class F>
static void host_parallel(size_t n, F fn)
{
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (n < 65536) nt = 1;
    if (nt == 1) {
        for (size_t i = 0; i < n; i++) fn(i);
        return;
    }
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++)
        th.emplace_back([&, t]() { for (size_t i = t; i < n; i += nt) fn(i); });
    for (auto &t : th) t.join();
}

static void lines(double tune, double epsilon, const double *coef, const double *freq, int n_lines, int start,
                  int end, double *out)
{
    if (!(start < end)) throw Error(HM_ERR_ARG, "modulation: start must be smaller than end");
    if (n_lines < 0 || (n_lines > 0 && (!coef || !freq))) throw Error(HM_ERR_ARG, "modulation: bad line set");
    host_parallel((size_t)(end - start), [&](size_t ii) {
        const unsigned int i = (unsigned int)ii;
        auto sum = 0.0;
        for (int j = 0; j < n_lines; j++) sum += coef[j] * std::cos(freq[j] * (i + start));
        out[i] = (2 * M_PI * tune * (1 + epsilon * sum));
    });
}

}  // namespace hm

using namespace hm;

extern "C" {

const char *hm_last_error(void) { return g_last_error.c_str(); }

const char *hm_version(void) { return "henon_map_cpp_amd 0.1 (gfx950)"; }

int hm_device_count(int *count)
{
    return guarded([&] {
        if (!count) throw Error(HM_ERR_ARG, "count is NULL");
        int c = 0;
        if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
        (void)hipGetLastError();
        *count = c;
    });
}

int hm_mem_info(size_t *free_bytes, size_t *total_bytes)
{
    return guarded([&] {
        if (!free_bytes || !total_bytes) throw Error(HM_ERR_ARG, "hm_mem_info: NULL pointer");
        hip_check(hipMemGetInfo(free_bytes, total_bytes), "hipMemGetInfo");
    });
}

int hm_malloc(void **ptr, size_t bytes)
{
    return guarded([&] {
        if (!ptr) throw Error(HM_ERR_ARG, "ptr is NULL");
        *ptr = nullptr;
        if (bytes == 0) return;
        hipError_t e = hipMalloc(ptr, bytes);
        if (e != hipSuccess) throw Error(HM_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
    });
}

int hm_free(void *ptr)
{
    return guarded([&] {
        if (ptr) hip_check(hipFree(ptr), "hipFree");
    });
}

int hm_memcpy(void *dst, const void *src, size_t bytes, int kind, void *stream)
{
    return guarded([&] {
        if (bytes == 0) return;
        if (!dst || !src) throw Error(HM_ERR_ARG, "hm_memcpy: NULL pointer");
        hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost
                                                            : hipMemcpyDeviceToDevice;
        if (kind < 0 || kind > 2) throw Error(HM_ERR_ARG, "hm_memcpy: bad kind");
        hip_check(hipMemcpyAsync(dst, src, bytes, k, as_stream(stream)), "hipMemcpyAsync");
    });
}

int hm_memset(void *dst, int value, size_t bytes, void *stream)
{
    return guarded([&] {
        if (bytes == 0) return;
        if (!dst) throw Error(HM_ERR_ARG, "hm_memset: NULL pointer");
        hip_check(hipMemsetAsync(dst, value, bytes, as_stream(stream)), "hipMemsetAsync");
    });
}

int hm_stream_synchronize(void *stream)
{
    return guarded([&] { hip_check(hipStreamSynchronize(as_stream(stream)), "hipStreamSynchronize"); });
}

int hm_take_status(uint32_t *flags, void *stream)
{
    return guarded([&] {
        if (!flags) throw Error(HM_ERR_ARG, "flags is NULL");
        uint32_t *w = status_word();
        hipStream_t s = as_stream(stream);
        hip_check(hipMemcpyAsync(flags, w, sizeof(uint32_t), hipMemcpyDeviceToHost, s), "hipMemcpyAsync(status)");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        hip_check(hipMemsetAsync(w, 0, sizeof(uint32_t), s), "hipMemsetAsync(status)");
    });
}

int hm_basic_modulation(double tune, double omega, double epsilon, int start, int end, double *out)
{
    // src/modulation.cpp:3-12: index i, not i + start (kept)
    return guarded([&] {
        if (!(start < end)) throw Error(HM_ERR_ARG, "modulation: start must be smaller than end");
        if (!out) throw Error(HM_ERR_ARG, "out is NULL");
        for (int i = 0; i < end - start; i++)
            out[i] = (2 * M_PI * tune * (1 + epsilon * std::cos(2 * M_PI * omega * i)));
    });
}

int hm_sps_modulation(double tune, double epsilon, int start, int end, double *out)
{
    return guarded([&] {
        if (!out) throw Error(HM_ERR_ARG, "out is NULL");
        lines(tune, epsilon, kSpsCoef, kSpsFreq, 7, start, end, out);
    });
}

int hm_lines_modulation(double tune, double epsilon, const double *coef, const double *freq, int n_lines,
                        int start, int end, double *out)
{
    return guarded([&] {
        if (!out) throw Error(HM_ERR_ARG, "out is NULL");
        lines(tune, epsilon, coef, freq, n_lines, start, end, out);
    });
}

int hm_sps_lines(double *coef, double *freq)
{
    return guarded([&] {
        if (!coef || !freq) throw Error(HM_ERR_ARG, "NULL output");
        std::memcpy(coef, kSpsCoef, sizeof(kSpsCoef));
        std::memcpy(freq, kSpsFreq, sizeof(kSpsFreq));
    });
}

int hm_gaussian_modulation(double tune, double sigma, int start, int end, uint64_t seed, double *out)
{
    // src/modulation.cpp:62-78 with an explicit seed
    return guarded([&] {
        if (!(start < end)) throw Error(HM_ERR_ARG, "modulation: start must be smaller than end");
        std::mt19937 gen((std::mt19937::result_type)seed);
        std::normal_distribution<double> dist(tune, sigma);
        for (int i = 0; i < end - start; i++) out[i] = (2 * M_PI * dist(gen));
    });
}

int hm_uniform_modulation(double from, double to, int start, int end, uint64_t seed, double *out)
{
    // src/modulation.cpp:80-96 with an explicit seed
    return guarded([&] {
        if (!(start < end)) throw Error(HM_ERR_ARG, "modulation: start must be smaller than end");
        std::mt19937 gen((std::mt19937::result_type)seed);
        std::uniform_real_distribution<double> dist(from, to);
        for (int i = 0; i < end - start; i++) out[i] = (2 * M_PI * dist(gen));
    });
}

int hm_build_table(const double *omega_x, const double *omega_y, size_t n, double *table)
{
    // src/henon.cu:2670-2692: sin/cos of each omega, glibc libm
    return guarded([&] {
        if (n && (!omega_x || !omega_y || !table)) throw Error(HM_ERR_ARG, "NULL pointer");
        host_parallel(n, [&](size_t i) {
            table[4 * i + 0] = std::sin(omega_x[i]);
            table[4 * i + 1] = std::cos(omega_x[i]);
            table[4 * i + 2] = std::sin(omega_y[i]);
            table[4 * i + 3] = std::cos(omega_y[i]);
        });
    });
}

int hm_birkhoff_weights(uint32_t n_weights, double *out)
{
    // src/dynamic_indicator.cpp:63-79
    return guarded([&] {
        if (n_weights && !out) throw Error(HM_ERR_ARG, "out is NULL");
        for (uint32_t i = 0; i < n_weights; i++) {
            double t = (double)i / (double)(n_weights);
            out[i] = std::exp(-1 / (t * (1.0 - t)));
        }
        double sum = 0.0;
        for (uint32_t i = 0; i < n_weights; i++) sum = sum + out[i];
        for (uint32_t i = 0; i < n_weights; i++) out[i] = out[i] / sum;
    });
}

int hm_profile_enable(int on)
{
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_prof_mu);
        g_prof_on = on != 0;
    });
}

int hm_profile_reset(void)
{
    return guarded([&] {
        std::lock_guard<std::mutex> lk(g_prof_mu);
        drain_pending_locked();
        g_totals.clear();
    });
}

int hm_profile_read(const char *name, double *total_ms, uint64_t *launches)
{
    return guarded([&] {
        if (!total_ms || !launches) throw Error(HM_ERR_ARG, "NULL output");
        std::lock_guard<std::mutex> lk(g_prof_mu);
        drain_pending_locked();
        double ms = 0.0;
        uint64_t cnt = 0;
        for (auto &kv : g_totals) {
            if (!name || kv.first.find(name) != std::string::npos) {
                ms += kv.second.ms;
                cnt += kv.second.launches;
            }
        }
        *total_ms = ms;
        *launches = cnt;
    });
}

}  // extern "C"
