// Host-side helpers shared by the library's translation units: error state,
// HIP error checks, kernel-launch instrumentation and the device workspace.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>
#include <stdexcept>

#include "henon_hip.h"

namespace hm {

// An error carrying an HM_ERR_* status; the C entry points catch it.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_error(const std::string &msg);

inline void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        throw Error(HM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Runs `body`, mapping exceptions to status codes and the thread's message.
template <class F>
int guarded(F &&body)
{
    try {
        body();
        return HM_OK;
    } catch (const Error &e) {
        set_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return HM_ERR_ALLOC;
    } catch (const std::exception &e) {
        set_error(e.what());
        return HM_ERR_ARG;
    }
}

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Kernel-launch instrumentation (hm_profile_*): a scope that records HIP events
// around one launch when profiling is enabled.
class LaunchTimer {
  public:
    LaunchTimer(const char *name, hipStream_t s);
    ~LaunchTimer();
  private:
    const char *name_;
    hipStream_t stream_;
    hipEvent_t start_ = nullptr;
};

// Checks the launch status right after a <<<>>> launch.
inline void launch_check(const char *name)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error(HM_ERR_HIP, std::string("launch of ") + name + ": " + hipGetErrorString(e));
}

// Library-owned device words, per device: the status flags word (see
// hm_take_status; static device memory, never allocated) and growable scratch
// buffers, one per slot: 0 tracking live
// lists, 1 Birkhoff weights of the indicator kernels, 2 tune-window weights,
// 3 DFT magnitudes, 4 batched-job lists and plans, 5-6 autotuner trial state,
// 7 reciprocal table, 8 FFT signal rows, 9 batched MEGNO state, 10-11
// indicator-kernel state and live lists. Growing a slot synchronises the
// device. Calls that overlap in time on one device share these buffers.
constexpr int kWorkspaceSlots = 12;
uint32_t *status_word();
void *workspace(size_t bytes, int slot);
// SIMDs of the current device (CUs x 4), cached per device.
uint32_t device_simds();

constexpr size_t kTableRecord = 4;  // doubles per turn in the interleaved table

// Device table of 1.0 / k for k = 1..n (at least n entries; workspace slot 7).
const double *reciprocal_table(uint32_t n, hipStream_t s);

}  // namespace hm
