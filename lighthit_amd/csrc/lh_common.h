// Shared definitions of the lighthit_amd C-ABI library (gfx950 only).
//
// Every entry point returns an lh_status; on failure the message of the
// calling thread is available from lh_last_error(). Device pointers are
// owned by the caller; no entry point allocates device memory.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "../../include/lighthit_amd.h"

namespace lh {

// Records a printf-style message for lh_last_error() of this thread.
void set_error(const char* fmt, ...);

inline int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("%s launch failed: %s", what, hipGetErrorString(e));
        return LH_ERR_HIP;
    }
    return LH_OK;
}

// Kernel timer: while enabled (lh_timing_enable), a KernelMark records a HIP
// event on its stream when built and another when destroyed, under one
// kernel name. Scope it around exactly one launch.
bool timing_enabled();
void timing_record(const char* name, hipStream_t stream, hipEvent_t start);
hipEvent_t timing_start(hipStream_t stream);

class KernelMark {
  public:
    KernelMark(const char* name, hipStream_t stream)
        : name_(name), stream_(stream), start_(timing_enabled() ? timing_start(stream) : nullptr) {}
    ~KernelMark() {
        if (start_ != nullptr) timing_record(name_, stream_, start_);
    }
    KernelMark(const KernelMark&) = delete;
    KernelMark& operator=(const KernelMark&) = delete;

  private:
    const char* name_;
    hipStream_t stream_;
    hipEvent_t start_;
};

}  // namespace lh

#define LH_REQUIRE(cond, ...)                                   \
    do {                                                        \
        if (!(cond)) {                                          \
            lh::set_error(__VA_ARGS__);                         \
            return LH_ERR_ARGUMENT;                             \
        }                                                       \
    } while (0)

#define LH_HIP_CALL(call)                                                     \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                               \
            lh::set_error("%s failed: %s", #call, hipGetErrorString(e_));     \
            return LH_ERR_HIP;                                                \
        }                                                                     \
    } while (0)
