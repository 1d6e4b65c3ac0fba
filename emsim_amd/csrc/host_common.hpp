// Host-side helpers shared by the library's translation units: the
// thread-local error string, HIP status checks, and kernel launches with
// optional per-kernel event timing.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/emsim_hip.h"

namespace emsim_host {

// Records the message for emsim_hip_last_error() and returns code.
int fail(int code, const char* fmt, ...);

// Event timing around one launch (no-ops unless emsim_hip_timing_enable(1)).
hipEvent_t timing_begin(hipStream_t s);
void timing_end(hipEvent_t start, const char* name, int n, hipStream_t s);

}  // namespace emsim_host

#define EMSIM_HIP_CHECK(call)                                                          \
    do {                                                                                \
        hipError_t err_ = (call);                                                       \
        if (err_ != hipSuccess)                                                         \
            return emsim_host::fail(EMSIM_HIP_EDEVICE, "%s failed: %s", #call,          \
                                    hipGetErrorString(err_));                           \
    } while (0)

#define EMSIM_LAUNCH(tag_n, kernel, grid, block, shmem, stream, ...)                      \
    do {                                                                                \
        hipEvent_t tk_ = emsim_host::timing_begin(stream);                              \
        hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__);            \
        hipError_t err_ = hipGetLastError();                                            \
        if (err_ != hipSuccess)                                                         \
            return emsim_host::fail(EMSIM_HIP_EDEVICE, "launch of %s failed: %s",       \
                                    #kernel, hipGetErrorString(err_));                  \
        emsim_host::timing_end(tk_, #kernel, tag_n, stream);                            \
    } while (0)
