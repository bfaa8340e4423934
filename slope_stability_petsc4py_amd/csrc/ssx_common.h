// Shared helpers for the ssx HIP library: error channel, launch checks, grid sizing.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

namespace ssx {

// Last error message of the calling host thread (returned by ssx_last_error()).
void set_error(const char *fmt, ...);

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Number of workgroups for n items at `block` items per workgroup (at least 1).  Not clamped: the
// largest launch at one GPU's sizes (250M nonzeros at 256 per workgroup) stays far below the HIP
// limit of 2^31 - 1 workgroups.
inline unsigned grid_for(int64_t n, int block) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    return (unsigned)g;
}

}  // namespace ssx

#define SSX_OK 0
#define SSX_EINVAL 1
#define SSX_ELAUNCH 2
#define SSX_ERUNTIME 3

// Check the launch that was just issued; on failure record the message and return.
#define SSX_CHECK_LAUNCH(name)                                                                 \
    do {                                                                                       \
        hipError_t _e = hipGetLastError();                                                     \
        if (_e != hipSuccess) {                                                                \
            ssx::set_error("%s: launch failed: %s", name, hipGetErrorString(_e));              \
            return SSX_ELAUNCH;                                                                \
        }                                                                                      \
    } while (0)

#define SSX_REQUIRE(cond, ...)                                                                 \
    do {                                                                                       \
        if (!(cond)) {                                                                         \
            ssx::set_error(__VA_ARGS__);                                                       \
            return SSX_EINVAL;                                                                 \
        }                                                                                      \
    } while (0)
