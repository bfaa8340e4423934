// Library-level entry points: version, error channel, device synchronisation.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "ssx_common.h"

namespace {
thread_local char g_err[512] = "";
}

namespace ssx {
void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}
}  // namespace ssx

extern "C" {

// ABI version of this library: bumped whenever an entry point's signature changes.
int ssx_version(void) { return 4; }

const char *ssx_last_error(void) { return g_err; }

int ssx_stream_sync(void *stream) {
    hipError_t e = hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream));
    if (e != hipSuccess) {
        ssx::set_error("ssx_stream_sync: %s", hipGetErrorString(e));
        return SSX_ERUNTIME;
    }
    return SSX_OK;
}

}  // extern "C"
