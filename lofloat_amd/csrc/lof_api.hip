// lof_api.hip -- library-wide C-ABI bits: error reporting and version.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>

#include "lof_host.hpp"

namespace lofamd {

namespace {
thread_local char g_last_error[512] = "";
}

lof_status set_error(lof_status st, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return st;
}

lof_status check_hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return LOF_OK;
    return set_error(LOF_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

}  // namespace lofamd

extern "C" {

LOF_API const char* lof_last_error(void) { return lofamd::g_last_error; }

LOF_API int32_t lof_version(void) { return 1; }

}  // extern "C"
