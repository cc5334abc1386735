// SPDX-License-Identifier: 0BSD
// Library-wide error status of the C ABI (first error since the last query).
#include <hip/hip_runtime_api.h>

#include <atomic>

#include "lpbf_kernels.h"

namespace lpbf_amd {

namespace {
std::atomic<int> g_last_error{0};
}

void record_error(int err) {
    int expected = 0;
    g_last_error.compare_exchange_strong(expected, err);
}

}  // namespace lpbf_amd

extern "C" {

int lpbf_get_last_error(void) { return lpbf_amd::g_last_error.exchange(0); }

const char* lpbf_error_string(int err) {
    if (err == 0) return "success";
    if (err == LPBF_ERR_INVALID_ARGUMENT) return "invalid argument to an lpbf_* call";
    if (err == LPBF_ERR_SCRATCH_LIMIT)
        return "lpbf_launch_step_f32_noalloc: the launch needs more workspace than the "
               "library's fixed scratch (LPBF_F32_NOALLOC_SCRATCH_BYTES); use "
               "lpbf_launch_step_f32 with a caller workspace, or taller bands";
    return hipGetErrorString(static_cast<hipError_t>(err));
}

}  // extern "C"
