// SPDX-License-Identifier: 0BSD
// Error plumbing shared by the host-side orchestration (DeviceRun, Ensemble)
// and their C ABIs: HIP failures become exceptions carrying the hipError_t
// code, and the C entry points turn exceptions back into return codes that
// are also recorded for lpbf_get_last_error().
#pragma once

#include <hip/hip_runtime_api.h>

#include <mutex>
#include <new>
#include <stdexcept>
#include <string>

#include "lpbf_kernels.h"

namespace lpbf_amd {
void record_error(int err);
}

namespace lpbf {

struct HipFailure : std::runtime_error {
    int code;
    HipFailure(int c, std::string const& m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char* who, const char* what) {
    if (e != hipSuccess)
        throw HipFailure(static_cast<int>(e),
                         std::string(who) + ": " + what + ": " + hipGetErrorString(e));
}

// Launch errors are recorded by the library; surface them per batch.
inline void lib_check(const char* who, const char* what) {
    const int e = lpbf_get_last_error();
    if (e != 0)
        throw HipFailure(e, std::string(who) + ": " + what + ": " + lpbf_error_string(e));
}

// Runs on several host threads (one stream each) may share the device.  A
// graph capture must not overlap another thread's allocation, free or
// device-wide call, so capture, set-up and teardown take this lock; the
// steady state (graph launches, batch copies and syncs) does not.
std::mutex& setup_mutex();

// Runs f() for a C entry point: 0, or the error code (also recorded).
template <class F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (HipFailure const& e) {
        lpbf_amd::record_error(e.code);
        return e.code;
    } catch (std::bad_alloc const&) {
        lpbf_amd::record_error(static_cast<int>(hipErrorOutOfMemory));
        return static_cast<int>(hipErrorOutOfMemory);
    } catch (...) {
        lpbf_amd::record_error(LPBF_ERR_INVALID_ARGUMENT);
        return LPBF_ERR_INVALID_ARGUMENT;
    }
}

}  // namespace lpbf
