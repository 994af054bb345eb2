// Shared host-side helpers of the fadesim HIP library: error reporting and
// the library-owned device workspace.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>

namespace fadesim {

// Records `msg` as the calling thread's last error and returns `code`.
int set_error(int code, const std::string &msg);

// Grows (never shrinks) the per-device workspace to at least `bytes` and
// returns it, or nullptr with the last error set.  The buffer is reused by
// every call on that device; calls on one device are stream-ordered by the
// caller.
void *workspace(size_t bytes);
// A second, independent buffer with the same rules (for kernels that run
// while the first one is in use by the same call).
void *workspace_aux(size_t bytes);
// A third one (the warm starts the in-kernel network start hands on).
void *workspace_warm(size_t bytes);

#define FADESIM_HIP_CHECK(expr)                                                   \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess)                                                     \
            return ::fadesim::set_error(FADESIM_E_HIP, std::string(#expr ": ") + \
                                                           hipGetErrorString(e_)); \
    } while (0)

} // namespace fadesim
