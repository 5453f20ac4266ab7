// Host-side status plumbing shared by the kernel translation units of
// libpfh.so: every C entry point returns an int status and records a
// message for pfh_last_error() on failure.
#pragma once

#include <hip/hip_runtime.h>

namespace pfh_host {

// Records `what: <HIP error text>` and returns PFH_ERR_HIP unless e == hipSuccess.
int hip_status(hipError_t e, const char* what);
// Status of the last kernel launch on this thread.
int launch_status(const char* what);
// Records "invalid argument: what" and returns PFH_ERR_INVALID.
int invalid(const char* what);
// Records the message and returns PFH_ERR_UNSUPPORTED.
int unsupported(const char* what);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace pfh_host
