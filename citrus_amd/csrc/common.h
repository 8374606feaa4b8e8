// Shared host helpers of the citrus HIP library: per-thread error message and the failure helper.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

namespace citrus {

// Records `msg` as the calling thread's last error and returns `code`.
int fail(int code, const std::string& msg);

}  // namespace citrus
