// Error reporting and version of the citrus HIP library.
#include "common.h"

#include "citrus_hip.h"

namespace citrus {
namespace {
thread_local std::string last_error;
}

int fail(int code, const std::string& msg) {
  last_error = msg;
  return code;
}

}  // namespace citrus

extern "C" const char* citrus_last_error(void) { return citrus::last_error.c_str(); }

extern "C" int citrus_abi_version(void) { return 20; }
