// api_common.h -- error plumbing shared by the C ABI translation units.
#pragma once

#include "rise_hip.h"

#include <hip/hip_runtime_api.h>

#include <string>

namespace rise_hip {

// Records msg as the calling thread's last error and returns s.
rise_hip_status fail(rise_hip_status s, const std::string& msg);
rise_hip_status fail_hip(hipError_t e, const char* what);
void clear_error();

}  // namespace rise_hip
