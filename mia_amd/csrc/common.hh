// Error reporting shared by the library's translation units: a failing call
// records its message (mia_hip_last_error()) and returns the status code.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/mia_hip.h"

namespace mia_hip_rt {
int fail(int code, const std::string &msg);
} // namespace mia_hip_rt

#define HIP_TRY(expr)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return mia_hip_rt::fail(MIA_HIP_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define LAUNCH_CHECK(name)                                                                              \
    do {                                                                                                \
        hipError_t e_ = hipGetLastError();                                                              \
        if (e_ != hipSuccess)                                                                           \
            return mia_hip_rt::fail(MIA_HIP_ERR_DEVICE,                                                 \
                                    std::string("launch ") + name + ": " + hipGetErrorString(e_));      \
    } while (0)
