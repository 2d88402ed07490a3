// Error reporting shared by the library's translation units: a failing call
// records its message (mia_hip_last_error()) and returns the status code.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/mia_hip.h"

namespace mia_hip_rt {
int fail(int code, const std::string &msg);

// The divcurl penalty object's coefficient grid (x, y, z).
int divcurl_grid(const mia_hip_divcurl *p, uint32_t cs[3]);

// Instance `inst` of a set-up context: dimension, image size (x, y, z; z = 1
// in 2D) and the context's device.
int instance_geometry(const mia_hip_ctx *ctx, int inst, int *dim, uint32_t size[3], int *device);

// Enqueue on stream s: the penalty of the double parameters `params` (one
// instance's, rounded to float as the transformation stores them), its
// unweighted value into *value and -weight times its gradient into wgrad
// (each product rounded on its own).  Uses the penalty object's partial-sum
// buffer: evaluations of one penalty object must not overlap.
int divcurl_enqueue(mia_hip_divcurl *p, hipStream_t s, const double *params, double weight, double *wgrad,
                    double *value);

// A non-blocking stream of `device` (the current device): one that an earlier
// user gave back, else a new one.  Creating a stream can take milliseconds
// (the runtime may set up a hardware queue), and a registration creates and
// destroys contexts level by level.
hipError_t stream_acquire(int device, hipStream_t *s);

// Give a stream back: drained, then kept for stream_acquire (up to eight per
// device) or destroyed.
void stream_release(int device, hipStream_t s);
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
