/* bl_internal.h -- error plumbing shared by the library's translation units. */
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>

#include "../../include/bitlev_hip.h"

namespace bl {

/* Records a message for bl_last_error() and returns `code`. */
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

}  // namespace bl

#define BL_HIP(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return bl::fail(e_ == hipErrorOutOfMemory ? BL_EALLOC : BL_EHIP, "%s: %s", \
                            #call, hipGetErrorString(e_));                              \
    } while (0)

#define BL_LAUNCHED(what)                                                               \
    do {                                                                                \
        hipError_t e_ = hipGetLastError();                                              \
        if (e_ != hipSuccess)                                                           \
            return bl::fail(BL_EHIP, "%s launch: %s", what, hipGetErrorString(e_));    \
    } while (0)
