// gpe_hip_check.hpp -- turning a failed HIP runtime call into GPE_EDEVICE
// with the call, its error string and the source line as the thread's
// last error.
#pragma once
#include <hip/hip_runtime.h>

#include "gpe_internal.hpp"

#define GPE_HIP_TRY(expr)                                                                         \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            gpe::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
            return GPE_EDEVICE;                                                                   \
        }                                                                                         \
    } while (0)
