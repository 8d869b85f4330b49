// Shared helpers of the gfx950 FVM kernels: error reporting and launch checks.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../../include/fvm_hip.h"

namespace fvm {

void set_error(const char *fmt, ...);

inline int check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("%s: %s", what, hipGetErrorString(e));
        return FVM_ERR_HIP;
    }
    return FVM_OK;
}

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned int blocks_for(int64_t n, int threads) {
    return static_cast<unsigned int>((n + threads - 1) / threads);
}

}  // namespace fvm
