// Shared helpers for the libtrg HIP sources (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/trg.h"

namespace trg {

extern thread_local int g_last_hip_error;

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Entry points clear the thread's last HIP error before launching, so a stale
// error left by an unrelated runtime call on this thread (which its caller
// already saw) is not reported as this launch's failure.
inline void clear_hip_error() { (void)hipGetLastError(); }

// Launch-error check after a kernel launch: returns TRG_HIP_ERROR and records the code.
inline int check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        g_last_hip_error = static_cast<int>(e);
        return TRG_HIP_ERROR;
    }
    return TRG_OK;
}

inline long long cdiv(long long a, long long b) { return (a + b - 1) / b; }

}  // namespace trg

#define TRG_TRY(expr)                         \
    do {                                      \
        int _st = (expr);                     \
        if (_st != TRG_OK) return _st;        \
    } while (0)
