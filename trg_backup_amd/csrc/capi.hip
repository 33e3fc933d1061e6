// Library-level entry points: version, status strings, device properties.
#include "common.h"

// The ABI revision after "abi" changes whenever an entry point's signature does; the
// Python binding refuses a library built with a different one.
extern "C" const char *trg_version(void) { return "libtrg 0.8 abi 8 (gfx950)"; }

extern "C" const char *trg_status_string(int status) {
    switch (status) {
        case TRG_OK: return "ok";
        case TRG_BAD_ARG: return "bad argument";
        case TRG_SHAPE: return "shape beyond supported maximum";
        case TRG_OOM: return "workspace too small";
        case TRG_NUMERIC: return "numeric failure (NaN/Inf or non-positive pivot)";
        case TRG_UNSUPPORTED: return "unsupported dtype or device";
        case TRG_HIP_ERROR: return "HIP runtime error";
        default: return "unknown status";
    }
}

extern "C" int trg_device_info(int device, int *num_cus, int *lds_per_block, int *warp_size) {
    int v = 0;
    if (num_cus) {
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return TRG_HIP_ERROR;
        *num_cus = v;
    }
    if (lds_per_block) {
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) return TRG_HIP_ERROR;
        *lds_per_block = v;
    }
    if (warp_size) {
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeWarpSize, device) != hipSuccess) return TRG_HIP_ERROR;
        *warp_size = v;
    }
    return TRG_OK;
}
