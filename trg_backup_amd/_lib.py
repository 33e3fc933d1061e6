"""ctypes binding of libtrg.so (the C ABI declared in include/trg.h).

The library is built in-tree by ``trg_backup_amd.build``.  There is no CPU
fallback: every op of this package runs on the GPU through this library or
raises ``TrgError``.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtrg.so")

F32, F64, BF16, I32, BF16X3, BF16X6 = 0, 1, 2, 3, 4, 5
ABI = 8          # must match the "abi N" revision in trg_version()

c_int, c_ll, c_dbl, c_vp, c_ull, c_uint = (ctypes.c_int, ctypes.c_longlong, ctypes.c_double, ctypes.c_void_p,
                                           ctypes.c_ulonglong, ctypes.c_uint)
P_LL = ctypes.POINTER(ctypes.c_longlong)

# name -> argument types (restype is int unless listed in _RESTYPES)
SIGNATURES = {
    "trg_version": [],
    "trg_status_string": [c_int],
    "trg_last_hip_error": [],
    "trg_device_info": [c_int, c_vp, c_vp, c_vp],
    "trg_gemm": [c_int, c_int, c_int, c_int, c_int, c_int, c_dbl,
                 c_vp, c_ll, c_ll, c_ll, c_vp, c_ll, c_ll, c_ll,
                 c_dbl, c_vp, c_ll, c_ll, c_ll, c_vp],
    "trg_gemm_ex": [c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_dbl,
                    c_vp, c_ll, c_ll, c_ll, c_vp, c_ll, c_ll, c_ll,
                    c_dbl, c_vp, c_ll, c_ll, c_ll, c_int, c_ll, c_int, c_ll, c_vp],
    "trg_gemm_f64i8_workspace": [c_int, c_int, c_int, c_int, c_int, c_int],
    "trg_gemm_f64i8": [c_int, c_int, c_int, c_int, c_int, c_dbl,
                       c_vp, c_ll, c_ll, c_ll, c_vp, c_ll, c_ll, c_ll,
                       c_dbl, c_vp, c_ll, c_ll, c_ll, c_int, c_ll, c_int, c_ll, c_vp, c_ll, c_vp],
    "trg_gemm_config_count": [c_int, c_int, c_int],
    "trg_gemm_config_name": [c_int, c_int, c_int, c_int],
    "trg_copy_strided": [c_int, c_int, c_int, P_LL, c_vp, P_LL, c_vp, P_LL, c_vp],
    "trg_fill": [c_int, c_ll, c_dbl, c_vp, c_vp],
    "trg_maxabs": [c_int, c_int, c_ll, c_vp, c_ll, c_vp, c_vp, c_int, c_vp],
    "trg_trace": [c_int, c_int, c_int, P_LL, P_LL, c_vp, c_ll, c_vp, c_vp],
    "trg_normalize": [c_int, c_int, c_ll, c_vp, c_ll, c_vp, c_vp, c_ll, c_int, c_vp, c_vp],
    "trg_syevj_f64": [c_int, c_int, c_vp, c_ll, c_vp, c_vp, c_int, c_vp, c_vp],
    "trg_chol_inv_f64": [c_int, c_int, c_vp, c_ll, c_dbl, c_vp, c_vp, c_vp],
    "trg_gaussian": [c_int, c_int, c_ll, c_ull, c_uint, c_uint, c_ull, c_vp, c_ll, c_vp],
    "trg_eigh_topk_workspace": [c_int, c_int, c_int],
    "trg_eigh_topk_f64": [c_int, c_int, c_int, c_int, c_int, c_vp, c_ll, c_vp, c_vp, c_vp, c_ull,
                          c_vp, c_ll, c_vp, c_vp],
    "trg_eigh_topk_noalloc_max_bytes": [],
    "trg_eigh_topk_f64_noalloc": [c_int, c_int, c_int, c_int, c_int, c_vp, c_ll, c_vp, c_vp, c_vp, c_ull, c_vp, c_vp],
    "trg_select_isometry": [c_int, c_int, c_ll, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp],
    "trg_scale_rows": [c_int, c_int, c_ll, c_ll, c_ll, c_vp, c_ll, c_vp, c_ll, c_int, c_vp],
    "trg_lnz": [c_int, c_int, c_vp, c_ll, c_vp, c_dbl, c_vp, c_vp],
    "trg_gesvj_f64": [c_int, c_int, c_int, c_vp, c_ll, c_vp, c_vp, c_vp, c_int, c_vp, c_vp],
    "trg_truncate_rank": [c_int, c_vp, c_int, c_dbl, c_dbl, c_vp, c_vp, c_vp],
    "trg_potrf_psd_f64": [c_int, c_int, c_vp, c_ll, c_dbl, c_vp, c_vp, c_vp, c_vp],
    "trg_orthonormalize_workspace": [c_int, c_int, c_int],
    "trg_orthonormalize_f64": [c_int, c_int, c_int, c_vp, c_ll, c_vp, c_vp, c_ll, c_int, c_vp, c_vp],
    "trg_axpy_dev_f64": [c_int, c_ll, c_dbl, c_vp, c_vp, c_ll, c_vp, c_ll, c_vp],
    "trg_health_or": [c_ll, c_vp, c_ll, c_vp, c_vp, c_vp],
}
_RESTYPES = {"trg_version": ctypes.c_char_p, "trg_status_string": ctypes.c_char_p,
             "trg_gemm_config_name": ctypes.c_char_p,
             "trg_eigh_topk_workspace": c_ll, "trg_gemm_f64i8_workspace": c_ll, "trg_eigh_topk_noalloc_max_bytes": c_ll, "trg_orthonormalize_workspace": c_ll}


class TrgError(RuntimeError):
    pass


_lib = None


def hip_runtimes():
    """Paths of the HIP runtime libraries mapped into this process (Linux)."""
    found = set()
    try:
        with open("/proc/self/maps") as f:
            for line in f:
                path = line.split()[-1]
                if "libamdhip64" in os.path.basename(path):
                    found.add(path)
    except OSError:
        pass
    return sorted(found)


def load():
    """Load libtrg.so (raises if it has not been built).

    The library takes torch's device pointers and stream handles, so it must
    run on the HIP runtime torch loaded.  torch is therefore imported first:
    loaded before it, libtrg.so would bind its own copy of the runtime
    (/opt/rocm's), torch would then bring a second one, and the library's
    first launch on a torch stream fails (hipErrorNoDevice)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TrgError(f"{LIB_PATH} is not built; run `python -m trg_backup_amd.build` "
                       "(or __graft_entry__.build()) first")
    import torch  # noqa: F401  (binds the process to torch's HIP runtime before the library loads)
    lib = ctypes.CDLL(LIB_PATH)
    rts = hip_runtimes()
    if len(rts) > 1:
        raise TrgError(f"two HIP runtimes are mapped ({', '.join(rts)}): libtrg.so must share torch's")
    lib.trg_version.restype = ctypes.c_char_p
    got = lib.trg_version().decode()
    if f" abi {ABI} " not in got:
        raise TrgError(f"{LIB_PATH} reports {got!r}, this binding needs abi {ABI}; rebuild the library")
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, c_int)
    _lib = lib
    return lib


def call(name, *args):
    """Call an entry point and raise TrgError on a non-zero status."""
    lib = load()
    st = getattr(lib, name)(*args)
    if st != 0:
        msg = lib.trg_status_string(st).decode()
        extra = ""
        if st == 6:
            extra = f" (hip error {lib.trg_last_hip_error()})"
        raise TrgError(f"{name} failed: {msg}{extra}")
    return st


def ll_array(vals):
    arr = (ctypes.c_longlong * max(1, len(vals)))(*[int(v) for v in vals])
    return arr
