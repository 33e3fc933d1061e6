"""CPU checks of the C ABI: the library loads and exports every declared symbol."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "trg.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(trg_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = _declared()
    assert "trg_gemm" in names and "trg_eigh_topk_f64" in names and len(names) >= 15


def test_library_exports_every_declared_symbol():
    from trg_backup_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libtrg.so not built (run __graft_entry__.build())")
    lib = _lib.load()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_declared()) == set(_lib.SIGNATURES), set(_declared()) ^ set(_lib.SIGNATURES)
    assert lib.trg_version().decode().startswith("libtrg")
    assert lib.trg_status_string(4).decode().startswith("numeric")


def test_bad_arguments_are_rejected_without_a_gpu():
    from trg_backup_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libtrg.so not built")
    lib = _lib.load()
    # negative sizes / null pointers are refused before any launch
    assert lib.trg_gemm(0, 0, 1, -1, 1, 1, 1.0, None, 0, 1, 1, None, 0, 1, 1, 0.0, None, 0, 1, 1, None) == 1
    assert lib.trg_gemm(0, 0, 1, 4, 4, 4, 1.0, None, 0, 1, 4, None, 0, 4, 1, 0.0, None, 0, 4, 1, None) == 1
    assert lib.trg_syevj_f64(1, 300, None, 0, None, None, 10, None, None) == 2
    assert lib.trg_eigh_topk_f64(1, 10, 5, 20, 2, None, 0, None, None, None, 0, None, 0, None, None) == 2


def test_int8_gemm_workspace_and_argument_checks_without_a_gpu():
    """trg_gemm_f64i8's workspace: the int8 slices of A and B at a row pitch of
    K rounded up to 64 (one K chunk: slices 127^2 K < 2^31), then the int32 row
    exponents, each part 256-byte aligned; bad arguments return before any launch."""
    import ctypes
    from trg_backup_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libtrg.so not built")
    lib = _lib.load()
    al = lambda v: (v + 255) // 256 * 256
    ws = lambda S, ba, bb, M, N, kp: al(ba * S * M * kp) + al(bb * S * N * kp) + al(ba * M * 4) + al(bb * N * 4)
    assert lib.trg_gemm_f64i8_workspace(6, 1, 1, 64, 64, 64) == ws(6, 1, 1, 64, 64, 64)
    assert lib.trg_gemm_f64i8_workspace(6, 3, 1, 100, 37, 65) == ws(6, 3, 1, 100, 37, 128)
    chunk = (2 ** 31 - 1) // (6 * 127 * 127) // 64 * 64
    assert lib.trg_gemm_f64i8_workspace(6, 1, 1, 8, 8, 10 ** 6) == ws(6, 1, 1, 8, 8, chunk)
    assert lib.trg_gemm_f64i8_workspace(7, 1, 1, 0, 0, 0) == 0
    assert lib.trg_gemm_f64i8_workspace(3, 1, 1, 4, 4, 4) == -1 and lib.trg_gemm_f64i8_workspace(8, 1, 1, 4, 4, 4) == -1
    p = ctypes.c_void_p(4096)   # never dereferenced: every call below returns before a launch
    args = lambda S, ws_bytes: (S, 1, 64, 64, 64, 1.0, p, 0, 64, 1, p, 0, 64, 1, 0.0, p, 0, 64, 1, 0, 0, 0, 0,
                                p, ws_bytes, None)
    assert lib.trg_gemm_f64i8(*args(9, 1 << 20)) == 1
    assert lib.trg_gemm_f64i8(*args(6, 1000)) == 3
    assert lib.trg_gemm_f64i8(6, 1, -1, 4, 4, 1.0, p, 0, 4, 1, p, 0, 4, 1, 0.0, p, 0, 4, 1, 0, 0, 0, 0, p, 1 << 20,
                              None) == 1


def test_library_shares_torchs_hip_runtime_whatever_the_import_order():
    """Loading the library before anything imported torch must still leave a
    single HIP runtime in the process (torch's): the library takes torch's
    device pointers and stream handles.  Run in a fresh interpreter, since
    this one has imported torch already."""
    import subprocess
    import sys
    from trg_backup_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libtrg.so not built (run __graft_entry__.build())")
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from trg_backup_amd import _lib\n"
            "assert 'torch' not in sys.modules\n"
            "_lib.load()\n"
            "import torch\n"
            "print(_lib.hip_runtimes())\n") % ROOT
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    rts = eval(out.stdout.strip().splitlines()[-1])
    assert len(rts) == 1 and "torch" in rts[0], rts
