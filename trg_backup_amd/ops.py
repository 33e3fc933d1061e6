"""Thin wrappers of the libtrg entry points over torch device buffers.

torch is used for device memory and the current stream only; every
arithmetic operation below is a libtrg kernel.
"""
import contextlib
import math
import threading

import torch

from . import _lib, autotune
from ._lib import BF16, BF16X3, BF16X6, F32, F64, I32, call, ll_array

_DT = {torch.float32: F32, torch.float64: F64, torch.bfloat16: BF16}


def dtype_code(t):
    try:
        return _DT[t.dtype]
    except KeyError:
        raise _lib.TrgError(f"unsupported dtype {t.dtype}; libtrg works on float32/float64/bfloat16") from None


def require_device(*ts):
    for t in ts:
        if not (isinstance(t, torch.Tensor) and t.is_cuda):
            raise _lib.TrgError("libtrg ops need torch tensors on the GPU (there is no CPU path)")


def stream_ptr():
    return torch.cuda.current_stream().cuda_stream


def ptr(t):
    return t.data_ptr()


def empty(shape, dtype, device):
    return torch.empty(shape, dtype=dtype, device=device)


# ---------------------------------------------------------------------------
# copies / permutations
# ---------------------------------------------------------------------------
def copy_into(dst, src):
    """dst[...] = src[...] for same-shape strided views (dtype conversion f32<->f64)."""
    require_device(dst, src)
    if tuple(dst.shape) != tuple(src.shape):
        raise _lib.TrgError(f"copy_into shape mismatch {tuple(src.shape)} -> {tuple(dst.shape)}")
    if dst.numel() == 0:
        return dst
    dims, si, so = _squeeze_dims(list(src.shape), list(src.stride()), list(dst.stride()))
    if len(dims) > 8:
        raise _lib.TrgError("copy_into supports rank <= 8 after merging")
    call("trg_copy_strided", dtype_code(src), dtype_code(dst), len(dims), ll_array(dims), ptr(src), ll_array(si),
         ptr(dst), ll_array(so), stream_ptr())
    return dst


def _squeeze_dims(dims, si, so):
    """Drop size-1 dims and merge neighbours that are contiguous in both views."""
    d, a, b = [], [], []
    for n, x, y in zip(dims, si, so):
        if n == 1:
            continue
        if d and a[-1] == x * n and b[-1] == y * n:
            d[-1] *= n
            a[-1] = x
            b[-1] = y
        else:
            d.append(n)
            a.append(x)
            b.append(y)
    if not d:
        d, a, b = [1], [0], [0]
    return d, a, b


def contiguous(t, dtype=None):
    """Materialise a (possibly permuted/strided) view as a new contiguous tensor."""
    out = torch.empty(tuple(t.shape), dtype=dtype or t.dtype, device=t.device)
    return copy_into(out, t)


def permute(t, perm, dtype=None):
    return contiguous(t.permute(*perm), dtype)


def cast(t, dtype):
    return contiguous(t, dtype)


def fill_(t, value):
    require_device(t)
    if not t.is_contiguous():
        raise _lib.TrgError("fill_ needs a contiguous tensor")
    code = I32 if t.dtype == torch.int32 else dtype_code(t)
    call("trg_fill", code, t.numel(), float(value), ptr(t), stream_ptr())
    return t


# ---------------------------------------------------------------------------
# solve health: breakdown flags and bad normalisations, checked once per solve
# ---------------------------------------------------------------------------
_SCOPES = threading.local()


class Health:
    """A device-side failure word for one solve (or one captured graph).

    While the scope is entered on a host thread, every eigensolver, Jacobi SVD
    and orthonormalisation call on that thread ORs its breakdown flags into the
    word (bit 1), and every normalisation ORs in a NaN, infinite or zero norm
    (bit 2) -- all on the device, without a host sync, so a captured graph
    keeps doing it on every replay.  check() synchronises the stream once and
    raises TrgError if any bit is set."""

    BITS = {1: "an eigensolver / SVD / Cholesky breakdown flag", 2: "a NaN, infinite or zero normalisation constant"}

    def __init__(self, device, word=None):
        self.word = word if word is not None else fill_(torch.empty(1, dtype=torch.int32, device=device), 0)

    def __enter__(self):
        stack = getattr(_SCOPES, "stack", None)
        if stack is None:
            stack = _SCOPES.stack = []
        stack.append(self)
        return self

    def __exit__(self, *exc):
        _SCOPES.stack.remove(self)
        return False

    def value(self):
        return int(self.word.cpu().item())

    def check(self, what="solve"):
        v = self.value()
        if v:
            why = "; ".join(m for b, m in self.BITS.items() if v & b)
            raise _lib.TrgError(f"{what} failed on the device: {why}")


@contextlib.contextmanager
def isolated_health(h=None):
    """Run a block in which only `h` (or no scope at all) receives health notes
    on this thread, e.g. the warm-up and capture of a HIP graph."""
    old = getattr(_SCOPES, "stack", None)
    _SCOPES.stack = [h] if h is not None else []
    try:
        yield h
    finally:
        _SCOPES.stack = old


def note(info=None, vals=None):
    """OR breakdown flags (int32 device tensor) and normalisation constants
    (float64 device tensor) into every Health scope active on this thread."""
    stack = getattr(_SCOPES, "stack", None)
    if not stack:
        return
    ni = info.numel() if info is not None else 0
    nv = vals.numel() if vals is not None else 0
    if ni == 0 and nv == 0:
        return
    if info is not None and not info.is_contiguous():
        info = info.contiguous()
    for h in stack:
        call("trg_health_or", ni, ptr(info) if ni else None, nv, ptr(vals) if nv else None, ptr(h.word),
             stream_ptr())


# ---------------------------------------------------------------------------
# GEMM
# ---------------------------------------------------------------------------
class GemmRecorder:
    """Records HIP events around tagged GEMM launches on the launching stream
    (used to measure a kernel's live launch duration inside a timed region)."""

    def __init__(self, tag):
        self.tag = tag
        self.events = []   # (start, end, flops)
        self.event_configs = []   # (dtype, C dtype, compute, tile config, M, N, K, batch) per event

    def __enter__(self):
        _RECORDERS.append(self)
        return self

    def __exit__(self, *exc):
        _RECORDERS.remove(self)
        return False

    def summary(self):
        torch.cuda.synchronize()
        if not self.events:
            return {"launches": 0}
        ms = [a.elapsed_time(b) for a, b, _ in self.events]
        fl = [f for _, _, f in self.events]
        return {"launches": len(ms), "avg_ms": sum(ms) / len(ms), "avg_flops": sum(fl) / len(fl),
                "total_ms": sum(ms), "total_flops": sum(fl)}


_RECORDERS = []


def gemm(A, B, C=None, alpha=1.0, beta=0.0, compute=None, tag=None, out_layout=None):
    """C[b] = alpha A[b] B[b] + beta C[b] for 2D or batched 3D strided views.

    A is (M, K) or (batch, M, K), B is (K, N) or (batch, K, N).  Views may
    be transposed as long as one of the two matrix strides of A and of B is 1.
    compute: None (same as storage), "bf16" / "bf16x3" / "bf16x6" (f32 storage only) or a dtype code.
    out_layout: write the product into a permuted tensor instead of a matrix
    view: C is then the destination tensor (contiguous, batch-major) and
    out_layout = (rows, cols), each a list of (size, stride) from outer to
    inner of at most two levels, with the row / column index of the product
    running over them in that order.
    """
    require_device(A, B)
    batched = A.dim() == 3
    if not batched:
        A, B = A.unsqueeze(0), B.unsqueeze(0)
        if C is not None and out_layout is None:
            C = C.unsqueeze(0)
    bt, M, K = A.shape
    bt2, K2, N = B.shape
    if K != K2 or (bt != bt2 and bt2 != 1 and bt != 1):
        raise _lib.TrgError(f"gemm shape mismatch {tuple(A.shape)} x {tuple(B.shape)}")
    batch = max(bt, bt2)
    split = (0, 0, 0, 0)
    if out_layout is not None:
        if C is None or not C.is_contiguous() or C.numel() != batch * M * N:
            raise _lib.TrgError("out_layout needs a contiguous destination with batch*M*N elements")
        (rows, cols), lv = out_layout, []
        for dims, n in ((rows, M), (cols, N)):
            if not 1 <= len(dims) <= 2 or math.prod(d for d, _ in dims) != n:
                raise _lib.TrgError(f"out_layout {out_layout} does not factor the {M}x{N} product")
            lv.append((0, 0, dims[0][1]) if len(dims) == 1 else (dims[1][0], dims[0][1], dims[1][1]))
        (dm, sm, rsC), (dn, sn, csC) = lv
        split = (dm, sm, dn, sn)
        C = C.view(batch, M * N)
        Cstr = (M * N if batch > 1 else 0, rsC, csC)
    if C is None:
        C = torch.empty((batch, M, N), dtype=A.dtype, device=A.device)
        beta = 0.0
    if A.dtype != B.dtype:
        raise _lib.TrgError("gemm operands must share a dtype")
    sA = A.stride(0) if bt > 1 else 0
    sB = B.stride(0) if bt2 > 1 else 0
    rsA, csA = A.stride(1), A.stride(2)
    rsB, csB = B.stride(1), B.stride(2)
    if M == 1:
        rsA = 1 if csA != 1 else rsA
    if K == 1:
        csA = 1 if rsA != 1 else csA
        rsB = 1 if csB != 1 else rsB
    if N == 1:
        csB = 1 if rsB != 1 else csB
    if rsA != 1 and csA != 1:
        A = contiguous(A)
        rsA, csA, sA = A.stride(1), A.stride(2), (A.stride(0) if bt > 1 else 0)
    if rsB != 1 and csB != 1:
        B = contiguous(B)
        rsB, csB, sB = B.stride(1), B.stride(2), (B.stride(0) if bt2 > 1 else 0)
    dt, cdt = dtype_code(A), dtype_code(C)
    if compute == "f64i8" and dt == F64 and cdt == F64 and min(M, N, K) >= I8_MIN_K:
        return _gemm_f64i8(A, B, C, alpha, beta, batch, M, N, K, sA, rsA, csA, sB, rsB, csB,
                           Cstr if out_layout is not None else (C.stride(0) if batch > 1 else 0, C.stride(1), C.stride(2)),
                           split, tag, batched, out_layout is not None)
    if compute is None:
        cc = BF16 if dt == BF16 else dt
    elif compute == "bf16":
        cc = BF16
    elif compute == "bf16x3":
        cc = BF16X3
    elif compute == "bf16x6":
        cc = BF16X6
    elif compute == "f32":
        cc = F32
    elif compute in ("f64", "f64i8"):
        cc = F64
    else:
        cc = compute
    if dt == F64:
        cc = F64
    opnds = (ptr(A), sA, rsA, csA, ptr(B), sB, rsB, csB)
    if out_layout is None:
        Cstr = (C.stride(0) if batch > 1 else 0, C.stride(1), C.stride(2))
    config = -1
    if autotune.active():
        # tuned with the call's own destination layout (strides, two-level
        # index), so the epilogue it times is the one this call runs
        def launch(cfg, out):
            call("trg_gemm_ex", cfg, dt, cdt, cc, batch, M, N, K, float(alpha), *opnds,
                 0.0, ptr(out), *Cstr, *split, stream_ptr())
        layout = "" if (Cstr[1], Cstr[2], split) == (N, 1, (0, 0, 0, 0)) else \
            f"{Cstr[1]},{Cstr[2]},{','.join(str(v) for v in split)}"
        config = autotune.choose(dt, cc, batch, M, N, K, opnds, launch, C if beta == 0.0 else None, cdt=cdt,
                                 layout=layout, like=C)
    recs = [r for r in _RECORDERS if tag is not None and r.tag == tag]
    if recs:
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
    call("trg_gemm_ex", config, dt, cdt, cc, batch, M, N, K, float(alpha), *opnds,
         float(beta), ptr(C), *Cstr, *split, stream_ptr())
    if recs:
        ev1.record()
        for r in recs:
            r.events.append((ev0, ev1, 2.0 * M * N * K * batch))
            r.event_configs.append((dt, cdt, cc, config, M, N, K, batch))
    if out_layout is not None:
        return None
    return C if batched else C[0]


# float64 products on the int8 matrix cores: signed 7-bit slices per row of A
# and column of B (6 slices: 42 bits, 21 int8 MFMAs per product); below
# I8_MIN_K in any of M, N, K the product is bound by memory, the slicing
# passes would add to it, and the fp64 MFMA runs
I8_SLICES = 6
I8_MIN_K = 256


def _gemm_f64i8(A, B, C, alpha, beta, batch, M, N, K, sA, rsA, csA, sB, rsB, csB, Cstr, split, tag, batched,
                layout):
    ws = _lib.load().trg_gemm_f64i8_workspace(I8_SLICES, batch if sA else 1, batch if sB else 1, M, N, K)
    if ws < 0:
        raise _lib.TrgError("trg_gemm_f64i8_workspace rejected the shape")
    work = torch.empty(max(ws, 1), dtype=torch.uint8, device=A.device)
    recs = [r for r in _RECORDERS if tag is not None and r.tag == tag]
    if recs:
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
    call("trg_gemm_f64i8", I8_SLICES, batch, M, N, K, float(alpha), ptr(A), sA, rsA, csA, ptr(B), sB, rsB, csB,
         float(beta), ptr(C), *Cstr, *split, ptr(work), ws, stream_ptr())
    if recs:
        ev1.record()
        for r in recs:
            r.events.append((ev0, ev1, 2.0 * M * N * K * batch))
            r.event_configs.append((F64, F64, "f64i8", I8_SLICES, M, N, K, batch))
    del work   # the caching allocator keeps the block until the stream has used it
    if layout:
        return None
    return C if batched else C[0]


# ---------------------------------------------------------------------------
# reductions / normalisation
# ---------------------------------------------------------------------------
def maxabs(T, batch=1):
    """c[b] = max |T_b| (float64) for `batch` contiguous instances."""
    require_device(T)
    n = T.numel() // batch
    nchunk = max(1, min(1024, n // 16384))
    part = torch.empty(batch * nchunk, dtype=torch.float64, device=T.device)
    c = torch.empty(batch, dtype=torch.float64, device=T.device)
    call("trg_maxabs", dtype_code(T), batch, n, ptr(T), n, ptr(c), ptr(part), nchunk, stream_ptr())
    return c


def trace_pairs(T, batch=1):
    """Tr over paired legs (0,1), (2,3), ... of each instance: 'aabb' / 'xxyytt'."""
    require_device(T)
    shape = T.shape[1:] if batch > 1 or T.dim() % 2 == 1 else T.shape
    st = T.stride()[1:] if len(shape) < T.dim() else T.stride()
    dims, strides = [], []
    for k in range(0, len(shape), 2):
        if shape[k] != shape[k + 1]:
            raise _lib.TrgError("trace over legs of different size")
        dims.append(shape[k])
        strides.append(st[k] + st[k + 1])
    t = torch.empty(batch, dtype=torch.float64, device=T.device)
    sb = T.stride(0) if len(shape) < T.dim() else T.numel()
    call("trg_trace", dtype_code(T), batch, len(dims), ll_array(dims), ll_array(strides), ptr(T), sb, ptr(t),
         stream_ptr())
    return t


def trace_general(T, dims, strides, batch=1, sb=None):
    t = torch.empty(batch, dtype=torch.float64, device=T.device)
    call("trg_trace", dtype_code(T), batch, len(dims), ll_array(dims), ll_array(strides), ptr(T),
         T.numel() // batch if sb is None else sb, ptr(t), stream_ptr())
    return t


def normalize_(T, c, lnf=None, step=0, batch=1, lnf_stride=0, step_ptr=None):
    """T_b /= |c_b| and lnf[b*lnf_stride + step] = ln|c_b| / 2^step."""
    require_device(T, c)
    n = T.numel() // batch
    call("trg_normalize", dtype_code(T), batch, n, ptr(T), n, ptr(c), ptr(lnf) if lnf is not None else None,
         lnf_stride, int(step), ptr(step_ptr) if step_ptr is not None else None, stream_ptr())
    note(vals=c if c.dtype == torch.float64 and c.is_contiguous() else None)
    return T


def scale_rows_(Y, x, period, inner, mode, batch=1):
    require_device(Y, x)
    n = Y.numel() // batch
    call("trg_scale_rows", dtype_code(Y), batch, n, period, inner, ptr(Y), n, ptr(x), x.numel() // batch, mode,
         stream_ptr())
    return Y


# ---------------------------------------------------------------------------
# linear algebra
# ---------------------------------------------------------------------------
def syevj(A, batch=1, max_sweeps=40):
    """Eigen-decomposition of `batch` symmetric n x n float64 matrices (A is consumed).
    Returns (w ascending, V columns)."""
    require_device(A)
    n = A.shape[-1]
    w = torch.empty((batch, n), dtype=torch.float64, device=A.device)
    V = torch.empty((batch, n, n), dtype=torch.float64, device=A.device)
    info = torch.empty(batch, dtype=torch.int32, device=A.device)
    call("trg_syevj_f64", batch, n, ptr(A), n * n, ptr(w), ptr(V), max_sweeps, ptr(info), stream_ptr())
    note(info=info)
    return w, V, info


def chol_inv(G, shift_rel=0.0, batch=1):
    require_device(G)
    l = G.shape[-1]
    R = torch.empty((batch, l, l), dtype=torch.float64, device=G.device)
    info = torch.empty(batch, dtype=torch.int32, device=G.device)
    call("trg_chol_inv_f64", batch, l, ptr(G), l * l, float(shift_rel), ptr(R), ptr(info), stream_ptr())
    return R, info


def gaussian(shape, seed, dtype=torch.float64, device="cuda", stream_id=0, offset=0, batch=1, stream_stride=0):
    out = torch.empty((batch,) + tuple(shape), dtype=dtype, device=device)
    n = out.numel() // batch
    call("trg_gaussian", dtype_code(out), batch, n, int(seed), int(stream_id), int(stream_stride), int(offset),
         ptr(out), n, stream_ptr())
    return out if batch > 1 else out[0]


def eigh_topk(A, k, l, q, seed=20240, batch=1):
    """Top-k eigenpairs of `batch` PSD n x n float64 matrices (A is read only).
    Returns (w ascending (batch,k), V (batch,n,k), eps (batch,), info)."""
    require_device(A)
    n = A.shape[-1]
    lib = _lib.load()
    wsb = lib.trg_eigh_topk_workspace(batch, n, l)
    ws = torch.empty(wsb // 8 + 1, dtype=torch.float64, device=A.device)
    w = torch.empty((batch, k), dtype=torch.float64, device=A.device)
    V = torch.empty((batch, n, k), dtype=torch.float64, device=A.device)
    eps = torch.empty(batch, dtype=torch.float64, device=A.device)
    info = torch.empty(batch, dtype=torch.int32, device=A.device)
    call("trg_eigh_topk_f64", batch, n, k, l, q, ptr(A), n * n, ptr(w), ptr(V), ptr(eps), int(seed), ptr(ws),
         ws.numel() * 8, ptr(info), stream_ptr())
    note(info=info)
    return w, V, eps, info


def select_isometry(eps0, eps1, V0, V1, out_dtype, batch=1):
    require_device(eps0, eps1, V0, V1)
    nk = V0.numel() // batch
    U = torch.empty(V0.shape, dtype=out_dtype, device=V0.device)
    flag = torch.empty(batch, dtype=torch.int32, device=V0.device)
    call("trg_select_isometry", dtype_code(U), batch, nk, ptr(eps0), ptr(eps1), ptr(V0), ptr(V1), ptr(U), ptr(flag),
         stream_ptr())
    return U, flag


def lnz(lnf, tr, volume, batch=1):
    """ln Z / V = sum lnf + ln|Tr T| / V per instance (float64, on the device)."""
    require_device(lnf)
    nterms = lnf.numel() // batch
    out = torch.empty(batch, dtype=torch.float64, device=lnf.device)
    call("trg_lnz", batch, nterms, ptr(lnf), nterms, ptr(tr) if tr is not None else None, 1.0 / float(volume),
         ptr(out), stream_ptr())
    return out


def identity(n, dtype, device, batch=1):
    """batch copies of the n x n identity, built with libtrg fills."""
    I = torch.empty((batch, n, n), dtype=dtype, device=device)
    fill_(I, 0.0)
    ones = torch.empty(batch * n, dtype=dtype, device=device)
    fill_(ones, 1.0)
    diag = torch.as_strided(I, (batch, n), (n * n, n + 1))
    copy_into(diag, ones.view(batch, n))
    return I


def flip(t, dim):
    """Copy of t reversed along `dim` (a negative-stride read through libtrg)."""
    require_device(t)
    out = torch.empty(tuple(t.shape), dtype=t.dtype, device=t.device)
    n = t.shape[dim]
    if out.numel() == 0:
        return out
    si = list(t.stride())
    base = ptr(t) + (n - 1) * si[dim] * t.element_size()
    si[dim] = -si[dim]
    call("trg_copy_strided", dtype_code(t), dtype_code(out), t.dim(), ll_array(list(t.shape)), base, ll_array(si),
         ptr(out), ll_array(list(out.stride())), stream_ptr())
    return out
