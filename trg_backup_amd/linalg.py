"""Float64 decompositions composed from libtrg kernels.

  svd_small   one-sided Jacobi (trg_gesvj_f64), m x n with n <= m, fits LDS
  svd_thin    thin SVD when min(m, n) <= 128: orthonormal basis of the long
              side (shifted CholeskyQR3), projection, small Jacobi SVD
  svd_topk    randomized range finder + power iterations for the leading
              singular triplets of a large matrix (Halko, Martinsson, Tropp,
              SIAM Rev. 53 (2011)), finished with svd_thin
  chol_psd_upper  blocked positive-semidefinite Cholesky R^T R = A
  truncate    the rank rule of utility/truncated_svd.py:17-78, on the device

Every arithmetic step is a libtrg kernel; the host only reads back the
integer ranks that size the next tensors (the reference syncs there too).
"""
import math

import torch

from . import _lib, ops
from ._lib import TrgError, call
from .ops import gemm, ptr, stream_ptr

SMALL_MAX = 128        # largest n for the one-workgroup Jacobi SVD (n x n in LDS)
PERTURB = 1e-13        # relative Gaussian perturbation that keeps range-finder blocks of full rank
RSVD_SEED = 90210


def _f64(A):
    return A if A.dtype == torch.float64 else ops.cast(A, torch.float64)


def svd_small(A, batch=None):
    """Thin SVD of (B, m, n) float64, n <= m, m*n small enough for LDS.
    Returns (u (B,m,n), s (B,n) descending, vh (B,n,n))."""
    A = ops.contiguous(_f64(A))
    B, m, n = A.shape
    s = torch.empty((B, n), dtype=torch.float64, device=A.device)
    U = torch.empty((B, m, n), dtype=torch.float64, device=A.device)
    V = torch.empty((B, n, n), dtype=torch.float64, device=A.device)
    info = torch.empty(B, dtype=torch.int32, device=A.device)
    call("trg_gesvj_f64", B, m, n, ptr(A), m * n, ptr(s), ptr(U), ptr(V), 60, ptr(info), stream_ptr())
    ops.note(info=info)
    return U, s, V.transpose(1, 2)


def _fits_small(m, n):
    return n <= m <= 256 and m * (n | 1) * 8 + 16384 <= 163840


def orthonormalize(X, perturb=PERTURB, stream_id=7, passes=3, return_info=False):
    """Orthonormal basis (B, m, l) of the columns of X (B, m, l), l <= m.
    A relative Gaussian perturbation keeps rank-deficient blocks of full
    column rank (directions below perturb * max|X| are noise either way).
    passes=1: one shifted CholeskyQR pass, a basis of the same span that is
    only as well conditioned as the shift allows (not enough to feed a power
    iteration on a matrix of wide spectrum; the default is CholeskyQR3)."""
    X = ops.contiguous(_f64(X))
    B, m, l = X.shape
    if perturb > 0:
        c = ops.maxabs(X, batch=B)
        Om = ops.gaussian((m, l), seed=RSVD_SEED, device=X.device, stream_id=stream_id, batch=B, stream_stride=0)
        Om = Om.view(B, m, l)
        call("trg_axpy_dev_f64", B, m * l, float(perturb), ptr(c), ptr(Om), m * l, ptr(X), m * l, stream_ptr())
    lib = _lib.load()
    wsb = lib.trg_orthonormalize_workspace(B, m, l)
    ws = torch.empty(wsb // 8 + 1, dtype=torch.float64, device=X.device)
    Q = torch.empty((B, m, l), dtype=torch.float64, device=X.device)
    info = torch.empty(B, dtype=torch.int32, device=X.device)
    call("trg_orthonormalize_f64", B, m, l, ptr(X), m * l, ptr(Q), ptr(ws), ws.numel() * 8, passes, ptr(info),
         stream_ptr())
    # info[b] != 0: a Cholesky pivot fell to the rounding level (the block is
    # rank deficient beyond what the perturbation repairs).  A caller that asks
    # for the flags handles them; otherwise they go to the solve's health word
    if not return_info:
        ops.note(info=info)
    return (Q, info) if return_info else Q


def svd_thin(A):
    """Thin SVD of (B, m, n) float64 with min(m, n) <= 128.
    Returns (u (B,m,r), s (B,r), vh (B,r,n)), r = min(m, n)."""
    A = _f64(A)
    B, m, n = A.shape
    if m < n:
        u, s, vh = svd_thin(A.transpose(1, 2))
        return vh.transpose(1, 2), s, u.transpose(1, 2)
    if _fits_small(m, n):
        return svd_small(A)
    if n > SMALL_MAX:
        raise TrgError(f"svd_thin: short side {n} above {SMALL_MAX}")
    Q = orthonormalize(A)                          # (B, m, n)
    R = gemm(Q.transpose(1, 2), ops.contiguous(A))  # (B, n, n)
    uR, s, vh = svd_small(R)
    return gemm(Q, uR), s, vh


def svd_topk(A, l, q, seed=RSVD_SEED, Aop=None, ATop=None, shape=None):
    """Leading l singular triplets of A (B, m, n) (or of an operator given by
    Aop(X) = A X and ATop(Y) = A^T Y with shape=(B, m, n)).  Randomized range
    finder with q power iterations; returns (u (B,m,l), s (B,l), vh (B,l,n))."""
    if Aop is None:
        A = _f64(A)
        B, m, n = A.shape
        Aop = lambda X: gemm(A, X)
        ATop = lambda Y: gemm(A.transpose(1, 2), Y)
    else:
        B, m, n = shape
    Om = ops.gaussian((n, l), seed=seed, device="cuda", batch=B, stream_stride=0).view(B, n, l)
    # every iterate is fully orthonormalised (CholeskyQR3): a single shifted
    # pass leaves a basis whose next product with a matrix of wide spectrum
    # is too ill-conditioned for the unshifted passes of the final block
    Q = orthonormalize(Aop(Om))
    for _ in range(q):
        Z = orthonormalize(ATop(Q))
        Q = orthonormalize(Aop(Z))
    Bt = ATop(Q)                                   # (B, n, l) = (Q^T A)^T
    v, s, uh = svd_thin(Bt)                        # Bt = v s uh  ->  Q^T A = uh^T s v^T
    return gemm(Q, uh.transpose(1, 2)), s, v.transpose(1, 2)


def chol_psd_upper(A, nb=128, tol=1e-15):
    """Upper R (B, n, n) with R^T R = A for symmetric PSD float64 A; rows whose
    Schur pivot falls below tol * max_i A_ii are zeroed.  Blocked right-looking
    factorisation: potrf_psd on nb x nb diagonal blocks, GEMM panel and
    trailing updates."""
    A = _f64(A)
    B, n, _ = A.shape
    W = ops.contiguous(A)
    diag = ops.contiguous(torch.as_strided(W, (B, n), (n * n, n + 1)))
    scale = ops.maxabs(diag, batch=B)
    R = torch.empty((B, n, n), dtype=torch.float64, device=A.device)
    ops.fill_(R, 0.0)
    for k0 in range(0, n, nb):
        k1 = min(n, k0 + nb)
        b = k1 - k0
        Akk = ops.contiguous(W[:, k0:k1, k0:k1])
        Rkk = torch.empty((B, b, b), dtype=torch.float64, device=A.device)
        Xkk = torch.empty((B, b, b), dtype=torch.float64, device=A.device)
        call("trg_potrf_psd_f64", B, b, ptr(Akk), b * b, float(tol), ptr(scale), ptr(Rkk), ptr(Xkk), stream_ptr())
        ops.copy_into(R[:, k0:k1, k0:k1], Rkk)
        if k1 < n:
            R12 = gemm(Xkk.transpose(1, 2), ops.contiguous(W[:, k0:k1, k1:]))      # (B, b, n-k1)
            ops.copy_into(R[:, k0:k1, k1:], R12)
            gemm(R12.transpose(1, 2), R12, W[:, k1:, k1:], alpha=-1.0, beta=1.0)
    return R


def eigh_small(A):
    """Full eigendecomposition (ascending) of (B, n, n) symmetric float64, n <= 256."""
    B, n, _ = A.shape
    w, V, _ = ops.syevj(ops.contiguous(_f64(A)), batch=B)
    return w, V


def truncate(s, k, degeneracy_eps, truncate_eps):
    """utility/truncated_svd.py:17-78 on a descending device vector s (n,).
    Returns (chi: int, error: float); the arithmetic runs on the device."""
    if degeneracy_eps == 0 and truncate_eps == 0:
        return int(k), 0.0
    n = s.numel()
    if k > n:
        return n, 0.0
    s = ops.contiguous(_f64(s)).view(-1)
    chi = torch.empty(1, dtype=torch.int32, device=s.device)
    err = torch.empty(1, dtype=torch.float64, device=s.device)
    call("trg_truncate_rank", n, ptr(s), int(k), float(degeneracy_eps), float(truncate_eps), ptr(chi), ptr(err),
         stream_ptr())
    return int(chi.item()), float(err.item())


def truncate_known(s, tail_sq, total_sq, k, degeneracy_eps, truncate_eps):
    """truncate() over the leading values s of a longer spectrum whose
    remaining squares sum to tail_sq (device scalars tail_sq and total_sq =
    the sum of all squares): the rest enters as one trailing value
    sqrt(tail_sq), so the tail-norm rule of utility/truncated_svd.py:57-61
    sees the exact total; the degeneracy rule only looks at positions <= k,
    which must lie inside s.  A tail within rounding of the total (the
    difference of two sums that agree to the last bits) counts as zero, so an
    exactly rank-deficient spectrum keeps the reference's cut."""
    buf = torch.empty(2, dtype=torch.float64, device=s.device)
    ops.copy_into(buf[0:1], _f64(tail_sq).reshape(1))
    ops.copy_into(buf[1:2], _f64(total_sq).reshape(1))
    t2, tot = buf.tolist()
    if t2 <= 64 * 2.220446049250313e-16 * abs(tot):
        return truncate(s, k, degeneracy_eps, truncate_eps)
    ext = torch.empty(s.numel() + 1, dtype=torch.float64, device=s.device)
    ops.copy_into(ext[:-1], s)
    ops.fill_(ext[-1:], math.sqrt(t2))
    return truncate(ext, k, degeneracy_eps, truncate_eps)


def frobenius_sq(P):
    """||P||_F^2 of a float64 matrix as a (1, 1) device tensor (row norms by a
    batched dot, then their sum)."""
    m, n = P.shape
    Pc = ops.contiguous(P)
    rows = gemm(Pc.view(m, 1, n), Pc.view(m, n, 1)).view(1, m)
    ones = ops.fill_(torch.empty((m, 1), dtype=torch.float64, device=P.device), 1.0)
    return gemm(rows, ones)


def sum_sq(s):
    """sum s_i^2 of a device vector as a (1, 1) device tensor."""
    sv = ops.contiguous(_f64(s)).view(1, -1)
    return gemm(sv, ops.contiguous(sv.transpose(0, 1)))


def truncate_with_tail(s, P, k, degeneracy_eps, truncate_eps):
    """truncate() when only the leading singular values s of the float64
    matrix P are known: the discarded mass ||P||_F^2 - sum s_i^2 is the tail."""
    tot = frobenius_sq(P)
    tail = ops.contiguous(tot)
    one = ops.fill_(torch.empty((1, 1), dtype=torch.float64, device=s.device), 1.0)
    gemm(one, sum_sq(s), tail, alpha=-1.0, beta=1.0)
    return truncate_known(s, tail, tot, k, degeneracy_eps, truncate_eps)


def count_above(s, eps):
    """#(s / max s > eps) for a descending device vector (the eps cut of
    trg/ATRG_3d_new.py:130-134); uses the truncate kernel with
    degeneracy_eps < 0 so that only the tail rule applies."""
    s = ops.contiguous(_f64(s)).view(-1)
    n = s.numel()
    # number of leading entries with s_i / s_0 > eps, read back as one int
    flags = torch.empty(n, dtype=torch.float64, device=s.device)
    ops.copy_into(flags, s)
    c = ops.maxabs(flags)
    ops.normalize_(flags, c)
    # scan on the host of the already-normalised flags is a rank decision; do it on a copy
    vals = flags.cpu()
    return int((vals > eps).sum().item())
