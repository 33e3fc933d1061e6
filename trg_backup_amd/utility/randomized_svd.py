"""Drop-in for utility/randomized_svd.py on the GPU.

Same signatures as utility/randomized_svd.py:5-226.  The Gaussian sketch is
Philox4x32-10 keyed by `seed` (cuRAND's RandomState stream cannot be
reproduced off CUDA); QR steps are shifted CholeskyQR3 (same column span as
Householder QR); the small SVD is the libtrg one-sided Jacobi SVD.
`seed` may be an int or an object with a `seed` attribute.
"""
import torch

from .. import linalg, ops
from .._lib import TrgError
from .._io import from_device, to_device
from ..contract import contract


def _seed(seed):
    if isinstance(seed, int):
        return seed
    return int(getattr(seed, "seed", 1234))


def _finish(Q, Amul_T, k, kind):
    """B = Q^T A via (A^T Q)^T, thin SVD of B, u = Q u~ -> top k."""
    Bt = Amul_T(Q)                          # (N, l)
    v, s, uh = linalg.svd_thin(Bt[None])    # Bt = v s uh -> B = uh^T s v^T
    u = ops.gemm(Q, ops.contiguous(uh[0].transpose(0, 1)))
    vh = ops.contiguous(v[0].transpose(0, 1))
    k = min(k, s.shape[1])
    return (from_device(ops.contiguous(u[:, :k]), kind), from_device(ops.contiguous(s[0, :k]), kind),
            from_device(ops.contiguous(vh[:k]), kind))


def _orth(Y):
    """Orthonormal basis of the sketch's columns.  A sketch wider than the
    rank of the operator can break the range finder's Cholesky down at the
    rounding level (flagged by the library); it is then repaired with a
    perturbation at 1e-10 of its largest entry (directions that small are
    noise either way)."""
    X = ops.cast(Y, torch.float64)[None]
    Q, info = linalg.orthonormalize(X, return_info=True)
    if int(info.max().item()) != 0:
        Q, info = linalg.orthonormalize(X, perturb=1e-10, return_info=True)
        if int(info.max().item()) != 0:
            raise TrgError("randomized SVD: Cholesky breakdown on a rank-deficient sketch")
    return Q[0]


def rsvd(A, k, n_oversamples=0, n_power_iter=0, iterator="QR", seed=1234):
    """utility/randomized_svd.py:5-37."""
    Ad, kind = to_device(A)
    Ad = ops.cast(Ad, torch.float64)
    M, N = Ad.shape
    l = k + n_oversamples
    Om = ops.gaussian((N, l), seed=_seed(seed), device=Ad.device)
    Y = ops.gemm(Ad, Om)
    AT = lambda X: ops.gemm(Ad.transpose(0, 1), X)
    if iterator == "power":
        for _ in range(n_power_iter):
            Y = ops.gemm(Ad, AT(Y))
        Q = _orth(Y)
    else:
        Q = _orth(Y)
        for _ in range(n_power_iter):
            Q = _orth(AT(Q))
            Q = _orth(ops.gemm(Ad, Q))
    return _finish(Q, AT, k, kind)


def rsvd2(A, k, n_oversamples=0, seed=1, del_A=False):
    """utility/randomized_svd.py:39-70 (one pass)."""
    Ad, kind = to_device(A)
    Ad = ops.cast(Ad, torch.float64)
    M, N = Ad.shape
    Om = ops.gaussian((N, k + n_oversamples), seed=_seed(seed), device=Ad.device)
    Q = _orth(ops.gemm(Ad, Om))
    return _finish(Q, lambda X: ops.gemm(Ad.transpose(0, 1), X), k, kind)


def __QR_iteration__(A, Y, n_power_iter):
    """utility/randomized_svd.py:83-95."""
    Ad, kind = to_device(A)
    Yd, _ = to_device(Y)
    Ad = ops.cast(Ad, torch.float64)
    Q = _orth(Yd)
    for _ in range(n_power_iter):
        Q = _orth(ops.gemm(Ad.transpose(0, 1), Q))
        Q = _orth(ops.gemm(Ad, Q))
    return from_device(Q, kind)


def __power_iteration__(A, Y, n_power_iter):
    """utility/randomized_svd.py:72-81."""
    Ad, kind = to_device(A)
    Yd, _ = to_device(Y)
    Ad, Yd = ops.cast(Ad, torch.float64), ops.cast(Yd, torch.float64)
    for _ in range(n_power_iter):
        Yd = ops.gemm(Ad, ops.gemm(Ad.transpose(0, 1), Yd))
    return from_device(_orth(Yd), kind)


def _operator_rsvd(Mop, MTop, rows_shape, cols_shape, Dcut, l, n_power_iter, iterator, seed, kind):
    nr = 1
    for d in rows_shape:
        nr *= d
    nc = 1
    for d in cols_shape:
        nc *= d
    Om = ops.gaussian((nc, l), seed=_seed(seed), device="cuda")
    Y = Mop(Om)
    if iterator == "power":
        for _ in range(n_power_iter):
            Y = Mop(MTop(Y))
        Q = _orth(Y)
    else:
        Q = _orth(Y)
        for _ in range(n_power_iter):
            Q = _orth(MTop(Q))
            Q = _orth(Mop(Q))
    u, s, vh = _finish(Q, MTop, Dcut, "torch")
    u = u.reshape(tuple(rows_shape) + (u.shape[1],))
    vh = vh.reshape((vh.shape[0],) + tuple(cols_shape))
    return from_device(u, kind), from_device(s, kind), from_device(vh, kind)


def rsvd_for_3dATRG_tensor_init(A, B, Dcut, n_oversamples=10, n_power_iter=0, iterator="QR", seed=1234):
    """utility/randomized_svd.py:100-162: rSVD of sum_i A_{abci} B_{idef}."""
    (Ad, kind), (Bd, _) = to_device(A), to_device(B)
    Ad, Bd = ops.cast(Ad, torch.float64), ops.cast(Bd, torch.float64)
    r = Ad.shape[:3]
    c = Bd.shape[1:]
    Am = Ad.reshape(-1, Ad.shape[3])
    Bm = Bd.reshape(Bd.shape[0], -1)
    Mop = lambda X: ops.gemm(Am, ops.gemm(Bm, X))
    MTop = lambda X: ops.gemm(Bm.transpose(0, 1), ops.gemm(Am.transpose(0, 1), X))
    return _operator_rsvd(Mop, MTop, r, c, Dcut, Dcut + n_oversamples, n_power_iter, iterator, seed, kind)


def rsvd_for_3dATRG_tensor_inte(B, C, Dcut, n_oversamples=10, n_power_iter=0, iterator="QR", seed=1234):
    """utility/randomized_svd.py:165-226: rSVD of M_{ijk,abc} = sum_d B_{idbc} C_{djka}."""
    (Bd, kind), (Cd, _) = to_device(B), to_device(C)
    Bd, Cd = ops.cast(Bd, torch.float64), ops.cast(Cd, torch.float64)
    ni, nd, nbb, ncc = Bd.shape
    _, nj, nk, na = Cd.shape

    def Mop(X):                                   # X[(a,b,c), l] -> [(i,j,k), l]
        Xt = X.reshape(na, nbb, ncc, -1)
        W = contract("idbc,abcl->idal", Bd, Xt)
        return contract("djka,idal->ijkl", Cd, W).reshape(ni * nj * nk, -1)

    def MTop(Y):                                  # Y[(i,j,k), l] -> [(a,b,c), l]
        Yt = Y.reshape(ni, nj, nk, -1)
        W = contract("djka,ijkl->dail", Cd, Yt)
        return contract("idbc,dail->abcl", Bd, W).reshape(na * nbb * ncc, -1)

    return _operator_rsvd(Mop, MTop, (ni, nj, nk), (na, nbb, ncc), Dcut, Dcut + n_oversamples, n_power_iter,
                          iterator, seed, kind)
