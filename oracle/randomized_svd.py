"""NumPy restatement of utility/randomized_svd.py (test infrastructure only).

Sketch draws use the Philox stream of ``oracle.philox`` instead of CuPy's
cuRAND ``RandomState`` (not reproducible without CUDA); everything else
follows the reference line by line in meaning.
"""
import numpy as np

from .philox import standard_normal


def _qr(Y):
    q, _ = np.linalg.qr(Y)
    return q


def _herm(A):
    return np.conj(A.T) if np.iscomplexobj(A) else A.T


def __QR_iteration__(A, Y, n_power_iter):
    """utility/randomized_svd.py:83-95."""
    Q = _qr(Y)
    for _ in range(n_power_iter):
        Q = _qr(_herm(A) @ Q)
        Q = _qr(A @ Q)
    return Q


def __power_iteration__(A, Y, n_power_iter):
    """utility/randomized_svd.py:72-81."""
    for _ in range(n_power_iter):
        Y = A @ (_herm(A) @ Y)
    return _qr(Y)


def rsvd(A, k, n_oversamples=0, n_power_iter=0, iterator="QR", seed=1234):
    """utility/randomized_svd.py:5-37: Gaussian sketch, QR/power iteration, small SVD."""
    A = np.asarray(A)
    M, N = A.shape
    l = k + n_oversamples
    Y = A @ standard_normal((N, l), seed)
    if iterator == "power":
        Q = __power_iteration__(A, Y, n_power_iter)
    else:
        Q = __QR_iteration__(A, Y, n_power_iter)
    B = _herm(Q) @ A
    u_t, s, vh = np.linalg.svd(B, full_matrices=False)
    u = Q @ u_t
    return u[:, :k], s[:k], vh[:k, :]


def rsvd2(A, k, n_oversamples=0, seed=1, del_A=False):
    """utility/randomized_svd.py:39-70: one-pass sketch + QR + small SVD."""
    A = np.asarray(A)
    M, N = A.shape
    l = k + n_oversamples
    Y = A @ standard_normal((N, l), seed)
    Q = _qr(Y)
    B = _herm(Q) @ A
    u_t, s, vh = np.linalg.svd(B, full_matrices=False)
    u = Q @ u_t
    return u[:, :k], s[:k], vh[:k, :]


def rsvd_for_3dATRG_tensor_init(A, B, Dcut, n_oversamples=10, n_power_iter=0, iterator="QR", seed=1234):
    """utility/randomized_svd.py:100-162: rSVD of the product A_{abci} B_{idef}.

    Pinned by tests/test_oracle.py: with a sketch that covers the rank the
    result equals the dense truncated SVD of the operator built from the
    reference's own index strings (:119, :125), for both iterators; only the
    reference's cuRAND draw cannot be reproduced (gauge-free checks).
    """
    l = Dcut + n_oversamples
    d = A.shape[:3]
    e = B.shape[1:]
    na, nb = int(np.prod(d)), int(np.prod(e))
    Am = A.reshape(na, A.shape[3])
    Bm = B.reshape(B.shape[0], nb)
    Mop = Am @ Bm
    Y = Mop @ standard_normal((nb, l), seed)
    if iterator == "power":
        for _ in range(n_power_iter):
            Y = Mop @ (_herm(Mop) @ Y)
        Q = _qr(Y)
    else:
        Q = _qr(Y)
        for _ in range(n_power_iter):
            Q = _qr(_herm(Mop) @ Q)
            Q = _qr(Mop @ Q)
    X = _herm(Q) @ Mop
    u, s, vh = np.linalg.svd(X, full_matrices=False)
    u = (Q @ u).reshape(d + (u.shape[1],))
    vh = vh.reshape((vh.shape[0],) + e)
    return u[..., :Dcut], s[:Dcut], vh[:Dcut]


def rsvd_for_3dATRG_tensor_inte(B, C, Dcut, n_oversamples=10, n_power_iter=0, iterator="QR", seed=1234):
    """utility/randomized_svd.py:165-226: rSVD of M_{ijk,abc} = B_{idbc} C_{djka}.

    The reference contracts "iabc,ajkd,dbcl->ijkl" for the first sketch
    (:209) and "idbc,djka,abcl->ijkl" inside the iteration (:202); both are
    the same operator M with the row index (i,j,k) and the column index
    (a,b,c).  Pinned by tests/test_oracle.py: the four index strings of the
    reference are checked: :202, :209 and :222 give M, :196 gives M^dag; and
    with a sketch that covers the rank the result equals the dense
    truncated SVD of M, for both iterators.
    """
    l = Dcut + n_oversamples
    Mop = np.einsum("idbc,djka->ijkabc", B, C).reshape(
        B.shape[0] * C.shape[1] * C.shape[2], C.shape[3] * B.shape[2] * B.shape[3])
    Y = Mop @ standard_normal((Mop.shape[1], l), seed)
    if iterator == "power":
        for _ in range(n_power_iter):
            Y = Mop @ (_herm(Mop) @ Y)
        Q = _qr(Y)
    else:
        Q = _qr(Y)
        for _ in range(n_power_iter):
            Q = _qr(_herm(Mop) @ Q)
            Q = _qr(Mop @ Q)
    X = _herm(Q) @ Mop
    u, s, vh = np.linalg.svd(X, full_matrices=False)
    u = (Q @ u).reshape((B.shape[0], C.shape[1], C.shape[2], u.shape[1]))
    vh = vh.reshape((vh.shape[0], C.shape[3], B.shape[2], B.shape[3]))
    return u[..., :Dcut], s[:Dcut], vh[:Dcut]
