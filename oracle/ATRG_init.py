"""NumPy restatement of tensor_init/ATRG_init.py (test infrastructure only).

initial_tensor_for_ATRG (:7-254) factorises the initial tensor of a
quadrature-sampled model,

    T_{T X.. t x..} = sum_a w_a J_a At_{aT} As_{aX} .. Bt_{ta} Bs_{xa} ..,

as the matrix M[(T X..), (t x..)] by a randomized SVD: Gaussian sketch of
l = k + p columns, QR, q rounds of (M^T Q, QR, M Q, QR), X = Q^T M, SVD of X,
U = Q u.  The sketch is the Philox stream of ``oracle.philox`` instead of
CuPy's cuRAND RandomState (not reproducible without CUDA); real inputs only
(the reference conjugates, the identity here).

Pinned by tests/test_atrg_init.py: with a sketch at least the rank of M
(l >= number of quadrature points) the result is the dense truncated SVD of
M built by an einsum of the same factors.
"""
import numpy as np

from .philox import standard_normal

_LEGS = "XYZ"


def _specs(dim):
    """The reference's index strings for dim = 2, 3, 4 (:83-131)."""
    up = "T" + _LEGS[:dim - 1]
    lo = up.lower()
    fa = ",".join(["a" + c for c in up])                 # At, As, ..
    fb = ",".join([c + "a" for c in lo])                 # Bt, Bs, ..
    ty = f"a,a,{fa},{fb},{lo}l->{up}l"                   # M Y
    tdq = f"a,a,{fb},{fa},{up}l->{lo}l"                  # M^T Q
    qdt = f"a,a,{fb},{fa},{up}l->l{lo}"                  # Q^T M
    return ty, tdq, qdt


def _ops(dim, At, As, Bt, Bs):
    return [At] + [As] * (dim - 1), [Bt] + [Bs] * (dim - 1)


def _qr(Y):
    return np.linalg.qr(Y)[0]


def initial_tensor_for_ATRG(dim, J, w, As, Bs, At, Bt, k, p, q, seed=1234):
    """:7-254 -> (u (chi,)*dim + (k,), s (k,), vh (k,) + (chi,)*dim)."""
    if any(np.iscomplexobj(x) for x in (J, w, As, Bs, At, Bt)):
        raise ValueError("real inputs only")
    ty, tdq, qdt = _specs(dim)
    l = k + p
    chi, _ = Bs.shape
    A, B = _ops(dim, At, As, Bt, Bs)
    n = chi ** dim
    Y = standard_normal((n, l), seed).reshape((chi,) * dim + (l,))
    Y = np.einsum(ty, w, J, *A, *B, Y, optimize=True).reshape(n, l)
    Q = _qr(Y).reshape((chi,) * dim + (-1,))
    for _ in range(q):
        Y = np.einsum(tdq, w, J, *B, *A, Q, optimize=True).reshape(n, -1)
        Q = _qr(Y).reshape((chi,) * dim + (-1,))
        Y = np.einsum(ty, w, J, *A, *B, Q, optimize=True).reshape(n, -1)
        Q = _qr(Y).reshape((chi,) * dim + (-1,))
    X = np.einsum(qdt, w, J, *B, *A, Q, optimize=True).reshape(Q.shape[-1], n)
    u, s, vh = np.linalg.svd(X, full_matrices=False)
    u, s, vh = u[:, :k], s[:k], vh[:k]
    u = np.tensordot(Q, u, axes=([dim], [0]))
    return u, s, vh.reshape((vh.shape[0],) + (chi,) * dim)


def dense_matrix(dim, J, w, As, Bs, At, Bt):
    """M[(T X..), (t x..)] of the same factors, for the checks."""
    up = "T" + _LEGS[:dim - 1]
    lo = up.lower()
    fa = ",".join(["a" + c for c in up])
    fb = ",".join([c + "a" for c in lo])
    A, B = _ops(dim, At, As, Bt, Bs)
    M = np.einsum(f"a,a,{fa},{fb}->{up}{lo}", w, J, *A, *B, optimize=True)
    n = Bs.shape[0] ** dim
    return M.reshape(n, n)
