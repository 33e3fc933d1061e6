"""Drop-in for tensor_init/ATRG_init.py: the ATRG initial tensor of a
quadrature-sampled model, factorised on the GPU.

initial_tensor_for_ATRG (:7-254) treats

    T_{T X.. t x..} = sum_a w_a J_a At_{aT} As_{aX} .. Bt_{ta} Bs_{xa} ..

as the matrix M[(T X..), (t x..)] = P diag(w J) Q^T, with P and Q the
Khatri-Rao products of the leg factors (column a = the outer product of the
a-th rows of At, As, .. resp. Bt, Bs, ..), and factorises it by a randomized
SVD.  Here P^T and Q^T (one row per quadrature point) are built by batched
rank-one GEMMs, M and M^T are applied as two GEMMs around a row scaling, and
the range finder is the library's: Philox sketch keyed by `seed` (cuRAND's
RandomState stream cannot be reproduced off CUDA), shifted CholeskyQR3 for
each QR (same column span as Householder QR), the one-sided Jacobi SVD of the
l x chi^dim factor (l = k + p <= 128).  Real inputs only: the reference's
conjugations are the identity on them, and complex inputs raise.
"""
import torch

from .. import ops
from .._io import from_device, to_device
from .._lib import TrgError
from ..utility.randomized_svd import _operator_rsvd


def _khatri_rao(first, rest):
    """R[a, (i0 i1 ..)] = first[a, i0] * rest[0][a, i1] * ..  for (n_a, chi) factors."""
    R = ops.contiguous(first)
    na = R.shape[0]
    for F in rest:
        R = ops.gemm(R.reshape(na, -1, 1), ops.contiguous(F).reshape(na, 1, -1)).reshape(na, -1)
    return R


def initial_tensor_for_ATRG(dim, J, w, As, Bs, At, Bt, k, p, q, seed=1234):
    """:7-254 -> (u (chi,)*dim + (k,), s (k,), vh (k,) + (chi,)*dim), in the
    array kind of As.  dim = 2, 3 or 4; As, At are (n_a, chi) (As_{aX}),
    Bs, Bt are (chi, n_a) (Bs_{xa}); J and w have n_a entries."""
    if dim not in (2, 3, 4):
        raise TrgError(f"initial_tensor_for_ATRG: dim {dim} is not 2, 3 or 4")
    dev = []
    kind = None
    for x in (J, w, As, Bs, At, Bt):
        d, kd = to_device(x)               # raises on a non-zero imaginary part
        dev.append(ops.cast(d, torch.float64))
        kind = kind or (kd if x is As else None)
    Jd, wd, Asd, Bsd, Atd, Btd = dev
    chi, na = Bsd.shape
    if Asd.shape != (na, chi) or Atd.shape != (na, chi) or Btd.shape != (chi, na) or \
            Jd.numel() != na or wd.numel() != na:
        raise TrgError("initial_tensor_for_ATRG: As, At must be (n_a, chi), Bs, Bt (chi, n_a), J and w n_a long")
    l = k + p
    if l > 128:
        raise TrgError(f"initial_tensor_for_ATRG: k + p = {l} is above the 128-column range finder")
    c = ops.contiguous(wd.reshape(-1))
    ops.scale_rows_(c, ops.contiguous(Jd.reshape(-1)), na, 1, 0)           # c_a = w_a J_a
    Pt = _khatri_rao(Atd, [Asd] * (dim - 1))                             # (n_a, chi^dim) rows (T X ..)
    BtT, BsT = ops.contiguous(Btd.transpose(0, 1)), ops.contiguous(Bsd.transpose(0, 1))
    Qt = _khatri_rao(BtT, [BsT] * (dim - 1))                             # (n_a, chi^dim) cols (t x ..)

    def scaled(Z):
        ops.scale_rows_(Z, c, na, Z.shape[1], 0)
        return Z

    Mop = lambda X: ops.gemm(Pt.transpose(0, 1), scaled(ops.gemm(Qt, ops.contiguous(X))))
    MTop = lambda Y: ops.gemm(Qt.transpose(0, 1), scaled(ops.gemm(Pt, ops.contiguous(Y))))
    return _operator_rsvd(Mop, MTop, (chi,) * dim, (chi,) * dim, k, l, q, "QR", seed, kind)
