"""Drop-in for utility/truncated_svd.py on the GPU (libtrg decompositions).

Same signatures as utility/truncated_svd.py:17-208.  Limits of the device
decompositions (documented): `svd` computes the exact thin SVD when the
shorter side of the matrix is <= 128 and otherwise the leading
min(128, max(2k, k+8)) singular triplets by randomized subspace iteration
(the discarded mass enters the truncation rule exactly through the
Frobenius norm); `eigh` is exact for n <= 256 and uses the top-k subspace
eigensolver (positive semidefinite input) above that.
"""
from functools import reduce
from operator import mul

import torch

from .. import linalg, ops
from .._io import from_device, to_device
from .._lib import TrgError


def truncate(s0, k, degeneracy_eps, truncate_eps):
    """utility/truncated_svd.py:17-78 -> (chi, error); the rule runs on the device."""
    s, _ = to_device(s0)
    return linalg.truncate(s.reshape(-1), k, degeneracy_eps, truncate_eps)


def _matrix(A, shape):
    llen = len(shape[0])
    axes = tuple(shape[0] + shape[1])
    newshape = tuple(A.shape[i] for i in axes)
    m = reduce(mul, newshape[:llen], 1)
    n = reduce(mul, newshape[llen:], 1)
    return ops.permute(A, axes).reshape(m, n), newshape, llen


def svd(A, shape, k, truncate_eps=0, degeneracy_eps=0, split=False, return_err=False):
    """utility/truncated_svd.py:80-149."""
    Ad, kind = to_device(A)
    M, newshape, llen = _matrix(Ad, shape)
    m, n = M.shape
    if min(m, n) <= linalg.SMALL_MAX:
        u, s, vh = linalg.svd_thin(M[None])
        u, s, vh = u[0], s[0], vh[0]
        chi, error = linalg.truncate(s, k, degeneracy_eps, truncate_eps)
    else:
        kk = min(k, min(m, n))
        l = min(linalg.SMALL_MAX, max(2 * kk, kk + 8))
        if kk >= l:
            raise TrgError(f"svd: rank {k} of a {m}x{n} matrix is beyond the device SVD limits")
        u, s, vh = linalg.svd_topk(ops.cast(M, torch.float64)[None], l, 12)
        u, s, vh = u[0], s[0], vh[0]
        chi, error = linalg.truncate_with_tail(s, ops.cast(M, torch.float64), k, degeneracy_eps, truncate_eps) \
            if (degeneracy_eps or truncate_eps) else (k, 0.0)
    chi = min(chi, s.numel())
    u = ops.contiguous(u[:, :chi])
    s = ops.contiguous(s[:chi])
    vh = ops.contiguous(vh[:chi])
    if split:
        us = ops.contiguous(u)
        ops.scale_rows_(us, s, chi, 1, 1)
        svh = ops.contiguous(vh)
        ops.scale_rows_(svh, s, chi, vh.shape[1], 1)
        out = (from_device(us.view(newshape[:llen] + (chi,)), kind), from_device(svh.view((chi,) + newshape[llen:]), kind),
               from_device(s, kind))
    else:
        out = (from_device(u.view(newshape[:llen] + (chi,)), kind), from_device(s, kind),
               from_device(vh.view((chi,) + newshape[llen:]), kind))
    return out + (error,) if return_err else out


def eigh(A, shape, k, truncate_eps=0, degeneracy_eps=5e-2, return_err=False):
    """utility/truncated_svd.py:151-208 (|e| in reversed ascending order)."""
    Ad, kind = to_device(A)
    M, newshape, llen = _matrix(Ad, shape)
    n = M.shape[0]
    Md = ops.cast(M, torch.float64)
    if n <= 256:
        w, V = linalg.eigh_small(Md[None])
        w, V = w[0], V[0]
        e, U = _reverse(w), _reverse_cols(V)
    else:
        kk = min(k, n)
        l = min(256, max(2 * kk, kk + 8))
        w, V, _, _ = ops.eigh_topk(ops.contiguous(Md), kk, l, 12)
        e, U = _reverse(w[0]), _reverse_cols(V[0])
    e = ops.maxabs(e.view(-1, 1), batch=e.numel())          # |e| element by element
    chi, error = linalg.truncate(e, k, degeneracy_eps, truncate_eps)
    chi = min(chi, e.numel())
    U = ops.contiguous(U[:, :chi]).view(newshape[:llen] + (chi,))
    out = (from_device(ops.contiguous(e[:chi]), kind), from_device(U, kind))
    return out + (error,) if return_err else out


def _reverse(v):
    return ops.flip(ops.contiguous(v), 0)


def _reverse_cols(V):
    return ops.flip(ops.contiguous(V), 1)
