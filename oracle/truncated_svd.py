"""NumPy restatement of utility/truncated_svd.py (test infrastructure only).

Follows utility/truncated_svd.py:17-208 of the reference.
"""
from functools import reduce
from operator import mul

import numpy as np


def truncate(s0, k, degeneracy_eps, truncate_eps):
    """Rank selection from descending singular/eigen values.

    Follows utility/truncated_svd.py:17-78:
      * both eps zero -> (k, 0) even when k > len(s0)  (:36-37)
      * k > len(s0)   -> (len(s0), 0)                   (:38-39)
      * diff_i = |s_i - s_{i-1}| / s_i on s/s[0], diff[len] = 1  (:46-49)
      * chi1 = largest position <= k outside a multiplet     (:52-54)
      * chi2 = #(tail-norm errors > truncate_eps) in [:chi1] (:57-61)
      * chi  = first non-degenerate position in [chi2, chi1] (:64-71)
    """
    s0 = np.asarray(s0)
    n = len(s0)
    if degeneracy_eps == 0 and truncate_eps == 0:
        return k, 0
    if k > n:
        return n, 0

    s = np.array(s0, dtype=np.float64 if not np.iscomplexobj(s0) else s0.dtype, copy=True)
    s = s / s[0]

    diff = np.zeros(n + 1)
    with np.errstate(divide="ignore", invalid="ignore"):
        diff[1:n] = np.abs(s[1:] - s[:n - 1]) / np.abs(s[1:])
    diff[n] = 1.0
    delta = np.where(diff > degeneracy_eps, 1, 0)

    chi1 = k
    for chi1 in range(k, 0, -1):
        if delta[chi1] == 1:
            break

    s_sqr = np.abs(s) ** 2
    norms_sqr = np.sum(s_sqr)
    # tail norms sqrt(sum_{j>=i} s_j^2 / sum s^2), i = 0..n-1
    tail = np.cumsum(s_sqr[::-1])[::-1]
    errors = np.sqrt(tail / norms_sqr)
    errors_chi1 = errors[:chi1]
    chi2 = int(np.count_nonzero(errors_chi1 > truncate_eps))

    if chi2 == chi1:
        chi = chi1
    elif chi2 < chi1:
        chi = chi2
        for chi in range(chi2, chi1 + 1):
            if delta[chi] == 1:
                break
    else:  # pragma: no cover - impossible by construction, kept as in :70-71
        raise ValueError("chi2 should be less than or equal chi1")

    error = 0.0 if n == chi else float(errors[chi])
    return chi, error


def _to_matrix(A, shape):
    llen = len(shape[0])
    axes = tuple(shape[0] + shape[1])
    newshape = tuple(A.shape[i] for i in axes)
    mshape = (reduce(mul, newshape[:llen], 1), reduce(mul, newshape[llen:], 1))
    return np.transpose(A, axes).reshape(mshape), newshape, llen


def svd(A, shape, k, truncate_eps=0, degeneracy_eps=0, split=False, return_err=False):
    """Follows utility/truncated_svd.py:80-149 (transpose, reshape, SVD, truncate)."""
    M, newshape, llen = _to_matrix(np.asarray(A), shape)
    u, s, vh = np.linalg.svd(M, full_matrices=False)
    chi, error = truncate(s, k, degeneracy_eps, truncate_eps)
    u = u[:, :chi]
    s = s[:chi]
    vh = vh[:chi, :]
    chi = len(s)
    if split:
        rs = np.sqrt(s)
        us = (u * rs[None, :]).reshape(newshape[:llen] + (chi,))
        svh = (vh * rs[:, None]).reshape((chi,) + newshape[llen:])
        return (us, svh, s, error) if return_err else (us, svh, s)
    u = u.reshape(newshape[:llen] + (chi,))
    vh = vh.reshape((chi,) + newshape[llen:])
    return (u, s, vh, error) if return_err else (u, s, vh)


def eigh(A, shape, k, truncate_eps=0, degeneracy_eps=5e-2, return_err=False):
    """Follows utility/truncated_svd.py:151-208 (eigh, |e|, descending, truncate)."""
    M, newshape, llen = _to_matrix(np.asarray(A), shape)
    e, u = np.linalg.eigh(M)
    e = np.abs(e)[::-1]
    u = u[:, ::-1]
    chi, error = truncate(e, k, degeneracy_eps, truncate_eps)
    u = u[:, :chi]
    e = e[:chi]
    u = u.reshape(newshape[:llen] + (u.shape[1],))
    return (e, u, error) if return_err else (e, u)
