"""Pairwise tensor contraction on libtrg (transpose-transpose-GEMM-transpose).

``contract("abc,cbd->ad", A, B)`` groups the labels into batch / free /
contracted sets, views each operand as a (batch, rows, cols) matrix when its
strides allow it (no copy), otherwise materialises one permuted copy with
the libtrg strided-copy kernel, runs one batched libtrg GEMM and permutes the
result into the requested label order.  Repeated labels inside an operand
(traces) are not supported; ``ops.trace_pairs`` covers those.
"""
import math

import torch

from . import ops
from ._lib import TrgError


def _group(t, labels, group):
    """(size, innermost stride) of `group` in t if the labels are mergeable."""
    dims = []
    for c in group:
        i = labels.index(c)
        if t.shape[i] != 1:
            dims.append((t.shape[i], t.stride(i)))
    if not dims:
        return 1, None
    for (n0, s0), (n1, s1) in zip(dims[:-1], dims[1:]):
        if s0 != s1 * n1:
            return None
    return math.prod(n for n, _ in dims), dims[-1][1]


def _as_matrix(t, labels, batch, rows, cols):
    """View t as (batch, rows, cols) with one unit matrix stride, or None."""
    gb, gr, gc = _group(t, labels, batch), _group(t, labels, rows), _group(t, labels, cols)
    if gb is None or gr is None or gc is None:
        return None
    (nb, sb), (nr, sr), (nc, sc) = gb, gr, gc
    if sr is None and sc is None:
        sr, sc = nc, 1
    elif sr is None:
        sr = 1 if sc != 1 else nc
    elif sc is None:
        sc = 1 if sr != 1 else nr
    if sr != 1 and sc != 1:
        return None
    sb = 0 if sb is None else sb
    if nb == 1:
        sb = 0
    return torch.as_strided(t, (nb, nr, nc), (sb, sr, sc), t.storage_offset())


def contract(spec, A, B, compute=None, tag=None):
    ins, out = spec.replace(" ", "").split("->")
    la, lb = ins.split(",")
    if len(set(la)) != len(la) or len(set(lb)) != len(lb):
        raise TrgError(f"repeated label inside an operand in {spec!r}")
    batch = [c for c in out if c in la and c in lb]
    contr = [c for c in la if c in lb and c not in out]
    fa = [c for c in out if c in la and c not in lb]
    fb = [c for c in out if c in lb and c not in la]
    for c in la + lb:
        if c not in out and c not in contr:
            raise TrgError(f"label {c!r} summed over a single operand is not supported ({spec!r})")
    size = {}
    for lab, t in ((la, A), (lb, B)):
        for c, n in zip(lab, t.shape):
            if size.setdefault(c, n) != n:
                raise TrgError(f"size mismatch on label {c!r} in {spec!r}")
    # contracted order: keep the larger operand's own order
    if B.numel() > A.numel():
        contr = [c for c in lb if c in contr]
    Am = _as_matrix(A, la, batch, fa, contr)
    if Am is None:
        perm = [la.index(c) for c in batch + fa + contr]
        Ac = ops.permute(A, perm)
        Am = _as_matrix(Ac, batch + fa + contr, batch, fa, contr)
    Bm = _as_matrix(B, lb, batch, contr, fb)
    if Bm is None:
        perm = [lb.index(c) for c in batch + contr + fb]
        Bc = ops.permute(B, perm)
        Bm = _as_matrix(Bc, batch + contr + fb, batch, contr, fb)
    nb = math.prod(size[c] for c in batch) if batch else 1
    C = torch.empty([size[c] for c in batch + fa + fb], dtype=A.dtype, device=A.device)
    Cm = C.view(nb, Am.shape[1], Bm.shape[2])
    if Am.shape[0] == 1 and nb > 1:
        Am = Am.expand(nb, -1, -1)
    ops.gemm(Am, Bm, Cm, compute=compute, tag=tag)
    produced = batch + fa + fb
    if produced == list(out):
        return C
    return ops.permute(C, [produced.index(c) for c in out])


def contract_path(spec, *operands, compute=None):
    """Left-to-right pairwise evaluation of a multi-operand einsum spec."""
    ins, out = spec.replace(" ", "").split("->")
    labels = ins.split(",")
    if len(labels) != len(operands):
        raise TrgError("operand count does not match the spec")
    cur, cl = operands[0], labels[0]
    for k in range(1, len(operands)):
        rest = "".join(labels[k + 1:]) + out
        nl = labels[k]
        keep = [c for c in dict.fromkeys(cl + nl) if c in rest]
        cur = contract(f"{cl},{nl}->{''.join(keep)}", cur, operands[k], compute=compute)
        cl = "".join(keep)
    if cl != out:
        cur = ops.permute(cur, [cl.index(c) for c in out])
    return cur
