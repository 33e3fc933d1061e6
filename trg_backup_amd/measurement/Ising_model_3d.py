"""Initial tensors and ln Z/V of measurement/Ising_model_3d.py:15-137, on the GPU.

The bond weight M = exp(beta s s') (2 x 2) is split by its SVD as
u sqrt(s), sqrt(s) vh and the site weight exp(h s) is folded into the site
tensor (:15-41); the impurity carries one more factor s (:43-79).  Both are
real for real beta and h, so they are returned as float64 GPU tensors (the
reference casts them to complex128; the kernels here are real).  The 2 x 2
split and the 64-entry site tensors are built on the host; the ATRG form's
8 x 8 SVD runs on the GPU.  OUTPUT_DIR is not required (the reference reads
it at import, :12).
"""
import os

import numpy as np
import torch

from .. import linalg, ops
from .._io import device
from ..tensor_class.tensor_class import ATRG_Tensor as Tensor


def _factors(beta, h):
    s = np.array([1.0, -1.0])
    M = np.exp(beta * np.outer(s, s))
    u, sv, vh = np.linalg.svd(M)
    us = u * np.sqrt(sv)[None, :]
    svh = np.sqrt(sv)[:, None] * vh
    return s, np.exp(h * s), us, svh


def _site(weight, us, svh):
    # a,xa,aX,ya,aY,ta,aT->xXyYtT
    return np.einsum("a,xa,aX,ya,aY,ta,aT->xXyYtT", weight, svh, us, svh, us, svh, us)


def _atrg(T6, Dcut, impure, loc):
    """(5,1,3,4,0,2) transpose and SVD of the (TXY) x (txy) matrix (:32-41)."""
    Td = torch.as_tensor(T6).to(device())
    M = ops.permute(Td, (5, 1, 3, 4, 0, 2)).view(8, 8)
    u, s, vh = linalg.svd_small(M[None])
    return Tensor(u[0].reshape(2, 2, 2, 8), s[0], vh[0].reshape(8, 2, 2, 2), Dcut, 3, impure, loc)


def init_tensor(beta, h, Dcut, rgscheme):
    """:15-41: the pure tensor, (x,X,y,Y,t,T) for "hotrg", an ATRG_Tensor for "atrg"."""
    _, site, us, svh = _factors(beta, h)
    T = _site(site, us, svh)
    if rgscheme == "hotrg":
        return torch.as_tensor(T).to(device())
    if rgscheme == "atrg":
        return _atrg(T, Dcut, False, {})
    raise ValueError(f"unknown rgscheme {rgscheme!r}")


def init_tensor_imp(beta, h, Dcut, rgscheme):
    """:43-79: (pure, impurity) with the impurity carrying the spin s at its site."""
    s, site, us, svh = _factors(beta, h)
    T, Timp = _site(site, us, svh), _site(s * site, us, svh)
    if rgscheme == "hotrg":
        return torch.as_tensor(T).to(device()), torch.as_tensor(Timp).to(device())
    if rgscheme == "atrg":
        return _atrg(T, Dcut, False, {}), _atrg(Timp, Dcut, True, {"T": 0, "X": 0, "Y": 0})
    raise ValueError(f"unknown rgscheme {rgscheme!r}")


def ln_Z_over_V(beta, h, Dcut, XLOOPS, YLOOPS, TLOOPS, gilt_eps=0):
    """:101-137: RGSCHEME (environment, default "hotrg") selects HOTRG_3d_QR or
    ATRG_3d_new; ln Z/V = sum of the factors + ln Tr T / 2^(steps)."""
    rgscheme = os.environ.get("RGSCHEME", "hotrg")
    V = 2 ** (XLOOPS + YLOOPS + TLOOPS)
    if rgscheme == "hotrg":
        from ..trg import HOTRG_3d_QR as q
        T = q.pure_tensor({"Tensor": init_tensor(beta, h, Dcut, rgscheme), "loc": {}, "factor": {}}, Dcut,
                          {"X": XLOOPS, "Y": YLOOPS, "Z": TLOOPS})
        tr = float(ops.trace_pairs(ops.contiguous(T["Tensor"][None]))[0].item())
        return sum(T["factor"].values()) + np.log(tr) / V
    if rgscheme == "atrg":
        from ..trg import ATRG_3d_new as a
        T = a.pure_tensor(init_tensor(beta, h, Dcut, rgscheme), Dcut, {"X": XLOOPS, "Y": YLOOPS, "T": TLOOPS})
        return float(sum(T.normal_const.values())) + np.log(float(T.trace())) / V
    raise ValueError("Only support hotrg")


def _scheme():
    return os.environ.get("RGSCHEME", "hotrg")


def magnetization(beta, h, Dcut, XLOOPS, YLOOPS, TLOOPS, gilt_eps=0):
    """:140-204: (ln Z/V, |<s>|) from the one-point flow of the spin impurity
    at the origin (HOTRG_3d_QR or ATRG_3d_new by RGSCHEME)."""
    V = 2 ** (XLOOPS + YLOOPS + TLOOPS)
    rgscheme = _scheme()
    if rgscheme == "hotrg":
        from ..trg import HOTRG_3d_QR as q
        T, Ti = init_tensor_imp(beta, h, Dcut, rgscheme)
        steps = {"X": XLOOPS, "Y": YLOOPS, "Z": TLOOPS}
        T, T0 = q.one_point_function({"Tensor": T, "loc": {}, "factor": {}},
                                     {"Tensor": Ti, "loc": {"X": 0, "Y": 0, "Z": 0}, "factor": {}}, Dcut, steps)
        lnz, P = q.cal_field_expected_value(T, [T0], steps, False)
        return lnz, abs(P[0])
    if rgscheme == "atrg":
        from ..trg import ATRG_3d_new as a
        T, Ti = init_tensor_imp(beta, h, Dcut, rgscheme)
        steps = {"X": XLOOPS, "Y": YLOOPS, "T": TLOOPS}
        T, (T0,) = a.one_point_function(T, [Ti], Dcut, steps)
        lnz, P = a.cal_field_expected_value(T, [T0], steps)
        return lnz, abs(P[0])
    raise ValueError("Only support hotrg")


def internal_energy(beta, h, Dcut, XLOOPS, YLOOPS, TLOOPS, gilt_eps=0):
    """:207-268: (ln Z/V, <s_0 s_1>) for the nearest neighbours (0,0,0) and
    (0,0,1) along t, from the two-point flow.  The flow normalises each
    impurity by |Tr|, which vanishes for the bare spin at h = 0 (the
    reference divides by it as well): use a non-zero field."""
    V = 2 ** (XLOOPS + YLOOPS + TLOOPS)
    rgscheme = _scheme()
    if rgscheme == "hotrg":
        from ..trg import HOTRG_3d_QR as q
        T, Ti = init_tensor_imp(beta, h, Dcut, rgscheme)
        steps = {"X": XLOOPS, "Y": YLOOPS, "Z": TLOOPS}
        T, T0 = q.two_point_function({"Tensor": T, "loc": {}, "factor": {}},
                                     {"Tensor": Ti, "loc": {"X": 0, "Y": 0, "Z": 0}, "factor": {}},
                                     {"Tensor": Ti.clone(), "loc": {"X": 0, "Y": 0, "Z": 1}, "factor": {}}, Dcut,
                                     steps)
        lnz, G = q.cal_field_expected_value(T, [T0], steps, False)
        return lnz, G[0]
    if rgscheme == "atrg":
        from ..trg import ATRG_3d_new as a
        T, T0 = init_tensor_imp(beta, h, Dcut, rgscheme)
        _, Tn = init_tensor_imp(beta, h, Dcut, rgscheme)
        T0.loc, Tn.loc = {"X": 0, "Y": 0, "T": 0}, {"X": 0, "Y": 0, "T": 1}
        steps = {"X": XLOOPS, "Y": YLOOPS, "T": TLOOPS}
        T, T0 = a.two_point_function(T, T0, Tn, Dcut, steps)
        lnz, G = a.cal_field_expected_value(T, [T0], steps)
        return lnz, G[0]
    raise ValueError("Only support hotrg")


def write_result(outputname, beta, h, lnZ_OVER_V, observable=None):
    """One line of the 3D Ising driver's result file (Ising_model_3d.py:40-42,
    :51-55): "beta h lnZ/V" or "beta h lnZ/V m", each value of lnZ/V and m in
    the driver's {:19.12e} form (real parts), written fresh like the driver's
    open(outputname, "w")."""
    line = "{:} {:} {:19.12e}".format(beta, h, float(np.real(lnZ_OVER_V)))
    if observable is not None:
        line += " {:19.12e}".format(float(np.real(observable)))
    with open(outputname, "w") as out:
        out.write(line + "\n")
    return line
