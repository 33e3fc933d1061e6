"""Drop-in for tensor_class/tensor_class.py; traces and X ratios run on the GPU.

U, s, VH may be NumPy arrays or torch GPU tensors; results keep that kind.
"""
import numpy as np
import torch

from .. import atrg, ops
from .._io import from_device, to_device
from ..contract import contract


class HOTRG_Tensor:
    """tensor_class/tensor_class.py:5-29 with the intended constructor (the
    reference misspells it ``__int__``, :6)."""

    def __init__(self, T, Ttype="pure", loc=None):
        self.T = T
        self.Tensor = T
        self.dim = len(T.shape) // 2
        self.Ttype = Ttype
        self.loc = dict(loc or {})
        self.rgstep = {"X": 0, "Y": 0, "Z": 0}

    def trace(self):
        Td, kind = to_device(self.T)
        t = ops.trace_pairs(ops.contiguous(Td)[None])[0]
        return float(t.item()) if kind == "numpy" else t

    def cal_location(self, direction):
        self.loc[direction] //= 2
        self.rgstep[direction] += 1

    def update(self, T):
        self.T = T
        self.Tensor = T
        return self


class Tensor:
    """tensor_class/tensor_class.py:31-39: plain (U, s, VH) holder."""

    def __init__(self, U, s, VH):
        self.U = U
        self.s = s
        self.VH = VH


class ATRG_Tensor:
    """tensor_class/tensor_class.py:41-155 (dim 3 on the path)."""

    def __init__(self, U, s, VH, Dcut, dim, is_impure, loc):
        self.U = U
        self.s = s
        self.VH = VH
        self.Dcut = Dcut
        self.dim = dim
        self.loc = loc
        self.is_impure = is_impure
        self.normal_const = {}

    def _dev(self):
        (U, kind), (s, _), (VH, _) = to_device(self.U), to_device(self.s), to_device(self.VH)
        return U, s, VH, kind

    def trace(self):
        """:66-86 -- Tr T = sum U_{txyi} s_i VH_{itxy} (float64 on the device)."""
        if self.dim != 3 or np.ndim(self.s) != 1:
            raise NotImplementedError("ATRG_Tensor.trace on the GPU covers dim 3 with a diagonal s")
        U, s, VH, kind = self._dev()
        t = atrg.trace(U, s, VH)[0]
        return float(t.item()) if kind == "numpy" else t

    def cal_X(self):
        """:88-141 (dim 3): X_t, X_x, X_y = (Tr T)^2 / Tr(T T) along each axis."""
        U, s, VH, kind = self._dev()
        U, VH = atrg._f64(U), atrg._f64(VH)
        s = atrg._f64(s)
        TrT = float(atrg.trace(U, s, VH)[0].item())
        sVH = ops.contiguous(VH)
        ops.scale_rows_(sVH, s, s.numel(), VH.numel() // VH.shape[0], 0)
        out = []
        for (l1, l2, r1, r2) in (("Txyi", "itxy", "txyi", "iTxy"),      # time
                                 ("tXyi", "itxy", "txyi", "itXy"),      # x
                                 ("txYi", "itxy", "txyi", "itxY")):     # y
            a = contract(f"{l1},{l2}->{_free(l1, l2)}", U, sVH)
            b = contract(f"{r1},{r2}->{_free(r1, r2)}", U, sVH)
            tt = _pair_trace(a, b)
            out.append(TrT ** 2 / tt)
        return tuple(out)

    def update(self, U, s, VH):
        self.U, self.s, self.VH = U, s, VH
        return self

    def append_normalization_const(self, key, value):
        self.normal_const[key] = value

    def get_normalization_const(self):
        return np.asarray(list(self.normal_const.values()))


def _free(a, b):
    return "".join(c for c in a + b if (a + b).count(c) == 1)


def _pair_trace(a, b):
    """sum_{uv} a[u, v] b[v, u] for the 2-leg intermediates of cal_X."""
    ab = ops.gemm(a, b)
    return float(ops.trace_general(ab, [ab.shape[0]], [ab.shape[0] + 1])[0].item())
