"""NumPy restatement of tensor_class/tensor_class.py (test infrastructure only)."""
import numpy as np


class HOTRG_Tensor:
    """tensor_class/tensor_class.py:5-29.  The reference misspells the
    constructor (``__int__``, :6) so it cannot be built with arguments; this
    restatement provides the intended ``__init__``."""

    def __init__(self, T, Ttype="pure", loc=None):
        self.T = T
        self.Tensor = T
        self.dim = T.ndim // 2
        self.Ttype = Ttype
        self.loc = dict(loc or {})
        self.rgstep = {"X": 0, "Y": 0, "Z": 0}

    def trace(self):
        if self.dim == 2:
            return np.einsum("xxyy", self.T)
        if self.dim == 3:
            return np.einsum("xxyyzz", self.T)
        return np.einsum("xxyyzztt", self.T)

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
    """tensor_class/tensor_class.py:41-155 for dim 3 (the only dim on the path)."""

    def __init__(self, U, s, VH, Dcut, dim, is_impure, loc):
        self.U = U
        self.s = s
        self.VH = VH
        self.Dcut = Dcut
        self.dim = dim
        self.loc = loc
        self.is_impure = is_impure
        self.normal_const = {}

    def trace(self):
        """:66-86 -- Tr T = sum_{txy,i} U_{txyi} s_i VH_{itxy}."""
        if self.dim == 2:
            if self.s.ndim == 1:
                return np.einsum("txi,i,itx", self.U, self.s, self.VH)
            return np.einsum("txi,ij,jtx", self.U, self.s, self.VH)
        if self.dim == 3:
            if self.s.ndim == 1:
                return np.einsum("txyi,i,itxy", self.U, self.s, self.VH)
            return np.einsum("txyi,ij,jtxy", self.U, self.s, self.VH)
        if self.s.ndim == 1:
            return np.einsum("txyzi,i,itxyz", self.U, self.s, self.VH)
        return np.einsum("txyzi,ij,jtxyz", self.U, self.s, self.VH)

    def cal_X(self):
        """:88-141 (dim 3) -- X_t, X_x, X_y = (Tr T)^2 / Tr(T T) along each axis."""
        TrT = self.trace()
        U, s, VH = self.U, self.s, self.VH
        if self.dim == 2:
            TTx = np.einsum("txi,i,itX,TXj,j,jTx", U, s, VH, U, s, VH)
            TTt = np.einsum("txi,i,iTx,TXj,j,jtX", U, s, VH, U, s, VH)
            return TrT ** 2 / TTt, TrT ** 2 / TTx
        X1 = np.einsum("tXyi,i,itxy->Xx", U, s, VH)
        X2 = np.einsum("txyi,i,itXy->xX", U, s, VH)
        TTx = np.einsum("Xx,xX", X1, X2)
        Y1 = np.einsum("txYi,i,itxy->Yy", U, s, VH)
        Y2 = np.einsum("txyi,i,itxY->yY", U, s, VH)
        TTy = np.einsum("Yy,yY", Y1, Y2)
        T1 = np.einsum("Txyi,i,itxy->Tt", U, s, VH)
        T2 = np.einsum("txyi,i,iTxy->tT", U, s, VH)
        TTt = np.einsum("Tt,tT", T1, T2)
        return TrT ** 2 / TTt, TrT ** 2 / TTx, TrT ** 2 / TTy

    def update(self, U, s, VH):
        self.U, self.s, self.VH = U, s, VH
        return self

    def append_normalization_const(self, key, value):
        self.normal_const[key] = value

    def get_normalization_const(self):
        return np.asarray(list(self.normal_const.values()))
