"""Drop-in for tensor_init/Potts_model.py's ``infinite_density`` (:10-96) at h = 0 and mu = 0.

M = exp(k delta_{ss'}) (q x q) is split by its SVD as u sqrt(s), sqrt(s) vh
and the site tensor is sum_a prod_leg of the factors (:38-73); the ATRG form
transposes (5,1,3,4,0,2) and splits the q^3 x q^3 matrix by an SVD on the
GPU (:75-94).  A field h or chemical potential mu makes the tensors complex
(:40-42); the kernels here are real, so those raise NotImplementedError.
"""
import numpy as np
import torch

from .. import linalg, ops
from .._io import device
from ..tensor_class.tensor_class import ATRG_Tensor as Tensor


class infinite_density:
    def __init__(self, dim, Dcut, q, k, h, mu):
        self.q, self.k, self.h, self.mu, self.Dcut, self.dim = q, k, h, mu, Dcut, dim

    def potts_spin(self):
        """:19-22: the q-th roots of unity (complex, host)."""
        return np.exp(2j * np.pi * np.arange(self.q) / self.q)

    def tensor_component_3d(self):
        """:38-56: (u sqrt(s), sqrt(s) vh, Z) of M = exp(k delta) (h = 0)."""
        if self.h != 0 or self.mu != 0:
            raise NotImplementedError("a field or chemical potential makes the Potts tensor complex; "
                                      "the device kernels are real")
        M = np.exp(self.k * np.eye(self.q))
        u, s, vh = np.linalg.svd(M)
        return u * np.sqrt(s)[None, :], np.sqrt(s)[:, None] * vh, self.potts_spin()

    def cal_HOTRG_init_tensor(self, impure=False):
        """:58-73 for dim 2 or 3 (float64 GPU tensor)."""
        us, svh, Z = self.tensor_component_3d()
        if self.dim == 2:
            T = np.einsum("xa,aX,ya,aY->xXyY", svh, us, svh, us)
        elif self.dim == 3:
            T = np.einsum("xa,aX,ya,aY,ta,aT->xXyYtT", svh, us, svh, us, svh, us)
        else:
            raise NotImplementedError("dim 4 is not on the device path")
        T = torch.as_tensor(T).to(device())
        return (T, us, svh, Z) if impure else T

    def cal_ATRG_init_tensor(self, impure=False):
        """:75-94: ATRG_Tensor of the 3D tensor."""
        us, svh, Z = self.tensor_component_3d()
        q = self.q
        T = torch.as_tensor(np.einsum("xa,aX,ya,aY,ta,aT->xXyYtT", svh, us, svh, us, svh, us)).to(device())
        Mq = ops.permute(T, (5, 1, 3, 4, 0, 2)).view(q ** 3, q ** 3)
        U, S, VH = linalg.svd_small(Mq[None])
        T = Tensor(U[0].reshape(q, q, q, q ** 3), S[0], VH[0].reshape(q ** 3, q, q, q), self.Dcut, self.dim, False,
                   {})
        return (T, us, svh, Z) if impure else T
