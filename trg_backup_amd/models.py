"""Initial tensors of the workload models (the step before the coarse-graining path).

Follow measurement/Ising_model.py:14-24, measurement/Ising_model_3d.py:15-41 and
tensor_init/Potts_model.py:38-96 at h = 0: the bond weight Q = W W^T is split
symmetrically and T_{legs} = sum_s prod_leg W_{s,leg}.  The weights are 2x2 /
3x3 closed forms; the site tensors (64 and 729 entries) are built on the host
and copied to the GPU, where everything else runs.  The ATRG form's SVD of the
(q^3 x q^3) matrix runs on the GPU (libtrg one-sided Jacobi).
"""
import math

import numpy as np
import torch

from . import linalg, ops
from ._io import device

BETA_C = {"ising2d": 0.5 * math.log(1.0 + math.sqrt(2.0)), "ising3d": 0.2216546, "potts3d": 0.550565}


def ising_W(beta):
    ch, sh = math.sqrt(math.cosh(beta)), math.sqrt(math.sinh(beta))
    return np.array([[ch, sh], [ch, -sh]])


def potts_W(beta, q=3):
    """exp(beta delta) = (e^b - 1) I + J has eigenvalues e^b + q - 1 (uniform
    vector) and e^b - 1 (q - 1 fold); W = V sqrt(lambda)."""
    lam = [math.exp(beta) + q - 1] + [math.exp(beta) - 1] * (q - 1)
    V = np.zeros((q, q))
    V[:, 0] = 1.0 / math.sqrt(q)
    for j in range(1, q):            # orthonormal complement (Helmert basis)
        V[:j, j] = 1.0 / math.sqrt(j * (j + 1))
        V[j, j] = -j / math.sqrt(j * (j + 1))
    return V * np.sqrt(np.array(lam))[None, :]


def site_tensor(W, nlegs):
    q, D = W.shape
    T = np.zeros((D,) * nlegs)
    for s in range(q):
        t = W[s]
        for _ in range(nlegs - 1):
            t = np.multiply.outer(t, W[s])
        T += t
    return T


def initial_tensor(model, beta, dev=None):
    """Device float64 tensor: 2D (x,x',y,y') or 3D (x,x',y,y',z,z')."""
    dev = dev or device()
    if model == "ising2d":
        T = site_tensor(ising_W(beta), 4)
    elif model == "ising3d":
        T = site_tensor(ising_W(beta), 6)
    elif model == "potts3d":
        T = site_tensor(potts_W(beta), 6)
    else:
        raise ValueError(f"unknown model {model!r}")
    return torch.as_tensor(T).to(dev)


def atrg_initial(model, beta, dev=None):
    """(U (T,X,Y,i), s (i,), VH (i,t,x,y)) on the device: T6 transposed
    (5,1,3,4,0,2) and split by an SVD (measurement/Ising_model_3d.py:32-41)."""
    T = initial_tensor(model, beta, dev)
    D = T.shape[0]
    M = ops.permute(T, (5, 1, 3, 4, 0, 2)).view(D ** 3, D ** 3)
    u, s, vh = linalg.svd_small(M[None])
    return u[0].reshape(D, D, D, D ** 3), s[0], vh[0].reshape(D ** 3, D, D, D)
