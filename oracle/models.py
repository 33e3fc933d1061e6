"""Initial tensors of the three workload models (test infrastructure only).

Follow measurement/Ising_model.py:14-24 (2D Ising), measurement/Ising_model_3d.py:15-41
(3D Ising, HOTRG and ATRG layouts) and tensor_init/Potts_model.py:38-96 (q-state
Potts) with h = 0: every bond weight Q = W W^T is split symmetrically and
T_{legs} = sum_s prod_leg W_{s,leg}.
"""
import numpy as np

from .tensor_class import ATRG_Tensor


def ising_W(beta):
    """W W^T = [[e^b, e^-b], [e^-b, e^b]] with W = [[sqrt ch, sqrt sh], [sqrt ch, -sqrt sh]]."""
    ch, sh = np.sqrt(np.cosh(beta)), np.sqrt(np.sinh(beta))
    return np.array([[ch, sh], [ch, -sh]], dtype=np.float64)


def potts_W(beta, q=3):
    """exp(beta delta_{ss'}) = V diag(lam) V^T, W = V diag(sqrt(lam))."""
    M = np.exp(beta * np.eye(q))
    lam, V = np.linalg.eigh(M)
    lam, V = lam[::-1], V[:, ::-1]
    return V * np.sqrt(np.clip(lam, 0.0, None))[None, :]


def site_tensor(W, nlegs):
    """T_{l1..ln} = sum_s prod_k W_{s,lk}."""
    q, D = W.shape
    T = np.zeros((D,) * nlegs)
    for s in range(q):
        t = W[s]
        for _ in range(nlegs - 1):
            t = np.multiply.outer(t, W[s])
        T += t
    return T


def ising2d_tensor(beta):
    return site_tensor(ising_W(beta), 4)


def ising3d_tensor(beta):
    return site_tensor(ising_W(beta), 6)


def potts3d_tensor(beta, q=3):
    return site_tensor(potts_W(beta, q), 6)


def to_atrg(T6, Dcut):
    """(x,X,y,Y,t,T) -> transpose (5,1,3,4,0,2) -> SVD of (TXY)x(txy)
    (measurement/Ising_model_3d.py:32-41, tensor_init/Potts_model.py:79-91)."""
    D = T6.shape[0]
    M = np.transpose(T6, (5, 1, 3, 4, 0, 2)).reshape(D ** 3, D ** 3)
    u, s, vh = np.linalg.svd(M)
    return ATRG_Tensor(u.reshape(D, D, D, len(s)), s, vh.reshape(len(s), D, D, D), Dcut, 3, False, {})


def zero_pad(T, D):
    """Zero-pad every leg to D (what the reference's callers do so that the
    bond equals Dcut, measurement/Potts_3d_inf_density_measurement.py:63-64)."""
    out = np.zeros((D,) * T.ndim, dtype=T.dtype)
    out[tuple(slice(0, n) for n in T.shape)] = T
    return out


def ising3d_field_tensors(beta, h):
    """measurement/Ising_model_3d.py:15-79 (real form): M = exp(beta s s') split
    by its SVD as u sqrt(s), sqrt(s) vh; site weight exp(h s); the impurity
    carries one more factor s.  Returns (T, Timp), legs (x,X,y,Y,t,T)."""
    s = np.array([1.0, -1.0])
    u, sv, vh = np.linalg.svd(np.exp(beta * np.outer(s, s)))
    us, svh = u * np.sqrt(sv)[None, :], np.sqrt(sv)[:, None] * vh
    site = np.exp(h * s)
    T = np.einsum("a,xa,aX,ya,aY,ta,aT->xXyYtT", site, svh, us, svh, us, svh, us)
    Timp = np.einsum("a,xa,aX,ya,aY,ta,aT->xXyYtT", s * site, svh, us, svh, us, svh, us)
    return T, Timp
