"""Host <-> device plumbing for the drop-in modules.

NumPy inputs are copied to the GPU, computed there, and copied back;
torch GPU tensors stay on the device.  Complex inputs are accepted when
their imaginary part is zero (the reference's callers cast real models to
complex128, e.g. measurement/Ising_model.py:23); the kernels are real.
"""
import numpy as np
import torch

from ._lib import TrgError


def device():
    if not torch.cuda.is_available():
        raise TrgError("no GPU visible: the trg_backup_amd modules run only on the GPU")
    return torch.device("cuda", torch.cuda.current_device())


def to_device(T, dtype=torch.float64):
    """Returns (device tensor, kind) with kind 'numpy' or 'torch'."""
    if isinstance(T, torch.Tensor):
        if T.is_complex():
            if torch.count_nonzero(T.imag).item() != 0:
                raise TrgError("complex tensors with a non-zero imaginary part are not supported")
            T = T.real
        if not T.is_cuda:
            T = T.to(device())
        if T.dtype not in (torch.float32, torch.float64):
            T = T.to(dtype)
        return T, "torch"
    A = np.asarray(T)
    if np.iscomplexobj(A):
        if np.any(A.imag != 0):
            raise TrgError("complex tensors with a non-zero imaginary part are not supported")
        A = A.real
    A = np.ascontiguousarray(A, dtype=np.float64)
    return torch.from_numpy(A).to(device()), "numpy"


def from_device(T, kind):
    if kind == "numpy":
        return T.detach().cpu().numpy()
    return T


def scalar(x):
    if isinstance(x, torch.Tensor):
        return float(x.item())
    return float(x)
