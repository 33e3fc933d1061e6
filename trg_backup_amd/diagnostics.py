"""Device diagnostics of a six-leg tensor T[x, X, y, Y, t, T] (HOTRG 3D).

The partial traces are trg_trace launches over strided index grids (no
contraction with an identity tensor), the ratios and spectra come from libtrg
GEMMs and Jacobi decompositions, and only the final handful of scalars is
read back.  Used by the HOTRG_3d / HOTRG_3d_QR drop-ins
(trg/HOTRG_3d_QR.py:559-713, trg/HOTRG_3d.py:240-245).
"""
import torch

from . import linalg, ops


def _f64c(T):
    return ops.contiguous(T, torch.float64)


def _strides(T):
    dx, dX, dy, dY, dt, dT = T.shape
    sT, st = 1, dT
    sY = dt * dT
    sy = dY * sY
    sX = dy * sy
    sx = dX * sX
    return (dx, dX, dy, dY, dt, dT), (sx, sX, sy, sY, st, sT)


def partial_trace(T, leg):
    """D[a, b] for leg "t" = sum_{x,y} T[x,x,y,y,a,b], "x" = sum_{y,t} T[a,b,y,y,t,t],
    "y" = sum_{x,t} T[x,x,a,b,t,t]; float64 (n, n) on the device."""
    T = _f64c(T)
    (dx, dX, dy, dY, dt, dT), (sx, sX, sy, sY, st, sT) = _strides(T)
    if leg == "t":
        n, m, dims, strides, sb = dt, dT, [dx, dy], [sx + sX, sy + sY], sT
    elif leg == "x":
        n, m, dims, strides, sb = dx, dX, [dy, dt], [sy + sY, st + sT], sX
    else:
        n, m, dims, strides, sb = dy, dY, [dx, dt], [sx + sX, st + sT], sY
    return ops.trace_general(T, dims, strides, batch=n * m, sb=sb).view(n, m)


def density_matrix(T, leg):
    """The reference's density matrices: "xxyytT->Tt", "xXyytt->Xx",
    "xxyYtt->Yy" (trg/HOTRG_3d_QR.py:617, :645, :673, :695), i.e. the
    transposed partial trace, materialised."""
    return ops.contiguous(partial_trace(T, leg).t())


def diagonal(T):
    """T[x,x,y,y,t,t] for all (x, y, t), float64, flattened in (x, y, t) order."""
    T = _f64c(T)
    (dx, _, dy, _, dt, _), (sx, sX, sy, sY, st, sT) = _strides(T)
    view = torch.as_strided(T, (dx, dy, dt), (sx + sX, sy + sY, st + sT))
    return ops.contiguous(view).view(-1)


def _trace_of_product(A, B):
    """sum_{uv} A[u, v] B[v, u]."""
    P = ops.gemm(ops.contiguous(A), ops.contiguous(B))
    return float(ops.trace_general(P, [P.shape[0]], [P.shape[0] + 1])[0].item())


def x_ratios(T):
    """(Tr T)^2 / TT along t, x, y: TT is the trace of two tensors stacked
    along that leg (trg/HOTRG_3d_QR.py:560-567)."""
    T = _f64c(T)
    tr = float(ops.trace_pairs(T)[0].item())
    out = []
    for leg in "txy":
        D = partial_trace(T, leg)
        out.append(tr * tr / _trace_of_product(D, D))
    return tuple(out)


def lower_symmetric(A):
    """The symmetric matrix of A's lower triangle (what cupy.linalg.eigh reads,
    UPLO="L"), built with device copies."""
    n = A.shape[0]
    S = ops.contiguous(A)
    for i in range(n - 1):
        ops.copy_into(S[i, i + 1:], A[i + 1:, i])
    return S


def eigvals_lower(A):
    """Eigenvalues (ascending, device Jacobi) of the lower triangle of A."""
    w, _ = linalg.eigh_small(lower_symmetric(A)[None])
    return w[0]


def singular_values(A):
    """Singular values (descending, device Jacobi) of a small square matrix."""
    _, s, _ = linalg.svd_thin(_f64c(A)[None])
    return s[0]


def tensor_spectrum(T, ncols, q=4):
    """Leading min(ncols, n) singular values of T as the (x y z) x (x' y' z')
    matrix (trg/HOTRG_3d.py:240-245): exact thin Jacobi SVD when n <= 128,
    else a randomized range finder of 128 columns with q power iterations.
    Returns (s[:ncols] / max s, max s) on the host."""
    dx, dX, dy, dY, dt, dT = T.shape
    M = ops.permute(_f64c(T), (0, 2, 4, 1, 3, 5)).view(1, dx * dy * dt, dX * dY * dT)
    n = min(M.shape[1], M.shape[2])
    if n <= linalg.SMALL_MAX:
        _, s, _ = linalg.svd_thin(M)
    else:
        _, s, _ = linalg.svd_topk(M, linalg.SMALL_MAX, q)
    s = s[0, :ncols].cpu().numpy()
    return s / s[0], float(s[0])
