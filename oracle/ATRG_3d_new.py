"""NumPy restatement of trg/ATRG_3d_new.py's pure-tensor path (test infrastructure only).

Tensor layout: T_{TXYtxy} = U_{TXYi} s_i VH_{itxy} (an ``ATRG_Tensor``).
One step (trg/ATRG_3d_new.py:441-482): leg transposition, B = s.VH and
C = U.s, leg swapping by randomized SVD, squeezer, coarse graining, restore.

Differences from the reference, all documented:
  * the rSVD sketch comes from the Philox stream (oracle/philox.py) rather
    than cuRAND ``RandomState(12345)`` (:90-91);
  * the rSVD parameters default to the reference's (k = 3 chi_k, p = 0,
    q = Dcut, :193) and can be switched to the workload's (k = chi_k,
    p = chi, q = 2) with ``set_rsvd_params``;
  * ``DGENERACY_EPS`` defaults to 0.0 (the reference requires the env var,
    :15); ``OUTPUT_DIR`` file output (:197-212, :736-766) is opt-in.
"""
import itertools
import math
import os

import numpy as np

from .philox import standard_normal
from .tensor_class import ATRG_Tensor as Tensor
from .truncated_svd import svd, truncate

degeneracy_eps = float(os.environ.get("DGENERACY_EPS", "0.0"))
truncate_eps = 1e-10

# rSVD parameters of the leg swap: rank multiplier, oversampling, QR iterations
# (None = Dcut as in the reference), seed.
_RSVD = {"k_factor": 3, "p": 0, "q": None, "seed": 12345}


# Basis of the leg swap: "svd" follows the reference (SVDs of B and C,
# trg/ATRG_3d_new.py:175-188); "identity" keeps the (x1,y1) / (X0,Y0) legs
# as they are.  Both give the same M = B~ C~ up to a rotation of its row and
# column legs; "identity" draws the rSVD sketch in the same basis as the GPU.
LEG_SWAP_BASIS = {"mode": "svd"}


def set_leg_swap_basis(mode):
    LEG_SWAP_BASIS["mode"] = mode


def set_rsvd_params(k_factor=3, p=0, q=None, seed=12345):
    """Select the leg-swap rSVD parameters (reference: 3, 0, Dcut, 12345)."""
    _RSVD.update(k_factor=k_factor, p=p, q=q, seed=seed)


def get_rsvd_params():
    return dict(_RSVD)


def _conj(A):
    return np.conj(A) if np.iscomplexobj(A) else A


def leg_transposition(A, do_what="transpose", direction="t"):
    """trg/ATRG_3d_new.py:18-65."""
    if direction in ("t", "T"):
        return A
    fwd = do_what == "transpose"
    if direction in ("x", "X"):
        ua, va = ((1, 2, 0, 3), (0, 2, 3, 1)) if fwd else ((2, 0, 1, 3), (0, 3, 1, 2))
    else:
        ua, va = ((2, 0, 1, 3), (0, 3, 1, 2)) if fwd else ((1, 2, 0, 3), (0, 2, 3, 1))
    A.U = np.ascontiguousarray(np.transpose(A.U, ua))
    A.VH = np.ascontiguousarray(np.transpose(A.VH, va))
    return A


def _qr(Y):
    q, _ = np.linalg.qr(Y)
    return q


def rsvd_for_ATRG_leg_swapping(B, C, chi_k, p=32, q=64, eps=1e-10, seed=12345):
    """trg/ATRG_3d_new.py:67-140: matrix-free rSVD of
    M_{(i,g),(j,b)} = sum_t B_{b,i,t} C_{t,j,g}.  Returns UM_{i,g,k}, sM_k, VMH_{k,j,b}.
    """
    nb, ni, nt = B.shape
    _, nj, ng = C.shape
    l = chi_k + p
    Y = standard_normal((nj, nb, l), seed)
    # Y'_{i g l} = sum B_{b i t} C_{t j g} Y_{j b l}
    W = np.tensordot(C, Y, axes=([1], [0]))                 # (t,g,b,l)
    Y = np.tensordot(B, W, axes=([0, 2], [2, 0]))           # (i,g,l)
    Q = _qr(Y.reshape(ni * ng, l))
    Bc, Cc = _conj(B), _conj(C)
    for _ in range(q):
        ls = Q.shape[1]
        Qt = Q.reshape(ni, ng, ls)
        # (M^dag Q)_{j b l} = sum B*_{b i t} C*_{t j g} Q_{i g l}
        W = np.tensordot(Bc, Qt, axes=([1], [0]))           # (b,t,g,l)
        Y = np.tensordot(Cc, W, axes=([0, 2], [1, 2]))      # (j,b,l)
        Q = _qr(Y.reshape(nj * nb, ls))
        ls = Q.shape[1]
        Qt = Q.reshape(nj, nb, ls)
        W = np.tensordot(C, Qt, axes=([1], [0]))            # (t,g,b,l)
        Y = np.tensordot(B, W, axes=([0, 2], [2, 0]))       # (i,g,l)
        Q = _qr(Y.reshape(ni * ng, ls))
    ls = Q.shape[1]
    Qt = Q.reshape(ni, ng, ls)
    # X_{l j b} = sum Q*_{i g l} B_{b i t} C_{t j g}
    W = np.tensordot(_conj(Qt), B, axes=([0], [1]))         # (g,l,b,t)
    X = np.tensordot(W, C, axes=([0, 3], [2, 0]))           # (l,b,j)
    X = np.transpose(X, (0, 2, 1)).reshape(ls, nj * nb)
    UX, sM, VMH = np.linalg.svd(X, full_matrices=False)
    r = len(sM)
    VMH = VMH.reshape(r, nj, nb)
    UM = (Q @ UX).reshape(ni, ng, r)
    s0 = sM / np.max(sM)
    k = min(int(np.count_nonzero(s0 > eps)), chi_k)
    return UM[:, :, :k], sM[:k], VMH[:k]


def leg_swapping(B, C, Dcut, ispure=True):
    """trg/ATRG_3d_new.py:142-235.  B_{i,t1,x1,y1}, C_{T0,X0,Y0,j} ->
    UM_{i,X0,Y0,k}, sM_k, VMH_{k,j,x1,y1}."""
    ni, nt1, nx1, ny1 = B.shape
    nT0, nX0, nY0, nj = C.shape
    if LEG_SWAP_BASIS["mode"] == "identity":
        return _leg_swapping_identity(B, C, Dcut)
    Bm = np.transpose(B, (2, 3, 0, 1)).reshape(nx1 * ny1, ni * nt1)
    UB, sB, VBH = np.linalg.svd(Bm, full_matrices=False)
    beta = len(sB)
    UB = UB.reshape(nx1, ny1, beta)
    Bt = sB[:, None, None] * VBH.reshape(beta, ni, nt1)
    Cm = np.transpose(C, (0, 3, 1, 2)).reshape(nT0 * nj, nX0 * nY0)
    UC, sC, VCH = np.linalg.svd(Cm, full_matrices=False)
    gamma = len(sC)
    VCH = VCH.reshape(gamma, nX0, nY0)
    Ct = UC.reshape(nT0, nj, gamma) * sC[None, None, :]
    chi_k = min(Dcut, nj * beta, ni * gamma)
    q = Dcut if _RSVD["q"] is None else _RSVD["q"]
    UM, sM, VMH = rsvd_for_ATRG_leg_swapping(Bt, Ct, chi_k=_RSVD["k_factor"] * chi_k,
                                             p=_RSVD["p"], q=q, eps=truncate_eps, seed=_RSVD["seed"])
    chi, _ = truncate(sM, k=chi_k, degeneracy_eps=degeneracy_eps, truncate_eps=truncate_eps)
    UM, sM, VMH = UM[:, :, :chi], sM[:chi], VMH[:chi]
    UM = np.tensordot(UM, VCH, axes=([1], [0]))             # (i,k,X,Y)
    UM = np.transpose(UM, (0, 2, 3, 1))
    VMH = np.tensordot(VMH, UB, axes=([2], [2]))            # (k,j,x,y)
    return UM, sM, VMH


def _leg_swapping_identity(B, C, Dcut):
    """leg_swapping with the (x1,y1) and (X0,Y0) legs kept as the beta and
    gamma bases; chi_k is still bounded by the reference's reduced sizes."""
    ni, nt1, nx1, ny1 = B.shape
    nT0, nX0, nY0, nj = C.shape
    beta_ref = min(nx1 * ny1, ni * nt1)
    gamma_ref = min(nT0 * nj, nX0 * nY0)
    chi_k = min(Dcut, nj * beta_ref, ni * gamma_ref)
    Bt = np.ascontiguousarray(np.transpose(B, (2, 3, 0, 1))).reshape(nx1 * ny1, ni, nt1)
    Ct = np.ascontiguousarray(np.transpose(C, (0, 3, 1, 2))).reshape(nT0, nj, nX0 * nY0)
    q = Dcut if _RSVD["q"] is None else _RSVD["q"]
    UM, sM, VMH = rsvd_for_ATRG_leg_swapping(Bt, Ct, chi_k=_RSVD["k_factor"] * chi_k,
                                             p=_RSVD["p"], q=q, eps=truncate_eps, seed=_RSVD["seed"])
    chi, _ = truncate(sM, k=chi_k, degeneracy_eps=degeneracy_eps, truncate_eps=truncate_eps)
    UM, sM, VMH = UM[:, :, :chi], sM[:chi], VMH[:chi]
    return UM.reshape(ni, nX0, nY0, chi), sM, VMH.reshape(chi, nj, nx1, ny1)


def squeezer_environments(A, UM, sM, VMH, D):
    """The four environments of trg/ATRG_3d_new.py:262-265 as matrices."""
    nX1, nY1 = A.shape[1], A.shape[2]
    nX0, nY0 = UM.shape[1], UM.shape[2]
    nx1, ny1 = VMH.shape[2], VMH.shape[3]
    nx0, ny0 = D.shape[2], D.shape[3]
    Ac, UMc, VMHc, Dc = _conj(A), _conj(UM), _conj(VMH), _conj(D)
    UMs = UM * sM[None, None, None, :]
    # XdagX[i,j,k,l] = sum A*[a,i,b,c] A[a,k,b,f] UM*[c,j,d,e] s[e] UM[f,l,d,e]
    P = np.tensordot(Ac, A, axes=([0, 2], [0, 2]))          # (i,c,k,f)
    Q = np.tensordot(UMc, UMs, axes=([2, 3], [2, 3]))       # (c,j,f,l)
    XdX = np.tensordot(P, Q, axes=([1, 3], [0, 2]))          # (i,k,j,l)
    XdX = np.transpose(XdX, (0, 2, 1, 3)).reshape(nX1 * nX0, nX1 * nX0)
    # YdagY[i,j,k,l] = sum A*[a,b,i,c] A[a,b,k,f] UM*[c,d,j,e] s[e] UM[f,d,l,e]
    P = np.tensordot(Ac, A, axes=([0, 1], [0, 1]))          # (i,c,k,f)
    Q = np.tensordot(UMc, UMs, axes=([1, 3], [1, 3]))       # (c,j,f,l)
    YdY = np.tensordot(P, Q, axes=([1, 3], [0, 2]))
    YdY = np.transpose(YdY, (0, 2, 1, 3)).reshape(nY1 * nY0, nY1 * nY0)
    VMHs = VMH * sM[:, None, None, None]
    # xxdag[i,j,k,l] = sum D[b,e,j,f] D*[d,e,l,f] VMH[a,b,i,c] s[a] VMH*[a,d,k,c]
    P = np.tensordot(D, Dc, axes=([1, 3], [1, 3]))          # (b,j,d,l)
    Q = np.tensordot(VMHs, VMHc, axes=([0, 3], [0, 3]))     # (b,i,d,k)
    xxd = np.tensordot(Q, P, axes=([0, 2], [0, 2]))          # (i,k,j,l)
    xxd = np.transpose(xxd, (0, 2, 1, 3)).reshape(nx1 * nx0, nx1 * nx0)
    # yydag[i,j,k,l] = sum D[b,e,f,j] D*[d,e,f,l] VMH[a,b,c,i] s[a] VMH*[a,d,c,k]
    P = np.tensordot(D, Dc, axes=([1, 2], [1, 2]))          # (b,j,d,l)
    Q = np.tensordot(VMHs, VMHc, axes=([0, 2], [0, 2]))     # (b,i,d,k)
    yyd = np.tensordot(Q, P, axes=([0, 2], [0, 2]))
    yyd = np.transpose(yyd, (0, 2, 1, 3)).reshape(ny1 * ny0, ny1 * ny0)
    return XdX, YdY, xxd, yyd


def _oblique_pair(Rbig, Rsmall, Dcut, shape_big, shape_small):
    """:278-292 -- u,s,vh = svd(RX@Rx, k) ; PX = Rx v s^-1/2 ; Px = s^-1/2 u^dag RX."""
    k1 = min(Rbig.shape[0], Rsmall.shape[1], Dcut)
    u, s, vh = svd(Rbig @ Rsmall, shape=[[0], [1]], k=k1,
                   degeneracy_eps=degeneracy_eps, truncate_eps=truncate_eps)
    uh = np.conj(u.T)
    v = np.conj(vh.T)
    sis = np.sqrt(1.0 / s)
    PX = (Rsmall @ v) * sis[None, :]
    Px = sis[:, None] * (uh @ Rbig)
    chi = min(len(s), Dcut)
    PX = PX[:, :chi].reshape(shape_big + (chi,))
    Px = Px[:chi, :].reshape((chi,) + shape_small)
    return Px, PX


def squeezer(A, UM, sM, VMH, D, Dcut):
    """trg/ATRG_3d_new.py:246-317 -> Px_{x,x1,x0}, PX_{X1,X0,X}, Py, PY."""
    XdX, YdY, xxd, yyd = squeezer_environments(A, UM, sM, VMH, D)
    _, sX, vXh = np.linalg.svd(XdX, full_matrices=False)
    _, sY, vYh = np.linalg.svd(YdY, full_matrices=False)
    ux, sx, _ = np.linalg.svd(xxd, full_matrices=False)
    uy, sy, _ = np.linalg.svd(yyd, full_matrices=False)
    RX = np.sqrt(sX)[:, None] * vXh
    Rx = ux * np.sqrt(sx)[None, :]
    Px, PX = _oblique_pair(RX, Rx, Dcut, (A.shape[1], UM.shape[1]), (VMH.shape[2], D.shape[2]))
    RY = np.sqrt(sY)[:, None] * vYh
    Ry = uy * np.sqrt(sy)[None, :]
    Py, PY = _oblique_pair(RY, Ry, Dcut, (A.shape[2], UM.shape[2]), (VMH.shape[3], D.shape[3]))
    return Px, PX, Py, PY


def leg_slicing(leg_size, slicing):
    """trg/ATRG_3d_new.py:319-355 (slice list of a leg).  A zero-size leg
    raises ValueError instead of the reference's sys.exit(0) (:330-334)."""
    leg_size, slicing = int(leg_size), int(slicing)
    if leg_size == 0:
        raise ValueError("leg_size or slicing is zero!")
    if leg_size < slicing:
        sl = [slice(i, i + 1) for i in range(leg_size)]
        return sl, len(sl)
    if slicing <= 1:
        return [slice(0, leg_size)], 1
    bs1, bs2 = leg_size // slicing + 1, leg_size // slicing
    n1 = (leg_size - bs2 * slicing) // (bs1 - bs2)
    sl = [slice(i, i + bs1) for i in range(0, n1 * bs1, bs1)]
    sl += [slice(i, i + bs2) for i in range(n1 * bs1, leg_size, bs2)]
    return sl, len(sl)


def contract_G(A, PX, UM, PY):
    """G_{T,X,Y,k} = sum A_{T,a,b,i} PX_{a,c,X} UM_{i,c,d,k} PY_{b,d,Y} (:379-384).

    Pairwise order A.PX (2 chi^6), .UM over (i,c) (2 chi^7), .PY over (b,d)."""
    W = np.tensordot(A, PX, axes=([1], [0]))                # (T,b,i,c,X)
    W = np.tensordot(W, UM, axes=([2, 3], [0, 1]))          # (T,b,X,d,k)
    W = np.tensordot(W, PY, axes=([1, 3], [0, 1]))          # (T,X,k,Y)
    return np.transpose(W, (0, 1, 3, 2))


def contract_H(VMH, Px, D, Py):
    """H_{k,t,x,y} = sum VMH_{k,j,a,b} Px_{x,a,c} D_{j,t,c,d} Py_{y,b,d} (:400-405).

    Pairwise order VMH.Px (2 chi^6), .D over (j,c) (2 chi^7), .Py over (b,d)."""
    W = np.tensordot(VMH, Px, axes=([2], [1]))              # (k,j,b,x,c)
    W = np.tensordot(W, D, axes=([1, 4], [0, 2]))           # (k,b,x,t,d)
    W = np.tensordot(W, Py, axes=([1, 4], [1, 2]))          # (k,x,t,y)
    return np.transpose(W, (0, 2, 1, 3))


def coarse_graining(A, UM, sM, VMH, D, Px, PX, Py, PY, slicing=100):
    """trg/ATRG_3d_new.py:357-438.  The reference's slice loops only bound
    memory; the sums are evaluated whole here."""
    nT, nt = A.shape[0], D.shape[1]
    nX, nY = PX.shape[2], PY.shape[2]
    nx, ny = Px.shape[0], Py.shape[0]
    nk = UM.shape[3]
    G = contract_G(A, PX, UM, PY)
    H = contract_H(VMH, Px, D, Py)
    kG = min(nT * nX * nY, nk)
    kH = min(nk, nt * nx * ny)
    UG, sG, VGH = svd(G, shape=[[0, 1, 2], [3]], k=kG, truncate_eps=truncate_eps, degeneracy_eps=degeneracy_eps)
    UH, sH, VHH = svd(H, shape=[[0], [1, 2, 3]], k=kH, truncate_eps=truncate_eps, degeneracy_eps=degeneracy_eps)
    K = (sG[:, None] * VGH * sM[None, :]) @ (UH * sH[None, :])
    kK = min(*K.shape)
    UK, sK, VKH = svd(K, shape=[[0], [1]], k=kK, truncate_eps=truncate_eps, degeneracy_eps=degeneracy_eps)
    U = np.tensordot(UG, UK, axes=([3], [0]))
    VH = np.tensordot(VKH, VHH, axes=([1], [0]))
    return U, sK, VH


def atrg_pure_tensor(T, Dcut, direction):
    """trg/ATRG_3d_new.py:441-482."""
    T = leg_transposition(T, "transpose", direction)
    A, D = T.U, T.VH
    B = T.s[:, None, None, None] * T.VH
    C = T.U * T.s[None, None, None, :]
    UM, sM, VMH = leg_swapping(B, C, Dcut, ispure=True)
    Px, PX, Py, PY = squeezer(A, UM, sM, VMH, D, Dcut)
    U, s, VH = coarse_graining(A, UM, sM, VMH, D, Px, PX, Py, PY, slicing=Dcut)
    T = T.update(U, s, VH)
    return leg_transposition(T, "restore", direction)


def normalization(T):
    """trg/ATRG_3d_new.py:954-958: c = |Tr T|, s /= c."""
    c = float(np.abs(T.trace()))
    T.s = T.s / c
    return T, c


def exec_pure_tensor_renorm(T, Dcut, direction, rgstep):
    """trg/ATRG_3d_new.py:967-972."""
    T = atrg_pure_tensor(T, Dcut, direction)
    key = f"({rgstep['X']},{rgstep['Y']},{rgstep['T']})"
    T, c = normalization(T)
    T.normal_const[key] = math.log(c) / 2 ** (sum(rgstep.values()))
    return T


def pure_tensor(T, Dcut, TOT_RGSTEPS):
    """trg/ATRG_3d_new.py:974-1039: cycle T, X, Y until the totals are reached."""
    rgstep = {"X": 0, "Y": 0, "T": 0}
    key = f"({rgstep['X']},{rgstep['Y']},{rgstep['T']})"
    T, c = normalization(T)
    T.normal_const[key] = math.log(c)
    total = sum(TOT_RGSTEPS.values())
    for direction in itertools.cycle("TXY"):
        if sum(rgstep.values()) >= total:
            break
        if rgstep[direction] < TOT_RGSTEPS[direction]:
            rgstep[direction] += 1
            T = exec_pure_tensor_renorm(T, Dcut, direction, rgstep)
    return T


def ln_Z_over_V(T, Dcut, TOT_RGSTEPS):
    """measurement/Potts3_model_3d.py:32-54: sum of normal_const + ln Tr T / V."""
    T = pure_tensor(T, Dcut, TOT_RGSTEPS)
    V = 2 ** sum(TOT_RGSTEPS.values())
    return float(np.real(np.sum(T.get_normalization_const()) + np.log(T.trace()) / V)), T


# ---------------------------------------------------------------------------
# one-point function (impurity) flow
# ---------------------------------------------------------------------------
def new_impuer_tensor_1point(T, T0, Dcut, direction):
    """trg/ATRG_3d_new.py:669-734 (the list form; :614-667 is the one-tensor
    case).  The squeezers come from the pure pair.  Each impurity forms its
    own pair with a pure tensor: it is the lower tensor (C, D) when its
    coordinate along `direction` is even and the upper one (A, B) when odd;
    the pair gets its own leg swap and the pure pair's coarse graining."""
    T = leg_transposition(T, "transpose", direction)
    T0 = [leg_transposition(t, "transpose", direction) for t in T0]
    A, D = T.U, T.VH
    B = T.s[:, None, None, None] * T.VH
    C = T.U * T.s[None, None, None, :]
    UM, sM, VMH = leg_swapping(B, C, Dcut, ispure=True)
    Px, PX, Py, PY = squeezer(A, UM, sM, VMH, D, Dcut)
    for i, t in enumerate(T0):
        if t.loc[direction] % 2 == 0:
            Ai, Bi = T.U, T.s[:, None, None, None] * T.VH
            Ci, Di = t.U * t.s[None, None, None, :], t.VH
        else:
            Ai, Bi = t.U, t.s[:, None, None, None] * t.VH
            Ci, Di = T.U * T.s[None, None, None, :], T.VH
        UMi, sMi, VMHi = leg_swapping(Bi, Ci, Dcut, ispure=False)
        U, s, VH = coarse_graining(Ai, UMi, sMi, VMHi, Di, Px, PX, Py, PY, slicing=Dcut)
        T0[i] = t.update(U, s, VH)
    U, s, VH = coarse_graining(A, UM, sM, VMH, D, Px, PX, Py, PY, slicing=Dcut)
    T = T.update(U, s, VH)
    T = leg_transposition(T, "restore", direction)
    T0 = [leg_transposition(t, "restore", direction) for t in T0]
    return T, T0


def field_expected_value(T, T0, nsteps):
    """trg/ATRG_3d_new.py:1043-1060: ln Z/V and P_i = exp(sum T0_i factors) Tr T0_i / Tr T
    after `nsteps` RG steps (the factor of step n of T is ln c / 2^n)."""
    trT = float(np.real(T.trace()))
    V = 2 ** nsteps
    lnz = float(np.real(np.sum(T.get_normalization_const()))) + math.log(trT) / V
    P = [math.exp(float(np.real(np.sum(t.get_normalization_const())))) * float(np.real(t.trace())) / trT
         for t in T0]
    return lnz, P


def one_point_function(T, T0, Dcut, TOT_RGSTEPS):
    """trg/ATRG_3d_new.py:1079-1189: cycle T, X, Y; after each step
    normalise T (factor ln c / 2^n) and each impurity (factor ln c0 - ln c),
    then halve the impurity coordinates along the step's direction.
    Returns (T, T0, lnZ/V, P)."""
    rgstep = {"X": 0, "Y": 0, "T": 0}
    key = lambda: f"({rgstep['X']},{rgstep['Y']},{rgstep['T']})"
    T, c = normalization(T)
    T.normal_const[key()] = math.log(c)
    for t in T0:
        t, c0 = normalization(t)
        t.normal_const[key()] = math.log(c0) - math.log(c)
    total = sum(TOT_RGSTEPS.values())
    for direction in itertools.cycle("TXY"):
        if sum(rgstep.values()) >= total:
            break
        if rgstep[direction] < TOT_RGSTEPS[direction]:
            rgstep[direction] += 1
            T, T0 = new_impuer_tensor_1point(T, T0, Dcut, direction)
            T, c = normalization(T)
            T.normal_const[key()] = math.log(c) / 2 ** sum(rgstep.values())
            for t in T0:
                t, c0 = normalization(t)
                t.normal_const[key()] = math.log(c0) - math.log(c)
                t.loc[direction] //= 2
    lnz, P = field_expected_value(T, T0, total)
    return T, T0, lnz, P


# ---------------------------------------------------------------------------
# two-point function (impurity) flow
# ---------------------------------------------------------------------------
def _pair_step(T, pairs, Dcut, direction):
    """One step of the pure tensor T and of the impurity pairs: each pair is
    (upper, lower) with None standing for the pure tensor; the upper tensor
    gives A = U and B = s VH, the lower one C = U s and D = VH; the squeezers
    come from the pure pair.  Returns (T, [new impurity per pair]), legs
    restored (tensors are transposed in place, as the reference does)."""
    T = leg_transposition(T, "transpose", direction)
    for up, lo in pairs:
        for t in (up, lo):
            if t is not None:
                leg_transposition(t, "transpose", direction)
    A, D = T.U, T.VH
    B = T.s[:, None, None, None] * T.VH
    C = T.U * T.s[None, None, None, :]
    UM, sM, VMH = leg_swapping(B, C, Dcut, ispure=True)
    Px, PX, Py, PY = squeezer(A, UM, sM, VMH, D, Dcut)
    outs = []
    for up, lo in pairs:
        u = T if up is None else up
        l = T if lo is None else lo
        Ai, Bi = u.U, u.s[:, None, None, None] * u.VH
        Ci, Di = l.U * l.s[None, None, None, :], l.VH
        UMi, sMi, VMHi = leg_swapping(Bi, Ci, Dcut, ispure=False)
        outs.append(coarse_graining(Ai, UMi, sMi, VMHi, Di, Px, PX, Py, PY, slicing=Dcut))
    U, s, VH = coarse_graining(A, UM, sM, VMH, D, Px, PX, Py, PY, slicing=Dcut)
    T = leg_transposition(T.update(U, s, VH), "restore", direction)
    return T, outs


def new_impuer_tensor_2points(T, T0, Tn, Dcut, direction):
    """trg/ATRG_3d_new.py:485-555 as the one-point step applies it: each
    impurity pairs with a pure tensor, as the upper tensor when its
    coordinate along `direction` is odd, the lower one when even.  (The
    reference's even branch, :519-521, contracts U with s and s with VH, and
    :545 passes T as an extra first argument of coarse_graining; both fail
    on the shapes.  This is the one-point step's pairing.)"""
    pairs = [((t, None) if t.loc[direction] % 2 == 1 else (None, t)) for t in (T0, Tn)]
    T, outs = _pair_step(T, pairs, Dcut, direction)
    T0 = leg_transposition(T0.update(*outs[0]), "restore", direction)
    Tn = leg_transposition(Tn.update(*outs[1]), "restore", direction)
    return T, T0, Tn


def new_impuer_tensor_absorb(T, T0, Tn, Dcut, direction):
    """trg/ATRG_3d_new.py:557-612: the impurities form one pair (the one at the
    odd coordinate is the upper tensor) and T0 becomes their merged impurity."""
    up, lo = (Tn, T0) if T0.loc[direction] % 2 == 0 else (T0, Tn)
    T, outs = _pair_step(T, [(up, lo)], Dcut, direction)
    leg_transposition(Tn, "restore", direction)
    T0 = leg_transposition(T0.update(*outs[0]), "restore", direction)
    return T, T0


def two_point_function(T, T0, Tn, Dcut, TOT_RGSTEPS):
    """trg/ATRG_3d_new.py:1333-1489: cycle T, X, Y.  While the impurities are
    apart after a step, the two-point step with T0's factor ln c0 + ln cn -
    2 ln c; when they meet, the absorb step (factor ln c0 - ln c); afterwards
    the one-point step.  Returns (T, T0, lnZ/V, G) with
    G = exp(sum T0 factors) Tr T0 / Tr T."""
    rgstep = {"X": 0, "Y": 0, "T": 0}
    key = lambda: f"({rgstep['X']},{rgstep['Y']},{rgstep['T']})"
    T, c = normalization(T)
    T0, c0 = normalization(T0)
    Tn, c1 = normalization(Tn)
    T.normal_const[key()] = math.log(c)
    T0.normal_const[key()] = math.log(c0) + math.log(c1) - 2 * math.log(c)
    total = sum(TOT_RGSTEPS.values())
    for direction in itertools.cycle("TXY"):
        if sum(rgstep.values()) >= total:
            break
        if rgstep[direction] >= TOT_RGSTEPS[direction]:
            continue
        rgstep[direction] += 1
        n = sum(rgstep.values())
        if Tn is not None:
            after0 = dict(T0.loc, **{direction: T0.loc[direction] // 2})
            aftern = dict(Tn.loc, **{direction: Tn.loc[direction] // 2})
            if sum((after0[k] - aftern[k]) ** 2 for k in after0) > 0:
                T, T0, Tn = new_impuer_tensor_2points(T, T0, Tn, Dcut, direction)
                T, c = normalization(T)
                T0, c0 = normalization(T0)
                Tn, c1 = normalization(Tn)
                T.normal_const[key()] = math.log(c) / 2 ** n
                T0.normal_const[key()] = math.log(c0) + math.log(c1) - 2 * math.log(c)
                T0.loc[direction] //= 2
                Tn.loc[direction] //= 2
                continue
            T, T0 = new_impuer_tensor_absorb(T, T0, Tn, Dcut, direction)
            Tn = None
        else:
            T, (T0,) = new_impuer_tensor_1point(T, [T0], Dcut, direction)
        T, c = normalization(T)
        T0, c0 = normalization(T0)
        T.normal_const[key()] = math.log(c) / 2 ** n
        T0.normal_const[key()] = math.log(c0) - math.log(c)
        T0.loc[direction] //= 2
    lnz, (G,) = field_expected_value(T, [T0], total)
    return T, T0, lnz, G


# ---------------------------------------------------------------------------
# diagnostics (trg/ATRG_3d_new.py:768-951); the tensor is T_{TXYtxy} = U s VH
# ---------------------------------------------------------------------------
def find_max(T, max_nth):
    """:768-816: the max_nth entries of the full tensor with the largest |value|
    and their 6-leg indices (T, X, Y, t, x, y).  The reference prints them and
    calls sys.exit(0) before returning; this returns them.  Evaluated by the
    same (x, y) slices, each Tsub[T,X,Y,t] = sum_i U s VH[:, :, x, y]."""
    chiT, chiX, chiY, _ = T.U.shape
    vals, idx = [], []
    for ix in range(chiX):
        for iy in range(chiY):
            sub = np.einsum("TXYi,i,it->TXYt", T.U, T.s, T.VH[:, :, ix, iy]).ravel()
            vals.append(sub)
            k = np.arange(sub.size)
            TT, XX, YY, tt = np.unravel_index(k, (chiT, chiX, chiY, chiT))
            idx.append(np.ravel_multi_index((TT, XX, YY, tt, np.full_like(k, ix), np.full_like(k, iy)),
                                            (chiT, chiX, chiY, chiT, chiX, chiY)))
    vals = np.concatenate(vals)
    idx = np.concatenate(idx)
    order = np.argsort(-np.abs(vals), kind="stable")[:max_nth]
    ind = np.array(np.unravel_index(idx[order], (chiT, chiX, chiY, chiT, chiX, chiY))).T
    return vals[order], ind


def tensor_diagonal(T):
    """T[i,j,k,i,j,k] = sum_a U[i,j,k,a] s_a VH[a,i,j,k] (:836-838)."""
    return np.einsum("ijka,a,aijk->ijk", T.U, T.s, T.VH).ravel()


def find_1stto3rdmax(T):
    """:818-852: the three diagonal entries of largest |value|."""
    d = tensor_diagonal(T)
    return d[np.argsort(-np.abs(d), kind="stable")[:3]]


def density_matrix(T, leg):
    """The partial traces of :856 and :884-923: leg "t" -> Tt (trace over x, y),
    "x" -> Xx (over t, y), "y" -> Yy (over t, x)."""
    spec = {"t": "TXYi,i,itXY->Tt", "x": "TXYi,i,iTxY->Xx", "y": "TXYi,i,iTXy->Yy"}[leg]
    return np.einsum(spec, T.U, T.s, T.VH)


def cal_scaling_dimensions(T):
    """:854-877: eigh of the T-density matrix (the lower triangle, like
    cupy.linalg.eigh), descending, floored at 1e-16, padded to Dcut,
    log(e0 / e) / (2 pi)."""
    e = np.linalg.eigh(density_matrix(T, "t"), UPLO="L")[0][::-1].copy()
    e[e < 1e-16] = 1e-16
    e = np.pad(e, (0, T.Dcut - len(e)), constant_values=1e-16)
    return np.log(np.max(e) / e) / (2 * np.pi)


def cal_coorelation(T, rgstep):
    """:879-951: singular values of the three density matrices normalised by
    their largest, and ln xi = n_leg ln 2 - ln(-1 / ln e_1) per leg.  As in the
    reference, the y branch tests the x leg's e_1 against 1e-200 (:944)."""
    out = {}
    for leg, key in (("t", "T"), ("x", "X"), ("y", "Y")):
        e = np.linalg.svd(density_matrix(T, leg), compute_uv=False)
        if len(e) < T.Dcut:
            e = np.pad(e, (0, T.Dcut - len(e)))
        out[leg] = e / np.max(e)
    lnxi = {}
    for leg, key in (("t", "T"), ("x", "X"), ("y", "Y")):
        cond = out["x"][1] if leg == "y" else out[leg][1]
        e1 = out[leg][1] if cond > 1e-200 else 1e-200
        lnxi[leg] = rgstep[key] * np.log(2) - np.log(-1 / np.log(e1))
    return out, lnxi
