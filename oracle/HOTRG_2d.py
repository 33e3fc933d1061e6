"""NumPy restatement of trg/HOTRG_2d.py (test infrastructure only).

Leg order of T is (x, x', y, y').  A "Y" step merges two tensors along y and
truncates the doubled x legs; an "X" step does the same after the
``abcd->cdab`` rearrangement (trg/HOTRG_2d.py:8-20).

Generalisation (documented, exercised by tests): the reference reshapes the
isometry to (Dcut, Dcut, Dcut) (:53), so callers must zero-pad the bond to
Dcut, and it leaves ``LDRU`` unassigned when the environment has no more than
Dcut eigenvalues (:39-40).  Here the isometry is (D, D, min(D^2, Dcut)) and the
untruncated case keeps all eigenvectors with ``LDRU = "LD"``.  Zero padding and
bond growth give the same ln Z/V (tests/test_oracle_hotrg.py).
"""
import numpy as np


def _conj(A):
    return np.conj(A) if np.iscomplexobj(A) else A


def __tensor_legs_rearrange__(T, direction):
    """trg/HOTRG_2d.py:8-13."""
    if direction in ("Y", "y"):
        return T
    return np.ascontiguousarray(np.transpose(T, (2, 3, 0, 1)))


def __tensor_legs_restore__(T, direction):
    """trg/HOTRG_2d.py:15-20 (``cdab->abcd`` is the same involution)."""
    return __tensor_legs_rearrange__(T, direction)


def environments(T):
    """The two environments of trg/HOTRG_2d.py:26-27 as (D^2 x D^2) matrices.

    MM1 = aijk,blkm,cijn,dlnm->abcd (left x legs), MM2 = iajk,lbkm,icjn,ldnm->abcd
    (right x legs).  Evaluated as two partial Gram products and one matrix
    product, e.g. for MM1: A[a,c,k,n] = sum_ij T[a,i,j,k] T*[c,i,j,n],
    B[b,d,k,n] = sum_lm T[b,l,k,m] T*[d,l,n,m], MM1[ab,cd] = sum_kn A B.
    """
    Dx, _, Dy, _ = T.shape
    Tc = _conj(T)
    out = []
    for leg in (0, 1):
        Tp = T if leg == 0 else np.transpose(T, (1, 0, 2, 3))
        Tpc = Tc if leg == 0 else np.transpose(Tc, (1, 0, 2, 3))
        # A[a,k,c,n] = sum_{i,j} T[a,i,j,k] T*[c,i,j,n]
        A = np.tensordot(Tp, Tpc, axes=([1, 2], [1, 2]))       # (a,k,c,n)
        # B[b,k,d,n] = sum_{l,m} T[b,l,k,m] T*[d,l,n,m]
        B = np.tensordot(Tp, Tpc, axes=([1, 3], [1, 3]))       # (b,k,d,n)
        # MM[a,b,c,d] = sum_{k,n} A[a,k,c,n] B[b,k,d,n]
        MM = np.tensordot(A, B, axes=([1, 3], [1, 3]))          # (a,c,b,d)
        MM = np.transpose(MM, (0, 2, 1, 3)).reshape(Dx * Dx, Dx * Dx)
        out.append(MM)
    return out


def __contraction_projector__(T, Dcut):
    """trg/HOTRG_2d.py:22-55: two environments, eigh, LD/RU choice.

    Returns (U, LDRU) with U of shape (Dx, Dx, chi'), chi' = min(Dx^2, Dcut);
    eigenvectors are kept in eigh's ascending order (:47,50).
    """
    Dx = T.shape[0]
    MM1, MM2 = environments(T)
    e1, v1 = np.linalg.eigh(MM1)
    e2, v2 = np.linalg.eigh(MM2)
    D = len(e1)
    if D <= Dcut:
        U = v1
        LDRU = "LD"
    else:
        eps1 = np.sum(e1[:D - Dcut])
        eps2 = np.sum(e2[:D - Dcut])
        if eps1 < eps2:
            LDRU = "LD"
            U = v1[:, D - Dcut:]
        else:
            LDRU = "RU"
            U = v2[:, D - Dcut:]
    return U.reshape(Dx, Dx, U.shape[1]), LDRU


def merge(T1, T2, Ua, Ub, nslice=None):
    """acke,bdel,abi,cdj->ijkl (trg/HOTRG_2d.py:65/67).

    Order: Z1 = T1 . Ua over a (chi^6 flop), Z2 = Z1 . T2 over (e, b)
    (2 chi^7 flop), out = Z2 . Ub over (c, d).  Sliced over the output
    index i to bound the chi^5 intermediates.
    """
    Dx = T1.shape[0]
    chi = Ua.shape[2]
    if nslice is None:
        per = T1.shape[1] * T1.shape[2] * T1.shape[3] * Dx * 16
        nslice = max(1, int(np.ceil(chi * per / 2 ** 30)))
    bounds = np.linspace(0, chi, min(nslice, chi) + 1).astype(int)
    out = np.empty((chi, Ub.shape[2], T1.shape[2], T2.shape[3]),
                   dtype=np.result_type(T1, T2, Ua, Ub))
    for i0, i1 in zip(bounds[:-1], bounds[1:]):
        Z1 = np.tensordot(T1, Ua[:, :, i0:i1], axes=([0], [0]))     # (c,k,e,b,i)
        Z2 = np.tensordot(Z1, T2, axes=([2, 3], [2, 0]))            # (c,k,i,d,l)
        R = np.tensordot(Z2, Ub, axes=([0, 3], [0, 1]))             # (k,i,l,j)
        out[i0:i1] = np.transpose(R, (1, 3, 0, 2))
    return out


def _merge_ldru(T1, T2, U, LDRU):
    if LDRU == "LD":
        return merge(T1, T2, _conj(U), U)
    return merge(T1, T2, U, _conj(U))


def new_pure_tensor(T, Dcut, direction):
    """trg/HOTRG_2d.py:57-72."""
    T = __tensor_legs_rearrange__(T, direction)
    U, LDRU = __contraction_projector__(T, Dcut)
    T = _merge_ldru(T, T, U, LDRU)
    return __tensor_legs_restore__(T, direction)


def new_impuer_tensor_2imp(T, Timp0, Timp1, nx, ny, Dcut, direction):
    """trg/HOTRG_2d.py:75-128."""
    T = __tensor_legs_rearrange__(T, direction)
    Timp0 = __tensor_legs_rearrange__(Timp0, direction)
    Timp1 = __tensor_legs_rearrange__(Timp1, direction)
    U, LDRU = __contraction_projector__(T, Dcut)
    Timp0 = _merge_ldru(Timp0, T, U, LDRU)
    isy = direction in ("Y", "y")
    odd = (ny % 2 == 1) if isy else (nx % 2 == 1)
    Timp1 = _merge_ldru(T, Timp1, U, LDRU) if odd else _merge_ldru(Timp1, T, U, LDRU)
    T = _merge_ldru(T, T, U, LDRU)
    return (__tensor_legs_rearrange__(T, direction), __tensor_legs_rearrange__(Timp0, direction),
            __tensor_legs_rearrange__(Timp1, direction))


def new_impuer_tensor_2to1imp(T, Timp0, Timp1, Dcut, direction):
    """trg/HOTRG_2d.py:130-153."""
    T = __tensor_legs_rearrange__(T, direction)
    Timp0 = __tensor_legs_rearrange__(Timp0, direction)
    Timp1 = __tensor_legs_rearrange__(Timp1, direction)
    U, LDRU = __contraction_projector__(T, Dcut)
    Timp0 = _merge_ldru(Timp0, Timp1, U, LDRU)
    T = _merge_ldru(T, T, U, LDRU)
    return __tensor_legs_rearrange__(T, direction), __tensor_legs_rearrange__(Timp0, direction)


def new_impuer_tensor_1imp(T, Timp0, Dcut, direction):
    """trg/HOTRG_2d.py:155-177."""
    T = __tensor_legs_rearrange__(T, direction)
    Timp0 = __tensor_legs_rearrange__(Timp0, direction)
    U, LDRU = __contraction_projector__(T, Dcut)
    Timp0 = _merge_ldru(Timp0, T, U, LDRU)
    T = _merge_ldru(T, T, U, LDRU)
    return __tensor_legs_rearrange__(T, direction), __tensor_legs_rearrange__(Timp0, direction)


def __tensor_normalization__(T):
    """trg/HOTRG_2d.py:181-185: c = max |T|."""
    c = np.max(np.abs(T))
    return T / c, c


def pure_tensor_renorm(T, Dcut, XLOOPS, YLOOPS):
    """trg/HOTRG_2d.py:187-216: Y then X steps, lnf[n] = ln c_n / 2^n."""
    nx = ny = n = 0
    lnf = np.zeros(XLOOPS + YLOOPS + 1, dtype=np.float64)
    T, c = __tensor_normalization__(T)
    lnf[0] = np.log(c)
    while nx < XLOOPS or ny < YLOOPS:
        if ny < YLOOPS:
            ny += 1
            n += 1
            T = new_pure_tensor(T, Dcut, "Y")
            T, c = __tensor_normalization__(T)
            lnf[n] = np.log(c) / 2 ** n
        if nx < XLOOPS:
            nx += 1
            n += 1
            T = new_pure_tensor(T, Dcut, "X")
            T, c = __tensor_normalization__(T)
            lnf[n] = np.log(c) / 2 ** n
    return T, lnf


def _norm3(T, Timp0, Timp1=None):
    T, c = __tensor_normalization__(T)
    Timp0, c0 = __tensor_normalization__(Timp0)
    if Timp1 is None:
        return T, Timp0, np.log(c), np.log(c0) - np.log(c)
    Timp1, c1 = __tensor_normalization__(Timp1)
    return T, Timp0, Timp1, np.log(c), np.log(c0) + np.log(c1) - 2 * np.log(c)


def two_point_func_renorm(T, Timp0, Timp1, nx, ny, Dcut, XLOOPS, YLOOPS):
    """trg/HOTRG_2d.py:219-297 (two impurities separated by (nx, ny))."""
    cx = cy = n = 0
    lnf = np.zeros(XLOOPS + YLOOPS + 1)
    lnf_imp = np.zeros(XLOOPS + YLOOPS + 1)
    T, Timp0, Timp1, lc, limp = _norm3(T, Timp0, Timp1)
    lnf[0], lnf_imp[0] = lc, limp
    while cx < XLOOPS or cy < YLOOPS:
        if cy < YLOOPS:
            cy += 1
            n += 1
            if int(nx ** 2 + ny ** 2) > 1 or (nx == 1 and ny == 0):
                T, Timp0, Timp1 = new_impuer_tensor_2imp(T, Timp0, Timp1, nx, ny, Dcut, "Y")
                T, Timp0, Timp1, lc, limp = _norm3(T, Timp0, Timp1)
                lnf[n], lnf_imp[n] = lc / 2 ** n, limp
            if nx == 0 and ny == 1:
                T, Timp0 = new_impuer_tensor_2to1imp(T, Timp0, Timp1, Dcut, "Y")
                T, Timp0, lc, limp = _norm3(T, Timp0)
                lnf[n], lnf_imp[n] = lc / 2 ** n, limp
            if nx == 0 and ny == 0:
                T, Timp0 = new_impuer_tensor_1imp(T, Timp0, Dcut, "Y")
                T, Timp0, lc, limp = _norm3(T, Timp0)
                lnf[n], lnf_imp[n] = lc / 2 ** n, limp
            ny = ny >> 1
        if cx < XLOOPS:
            cx += 1
            n += 1
            if int(nx ** 2 + ny ** 2) > 1 or (nx == 0 and ny == 1):
                T, Timp0, Timp1 = new_impuer_tensor_2imp(T, Timp0, Timp1, nx, ny, Dcut, "X")
                T, Timp0, Timp1, lc, limp = _norm3(T, Timp0, Timp1)
                lnf[n], lnf_imp[n] = lc / 2 ** n, limp
            if nx == 1 and ny == 0:
                T, Timp0 = new_impuer_tensor_2to1imp(T, Timp0, Timp1, Dcut, "X")
                T, Timp0, lc, limp = _norm3(T, Timp0)
                lnf[n], lnf_imp[n] = lc / 2 ** n, limp
            if nx == 0 and ny == 0:
                T, Timp0 = new_impuer_tensor_1imp(T, Timp0, Dcut, "X")
                T, Timp0, lc, limp = _norm3(T, Timp0)
                lnf[n], lnf_imp[n] = lc / 2 ** n, limp
            nx = nx >> 1
    return T, Timp0, lnf, lnf_imp


def ynearest_two_point_func_renorm(T, Timp0, Timp1, Dcut, XLOOPS, YLOOPS):
    """trg/HOTRG_2d.py:300-348 (nearest neighbours along y)."""
    cx = cy = n = 0
    lnf = np.zeros(XLOOPS + YLOOPS + 1)
    lnf_imp = np.zeros(XLOOPS + YLOOPS + 1)
    T, Timp0, Timp1, lc, limp = _norm3(T, Timp0, Timp1)
    lnf[0], lnf_imp[0] = lc, limp
    while cx < XLOOPS or cy < YLOOPS:
        if cy < YLOOPS:
            cy += 1
            n += 1
            if n == 1:
                T, Timp0 = new_impuer_tensor_2to1imp(T, Timp0, Timp1, Dcut, "Y")
            else:
                T, Timp0 = new_impuer_tensor_1imp(T, Timp0, Dcut, "Y")
            T, Timp0, lc, limp = _norm3(T, Timp0)
            lnf[n], lnf_imp[n] = lc / 2 ** n, limp
        if cx < XLOOPS:
            cx += 1
            n += 1
            T, Timp0 = new_impuer_tensor_1imp(T, Timp0, Dcut, "X")
            T, Timp0, lc, limp = _norm3(T, Timp0)
            lnf[n], lnf_imp[n] = lc / 2 ** n, limp
    return T, Timp0, lnf, lnf_imp


def ln_Z_over_V(T, Dcut, XLOOPS, YLOOPS):
    """Caller-side assembly (measurement/SU2_pcm_measurement.py:30-36):
    ln Z/V = sum lnf + ln Tr T / 2^(XLOOPS+YLOOPS), Tr over (xx)(yy)."""
    T, lnf = pure_tensor_renorm(T, Dcut, XLOOPS, YLOOPS)
    tr = np.einsum("aabb", T)
    return float(np.real(np.sum(lnf) + np.log(tr) / 2 ** (XLOOPS + YLOOPS)))
