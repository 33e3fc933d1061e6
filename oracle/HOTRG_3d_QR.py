"""NumPy restatement of trg/HOTRG_3d_QR.py (test infrastructure only).

The 3D HOTRG the reference's own 3D drivers run (measurement/Ising_model_3d.py:101-137,
measurement/Potts3_model_3d.py:22-54): a "Z" step merges a lower tensor T0 and
an upper tensor T1 along z, and the doubled x and y legs are cut by projectors
from one of three squeezers (trg/HOTRG_3d_QR.py:394-406):

* "original" (:120-193): isometries from `eigh` of the x-leg and y-leg
  environments with the truncation rule of utility/truncated_svd.py;
* "fix" (:74-118): isometries from the X-leg and y-leg environments;
* "QR" (:195-281): oblique projectors from square roots of the two
  environments on either side of each cut and a truncated SVD of their
  product (the `generalized_inverse`).

Leg order of T is (x, X, y, Y, z, Z).  Steps cycle "ZXY" (:780), and every
normalisation factor is stored as ln c / 2^(steps so far) under the key
"(X,Y,Z)" (:734-745).

Decided reference behaviour (documented):
* "original" compares the discarded sums eLX[k:Dcut*Dcut] and eRx[k:...] of
  eigenvalue vectors that `eigh` has already truncated to length <= k, so
  both sums are over empty slices and equal 0, `errLX < errRx` is False, and
  the x projectors always come from the x-leg environment (RxRx^dag) and the
  y projectors from the y-leg environment (LyLy^dag) (:164-189).
* The module reads DGENERACY_EPS, HOTRG_SQUEEZER and OUTPUT_DIR from the
  environment and Dcut from sys.argv[4] at import (:19-27); here they are
  arguments (`degeneracy_eps`, `squeezer_type`) and `slicing` is a
  parameter of `coarse_graining` that does not change the result.
* Real tensors only: the reference conjugates, which is the identity here.
* The per-step diagnostics (cal_X, find_1stto3rdmax, cal_scaling_dimensions,
  cal_coorelation, :559-713) are functions of the tensor below that return
  their values; the prints, timings and files are dropped.
"""
import math
from itertools import cycle, product

import numpy as np

from .HOTRG_3d import merge
from .truncated_svd import eigh, svd

TRUNCATE_EPS = 1e-10                 # trg/HOTRG_3d_QR.py:27


def leg_transposition(T, do_what="transpose", direction="Z"):
    """trg/HOTRG_3d_QR.py:29-72: X: xXyYzZ -> yYzZxX, Y: xXyYzZ -> zZxXyY."""
    if direction in ("Z", "z"):
        return T
    fwd = {"X": (2, 3, 4, 5, 0, 1), "Y": (4, 5, 0, 1, 2, 3)}
    bwd = {"X": (4, 5, 0, 1, 2, 3), "Y": (2, 3, 4, 5, 0, 1)}
    axes = (fwd if do_what == "transpose" else bwd)[direction.upper()]
    return np.ascontiguousarray(np.transpose(T, axes))


def environment(Tlow, Tup, leg):
    """E[(i,j),(k,l)] = sum Tlow[..i..,m] Tup[..j..,m,..] Tlow[..k..,n] Tup[..l..,n,..]
    over every leg except `leg` of both tensors, with Tlow's Z leg bonded to
    Tup's z leg (the four `oe.contract` calls of trg/HOTRG_3d_QR.py:212-215;
    leg 1 = LXdagLX, 0 = RxRxdag, 2 = LydagLy, 3 = RYRYdag).

    Evaluated as A[i,m,k,n] = Gram of Tlow over its four other legs (not Z),
    B[j,m,l,n] = Gram of Tup over its four other legs (not z), E = sum_{m,n} A B.
    """
    others = [a for a in range(4) if a != leg]
    A = np.tensordot(Tlow, Tlow, axes=(others + [4], others + [4]))      # (i, m, k, n)
    B = np.tensordot(Tup, Tup, axes=(others + [5], others + [5]))        # (j, m, l, n)
    E = np.tensordot(A, B, axes=([1, 3], [1, 3]))                        # (i, k, j, l)
    Di, Dj = Tlow.shape[leg], Tup.shape[leg]
    return np.transpose(E, (0, 2, 1, 3)).reshape(Di * Dj, Di * Dj)


def squeezer_fix(T0, T1, T2, T3, Dcut, degeneracy_eps=0.0):
    """trg/HOTRG_3d_QR.py:74-118."""
    LX = environment(T0, T1, 1)
    Ly = environment(T0, T1, 2)
    TdagT = np.trace(LX)
    eLX, ULX = eigh(LX, shape=[[0], [1]], k=min(*LX.shape, Dcut), degeneracy_eps=degeneracy_eps,
                    truncate_eps=TRUNCATE_EPS)
    eLy, ULy = eigh(Ly, shape=[[0], [1]], k=min(*Ly.shape, Dcut), degeneracy_eps=degeneracy_eps,
                    truncate_eps=TRUNCATE_EPS)
    k = min(len(eLX), Dcut)
    PX = ULX[:, :k].reshape(T0.shape[1], T1.shape[1], k)
    Px = ULX.T[:k].reshape(k, T0.shape[0], T1.shape[0])
    k = min(len(eLy), Dcut)
    PY = ULy.T[:k].reshape(k, T3.shape[3], T2.shape[3])
    Py = ULy[:, :k].reshape(T0.shape[2], T1.shape[2], k)
    return Px, PX, Py, PY, TdagT


def squeezer_original(T0, T1, T2, T3, Dcut, degeneracy_eps=0.0):
    """trg/HOTRG_3d_QR.py:120-193 (side choice decided as in the module
    docstring: the compared sums are over empty slices)."""
    LX = environment(T0, T1, 1)
    Rx = environment(T3, T2, 0)
    Ly = environment(T0, T1, 2)
    RY = environment(T3, T2, 3)
    TdagT = np.trace(LX)
    dec = []
    for E in (LX, Rx, Ly, RY):
        dec.append(eigh(E, shape=[[0], [1]], k=min(*E.shape, Dcut), degeneracy_eps=degeneracy_eps,
                        truncate_eps=TRUNCATE_EPS))
    (eLX, ULX), (eRx, URx), (eLy, ULy), (eRY, URY) = dec
    k = min(len(eLX), Dcut)
    errLX, errRx = np.sum(eLX[k:Dcut * Dcut]), np.sum(eRx[k:Dcut * Dcut])
    U = ULX if errLX < errRx else URx
    PX = U[:, :k].reshape(T0.shape[1], T1.shape[1], k)
    Px = U.T[:k].reshape(k, T0.shape[0], T1.shape[0])
    k = min(len(eLy), Dcut)
    errLy, errRY = np.sum(eLy[k:Dcut * Dcut]), np.sum(eRY[k:Dcut * Dcut])
    U = URY if errLy < errRY else ULy
    PY = U.T[:k].reshape(k, T3.shape[3], T2.shape[3])
    Py = U[:, :k].reshape(T0.shape[2], T1.shape[2], k)
    return Px, PX, Py, PY, TdagT


def _psd_roots(E):
    """(sqrt(S) V^H, U sqrt(S)) from the SVD of a PSD environment (:227-230, :242-243, :251-252)."""
    U, S, VH = np.linalg.svd(E)
    return np.sqrt(S)[:, None] * VH, U * np.sqrt(S)[None, :]


def squeezer_QR(T0, T1, T2, T3, Dcut, degeneracy_eps=0.0):
    """trg/HOTRG_3d_QR.py:195-281: PX = RRx Vx S^-1/2, Px = S^-1/2 Ux^H RLx with
    RLx = sqrt(S_LX) V_LX^H, RRx = U_Rx sqrt(S_Rx) and (Ux, S, Vx) the
    truncated SVD of RLx RRx (`generalized_inverse`, :249-256); likewise for y."""
    LX = environment(T0, T1, 1)
    Rx = environment(T3, T2, 0)
    Ly = environment(T0, T1, 2)
    RY = environment(T3, T2, 3)
    TdagT = np.trace(LX)

    def generalized_inverse(A):
        k = min(*A.shape, Dcut)
        U, S, VH = svd(A, shape=[[0], [1]], k=k, degeneracy_eps=degeneracy_eps, truncate_eps=TRUNCATE_EPS)
        return U.T, 1.0 / S, VH.T

    RLx, _ = _psd_roots(LX)
    _, RRx = _psd_roots(Rx)
    UxH, Sxinv, Vx = generalized_inverse(RLx @ RRx)
    PX = (RRx @ Vx) * np.sqrt(Sxinv)[None, :]
    Px = np.sqrt(Sxinv)[:, None] * (UxH @ RLx)
    PX = PX.reshape(T0.shape[1], T1.shape[1], PX.shape[1])
    Px = Px.reshape(Px.shape[0], T0.shape[0], T1.shape[0])

    RLy, _ = _psd_roots(Ly)
    _, RRy = _psd_roots(RY)
    UyH, Syinv, Vy = generalized_inverse(RLy @ RRy)
    Py = (RRy @ Vy) * np.sqrt(Syinv)[None, :]
    PY = np.sqrt(Syinv)[:, None] * (UyH @ RLy)
    Py = Py.reshape(T0.shape[2], T1.shape[2], Py.shape[1])
    PY = PY.reshape(PY.shape[0], T0.shape[3], T1.shape[3])
    return Px, PX, Py, PY, TdagT


def squeezer(T0, T1, T2, T3, Dcut, squeezer_type, degeneracy_eps=0.0):
    """trg/HOTRG_3d_QR.py:394-406."""
    f = {"original": squeezer_original, "QR": squeezer_QR, "fix": squeezer_fix}[squeezer_type]
    return f(T0, T1, T2, T3, Dcut, degeneracy_eps)


def slice_legs(leg_size, slicing):
    """trg/HOTRG_3d_QR.py:284-316: split range(leg_size) into `slicing`
    contiguous slices (n1 of length leg_size//slicing, the rest one longer);
    one slice per index when leg_size < slicing, a single slice when
    slicing <= 1."""
    leg_size, slicing = int(leg_size), int(slicing)
    if leg_size == 0:
        raise ValueError("leg_size or slicing is zero!")
    if leg_size < slicing:
        return [slice(i, i + 1) for i in range(leg_size)], leg_size
    if slicing <= 1:
        return [slice(0, leg_size)], 1
    bs1 = leg_size // slicing
    bs2 = bs1 + 1
    n1 = (leg_size - bs2 * slicing) // (bs1 - bs2)
    out = [slice(i, i + bs1) for i in range(0, n1 * bs1, bs1)]
    out += [slice(i, i + bs2) for i in range(n1 * bs1, leg_size, bs2)]
    return out, len(out)


def coarse_graining(T0, T1, Px, PX, Py, PY, slicing=None, Dcut=None):
    """adhezi,bcgfiZ,xab,dcX,hgy,Yef->xXyYzZ (trg/HOTRG_3d_QR.py:318-391).

    The reference sums the contraction over slices of five legs; the sum does
    not depend on the slicing, so this evaluates it with the pairwise chain of
    the HOTRG 3D merge (T0.Px -> .Py -> .T1 -> .PX -> .PY): in its
    ijklez,mnopzf,ima,jnb,koc,lpd form the projectors are
    ima = Px^(1,2,0), jnb = PX, koc = Py, lpd = PY^(1,2,0)."""
    return merge(T0, T1, np.transpose(Px, (1, 2, 0)), PX, Py, np.transpose(PY, (1, 2, 0)))


def coarse_graining_einsum(T0, T1, Px, PX, Py, PY):
    """The same contraction as one einsum (small sizes; cross-check)."""
    return np.einsum("adhezi,bcgfiZ,xab,dcX,hgy,Yef->xXyYzZ", T0, T1, Px, PX, Py, PY, optimize=True)


def normalization(T):
    """trg/HOTRG_3d_QR.py:715-719: c = |Tr T| over (xx)(yy)(zz)."""
    c = abs(float(np.einsum("xxyyzz", T)))
    return T / c, c


def new_pure_tensor(T, Dcut, direction, squeezer_type="QR", degeneracy_eps=0.0):
    """trg/HOTRG_3d_QR.py:408-425 on the tensor array (the reference passes the dict)."""
    T = leg_transposition(T, "transpose", direction)
    Px, PX, Py, PY, _ = squeezer(T, T, T, T, Dcut, squeezer_type, degeneracy_eps)
    T = coarse_graining(T, T, Px, PX, Py, PY)
    return leg_transposition(T, "restore", direction)


def pure_tensor(T, Dcut, TOT_RGSTEPS, squeezer_type="QR", degeneracy_eps=0.0):
    """trg/HOTRG_3d_QR.py:734-795 on the dict {'Tensor', 'loc', 'factor'}."""
    rgstep = {"X": 0, "Y": 0, "Z": 0}
    key = lambda: f"({rgstep['X']},{rgstep['Y']},{rgstep['Z']})"
    T["Tensor"], c = normalization(T["Tensor"])
    T["factor"][key()] = math.log(c) / 2 ** sum(rgstep.values())
    for direction in cycle("ZXY"):
        if sum(rgstep.values()) >= sum(TOT_RGSTEPS.values()):
            break
        if rgstep[direction] < TOT_RGSTEPS[direction]:
            rgstep[direction] += 1
            T["Tensor"] = new_pure_tensor(T["Tensor"], Dcut, direction, squeezer_type, degeneracy_eps)
            T["Tensor"], c = normalization(T["Tensor"])
            T["factor"][key()] = math.log(c) / 2 ** sum(rgstep.values())
    return T


def ln_Z_over_V(T0, Dcut, TOT_RGSTEPS, squeezer_type="QR", degeneracy_eps=0.0):
    """ln Z / V as measurement/Ising_model_3d.py:101-137 assembles it:
    sum of the factors + ln Tr T / 2^(total steps)."""
    T = pure_tensor({"Tensor": T0, "loc": {}, "factor": {}}, Dcut, dict(TOT_RGSTEPS), squeezer_type, degeneracy_eps)
    V = 2 ** sum(TOT_RGSTEPS.values())
    return float(sum(T["factor"].values()) + np.log(np.einsum("xxyyzz", T["Tensor"])) / V), T


# ---------------------------------------------------------------------------
# one-point function (trg/HOTRG_3d_QR.py:520-557, 798-908)
# ---------------------------------------------------------------------------
def new_impuer_tensor_1point(T, T0, loc, Dcut, direction, squeezer_type="QR", degeneracy_eps=0.0):
    """trg/HOTRG_3d_QR.py:520-557 on arrays: the projectors come from the pure
    tensor; the impurity T0 is the upper tensor of its pair when its
    coordinate along `direction` is odd, the lower one when even."""
    T = leg_transposition(T, "transpose", direction)
    T0 = leg_transposition(T0, "transpose", direction)
    Px, PX, Py, PY, _ = squeezer(T, T, T, T, Dcut, squeezer_type, degeneracy_eps)
    if loc[direction] % 2 == 1:
        T0 = coarse_graining(T, T0, Px, PX, Py, PY)
    else:
        T0 = coarse_graining(T0, T, Px, PX, Py, PY)
    T = coarse_graining(T, T, Px, PX, Py, PY)
    return leg_transposition(T, "restore", direction), leg_transposition(T0, "restore", direction)


def one_point_function(T, T0, Dcut, TOT_RGSTEPS, squeezer_type="QR", degeneracy_eps=0.0):
    """trg/HOTRG_3d_QR.py:859-908 on the dicts {'Tensor', 'loc', 'factor'}:
    T0['factor'] holds ln c0 - ln c per step; returns (T, T0, P) with
    P = exp(sum T0 factors) Tr T0 / Tr T after the last step
    (cal_field_expected_value, :798-831)."""
    rgstep = {"X": 0, "Y": 0, "Z": 0}
    key = lambda: f"({rgstep['X']},{rgstep['Y']},{rgstep['Z']})"
    T["Tensor"], c = normalization(T["Tensor"])
    T0["Tensor"], c0 = normalization(T0["Tensor"])
    T["factor"][key()] = math.log(c)
    T0["factor"][key()] = math.log(c0) - math.log(c)
    for direction in cycle("ZXY"):
        if sum(rgstep.values()) >= sum(TOT_RGSTEPS.values()):
            break
        if rgstep[direction] < TOT_RGSTEPS[direction]:
            rgstep[direction] += 1
            T["Tensor"], T0["Tensor"] = new_impuer_tensor_1point(T["Tensor"], T0["Tensor"], T0["loc"], Dcut,
                                                                 direction, squeezer_type, degeneracy_eps)
            T["Tensor"], c = normalization(T["Tensor"])
            T0["Tensor"], c0 = normalization(T0["Tensor"])
            T["factor"][key()] = math.log(c) / 2 ** sum(rgstep.values())
            T0["factor"][key()] = math.log(c0) - math.log(c)
            T0["loc"][direction] //= 2
    P = math.exp(sum(T0["factor"].values())) * float(np.einsum("xxyyzz", T0["Tensor"])) \
        / float(np.einsum("xxyyzz", T["Tensor"]))
    return T, T0, P


def network_value(tensors, shape):
    """Contraction of the periodic box `shape` (Lx, Ly, Lz) of six-leg tensors
    (x, X, y, Y, z, Z); `tensors` maps site (i, j, k) to its tensor.  Each
    site's X leg bonds to the x leg of its +x neighbour (likewise y, z); a
    length-2 axis therefore carries two bonds between the pair.  For tests
    on boxes of a few sites."""
    import itertools
    sites = list(itertools.product(*[range(L) for L in shape]))
    labels = {}

    def bond(site, d):
        key = (site, d)
        if key not in labels:
            labels[key] = len(labels)
        return labels[key]

    operands, subs = [], []
    for s in sites:
        legs = []
        for d in range(3):
            prev = list(s)
            prev[d] = (prev[d] - 1) % shape[d]
            legs += [bond(tuple(prev), d), bond(s, d)]
        operands.append(tensors[s])
        subs.append(legs)
    args = []
    for t, sb in zip(operands, subs):
        args += [t, sb]
    return float(np.einsum(*args, [], optimize="greedy"))


# ---------------------------------------------------------------------------
# two-point function (trg/HOTRG_3d_QR.py:428-518, 910-1033)
# ---------------------------------------------------------------------------
def sqr_distance(loc1, loc2):
    """Squared distance of two location dicts (trg/HOTRG_3d_QR.py, as in trg/ATRG_3d_new.py:960-965)."""
    return sum((loc1[k] - loc2[k]) ** 2 for k in loc1)


def new_impuer_tensor_2points(T, T0, Tn, loc0, locn, Dcut, direction, squeezer_type="QR", degeneracy_eps=0.0):
    """trg/HOTRG_3d_QR.py:428-475 on arrays: projectors from the pure tensor;
    each impurity is merged with a pure partner, as the upper tensor of the
    pair when its coordinate along `direction` is odd."""
    T = leg_transposition(T, "transpose", direction)
    T0 = leg_transposition(T0, "transpose", direction)
    Tn = leg_transposition(Tn, "transpose", direction)
    Px, PX, Py, PY, _ = squeezer(T, T, T, T, Dcut, squeezer_type, degeneracy_eps)
    T0 = coarse_graining(T, T0, Px, PX, Py, PY) if loc0[direction] % 2 == 1 else coarse_graining(T0, T, Px, PX, Py, PY)
    Tn = coarse_graining(T, Tn, Px, PX, Py, PY) if locn[direction] % 2 == 1 else coarse_graining(Tn, T, Px, PX, Py, PY)
    T = coarse_graining(T, T, Px, PX, Py, PY)
    return (leg_transposition(T, "restore", direction), leg_transposition(T0, "restore", direction),
            leg_transposition(Tn, "restore", direction))


def new_impuer_tensor_absorb(T, T0, Tn, loc0, locn, Dcut, direction, squeezer_type="QR", degeneracy_eps=0.0):
    """trg/HOTRG_3d_QR.py:477-518 on arrays: the two impurities form one pair
    (the even coordinate is the lower tensor) and become one impurity."""
    T = leg_transposition(T, "transpose", direction)
    T0 = leg_transposition(T0, "transpose", direction)
    Tn = leg_transposition(Tn, "transpose", direction)
    Px, PX, Py, PY, _ = squeezer(T, T, T, T, Dcut, squeezer_type, degeneracy_eps)
    if loc0[direction] % 2 == 0 and locn[direction] % 2 == 1:
        T0 = coarse_graining(T0, Tn, Px, PX, Py, PY)
    elif loc0[direction] % 2 == 1 and locn[direction] % 2 == 0:
        T0 = coarse_graining(Tn, T0, Px, PX, Py, PY)
    T = coarse_graining(T, T, Px, PX, Py, PY)
    return leg_transposition(T, "restore", direction), leg_transposition(T0, "restore", direction)


def two_point_function(T, T0, Tn, Dcut, TOT_RGSTEPS, squeezer_type="QR", degeneracy_eps=0.0):
    """trg/HOTRG_3d_QR.py:910-1033 on the dicts {'Tensor', 'loc', 'factor'}:
    while both impurities exist, T0['factor'] holds ln c0 + ln cn - 2 ln c per
    step; once their coordinates coincide after a step they are absorbed
    into one impurity (factor ln c0 - ln c from then on; the one-point step
    when Tn is gone).  Returns (T, T0, G) with G = exp(sum T0 factors)
    Tr T0 / Tr T after the last step."""
    rgstep = {"X": 0, "Y": 0, "Z": 0}
    key = lambda: f"({rgstep['X']},{rgstep['Y']},{rgstep['Z']})"
    T["Tensor"], c = normalization(T["Tensor"])
    T0["Tensor"], c0 = normalization(T0["Tensor"])
    Tn["Tensor"], c1 = normalization(Tn["Tensor"])
    T["factor"][key()] = math.log(c)
    T0["factor"][key()] = math.log(c0) + math.log(c1) - 2 * math.log(c)
    have_n = True
    for direction in cycle("ZXY"):
        if sum(rgstep.values()) >= sum(TOT_RGSTEPS.values()):
            break
        if rgstep[direction] >= TOT_RGSTEPS[direction]:
            continue
        rgstep[direction] += 1
        loc0, n = T0["loc"], sum(rgstep.values())
        if have_n:
            locn = Tn["loc"]
            after0 = dict(loc0, **{direction: loc0[direction] // 2})
            aftern = dict(locn, **{direction: locn[direction] // 2})
            if sqr_distance(after0, aftern) > 0:
                T["Tensor"], T0["Tensor"], Tn["Tensor"] = new_impuer_tensor_2points(
                    T["Tensor"], T0["Tensor"], Tn["Tensor"], loc0, locn, Dcut, direction, squeezer_type,
                    degeneracy_eps)
                T["Tensor"], c = normalization(T["Tensor"])
                T0["Tensor"], c0 = normalization(T0["Tensor"])
                Tn["Tensor"], c1 = normalization(Tn["Tensor"])
                T["factor"][key()] = math.log(c) / 2 ** n
                T0["factor"][key()] = math.log(c0) + math.log(c1) - 2 * math.log(c)
                loc0[direction] //= 2
                locn[direction] //= 2
                continue
            T["Tensor"], T0["Tensor"] = new_impuer_tensor_absorb(
                T["Tensor"], T0["Tensor"], Tn["Tensor"], loc0, locn, Dcut, direction, squeezer_type, degeneracy_eps)
            have_n = False
        else:
            T["Tensor"], T0["Tensor"] = new_impuer_tensor_1point(T["Tensor"], T0["Tensor"], loc0, Dcut, direction,
                                                                 squeezer_type, degeneracy_eps)
        T["Tensor"], c = normalization(T["Tensor"])
        T0["Tensor"], c0 = normalization(T0["Tensor"])
        T["factor"][key()] = math.log(c) / 2 ** n
        T0["factor"][key()] = math.log(c0) - math.log(c)
        loc0[direction] //= 2
    G = math.exp(sum(T0["factor"].values())) * float(np.einsum("xxyyzz", T0["Tensor"])) \
        / float(np.einsum("xxyyzz", T["Tensor"]))
    return T, T0, G


# ---------------------------------------------------------------------------
# diagnostics (trg/HOTRG_3d_QR.py:559-713); T has legs (x, X, y, Y, t, T)
# ---------------------------------------------------------------------------
def density_matrix(T, leg):
    """The partial traces the diagnostics use, in the reference's output order:
    "t" -> "xxyytT->Tt", "x" -> "xXyytt->Xx", "y" -> "xxyYtt->Yy" (:617, :645,
    :673, :695)."""
    spec = {"t": "xxyytT->Tt", "x": "xXyytt->Xx", "y": "xxyYtt->Yy"}[leg]
    return np.einsum(spec, T)


def cal_X(T):
    """:559-583: X_t, X_x, X_y = (Tr T)^2 / TT, TT the trace of two tensors
    stacked along the leg ("xxyytT,aabbTt", "xXyytt,Xxaabb", "xxyYtt,aaYybb")."""
    tr = np.einsum("xxyytt", T)
    TTt = np.einsum("xxyytT,aabbTt", T, T)
    TTx = np.einsum("xXyytt,Xxaabb", T, T)
    TTy = np.einsum("xxyYtt,aaYybb", T, T)
    return tr ** 2 / TTt, tr ** 2 / TTx, tr ** 2 / TTy


def find_1stto3rdmax(T):
    """:585-611: the three diagonal entries T[x,x,y,y,t,t] of largest |value|
    (the reference writes them and returns nothing)."""
    d = np.einsum("xxyytt->xyt", T).ravel()
    return d[np.argsort(-np.abs(d), kind="stable")[:3]]


def cal_scaling_dimensions(T, Dcut):
    """:613-637: eigh of the t-density matrix (its lower triangle, as
    cupy.linalg.eigh reads it), descending, floored at 1e-16, padded to Dcut,
    log(e0 / e) / (2 pi)."""
    e = np.linalg.eigh(density_matrix(T, "t"), UPLO="L")[0][::-1].copy()
    e[e < 1e-16] = 1e-16
    e = np.pad(e, (0, Dcut - len(e)), constant_values=1e-16)
    return np.log(np.max(e) / e) / (2 * np.pi)


def cal_coorelation(T, Dcut, rgstep):
    """:639-713: singular values of the three density matrices, zero-padded
    to Dcut^2 and normalised by their largest, and ln xi = n ln 2 - ln(-1 / ln e_1)
    with n = rgstep["Z"], ["X"], ["Y"]; as in the reference the y branch
    tests the x leg's e_1 against 1e-200 (:700).  Returns ({leg: e}, {leg: ln xi})."""
    ev = {}
    for leg in "txy":
        e = np.linalg.svd(density_matrix(T, leg), compute_uv=False)
        e = np.pad(e, (0, max(0, Dcut * Dcut - len(e))))
        ev[leg] = e / np.max(e)
    lnxi = {}
    for leg, key in (("t", "Z"), ("x", "X"), ("y", "Y")):
        cond = ev["x"][1] if leg == "y" else ev[leg][1]
        e1 = ev[leg][1] if cond > 1e-200 else 1e-200
        lnxi[leg] = rgstep[key] * np.log(2) - np.log(-1 / np.log(e1))
    return ev, lnxi


def tensor_spectrum(T, Dcut):
    """trg/HOTRG_3d.py:240-245 and trg/HOTRG_3d_QR.py:727-731: singular values of
    T as the (x y z) x (x' y' z') matrix; returns (s[:2 Dcut] / max s, max s)."""
    n = T.shape[0] * T.shape[2] * T.shape[4]
    s = np.linalg.svd(T.transpose(0, 2, 4, 1, 3, 5).reshape(n, -1), compute_uv=False)
    return s[:2 * Dcut] / np.max(s), np.max(s)
