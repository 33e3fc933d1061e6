"""ATRG 3D coarse graining on libtrg (trg/ATRG_3d_new.py:18-482, 954-1039).

Tensor: T_{TXYtxy} = U_{TXYi} s_i VH_{itxy}; U, s, VH are GPU tensors.
Contractions run in the configured storage dtype (float64 or float32 on
MFMA); every decomposition is float64.

Device-specific choices (documented; results are the same function of the
tensor up to the randomized sketch and rounding):
  * leg swapping works in the (x1, y1) / (X0, Y0) leg bases directly.  The
    reference's SVDs of B and C (:175-188) only rotate those legs (their
    factors UB, VCH are multiplied back at :232-233), so M = B~ C~ and its
    singular triplets are unchanged; the rank bound chi_k is still computed
    from the reference's reduced sizes.
  * range-finder blocks are orthonormalised with shifted CholeskyQR3 in
    place of Householder QR (same column span).
  * the squeezer's environment square roots R^T R = XdagX, Rx Rx^T = xxdag
    are positive-semidefinite Cholesky factors instead of sqrt(s) vh from
    full SVDs (:271-279); the oblique projectors PX Px depend only on the
    environments, not on which square root is used.
"""
import math

import torch

from . import linalg, ops
from ._lib import TrgError, call
from .contract import contract

degeneracy_eps = 0.0
truncate_eps = 1e-10
# leg-swap rSVD parameters (reference: k_factor 3, p 0, q = Dcut, :193)
RSVD = {"k_factor": 3, "p": 0, "q": None, "seed": 12345}
# called as hook(sM, k, ispure) with the leg-swap singular values of every
# step (k = max(Dcut, chi_k)); the drop-in writes the reference's
# leg_exchange_*/ files through it (trg/ATRG_3d_new.py:197-212)
LEG_SWAP_HOOK = None


def set_rsvd_params(k_factor=3, p=0, q=None, seed=12345):
    RSVD.update(k_factor=k_factor, p=p, q=q, seed=seed)


def _dt(prec):
    return torch.float64 if prec in (None, "f64") else torch.float32


# ---------------------------------------------------------------------------
# leg transposition (:18-65)
# ---------------------------------------------------------------------------
def leg_transposition(U, VH, do_what="transpose", direction="t"):
    if direction in ("t", "T"):
        return U, VH
    fwd = do_what == "transpose"
    if direction in ("x", "X"):
        ua, va = ((1, 2, 0, 3), (0, 2, 3, 1)) if fwd else ((2, 0, 1, 3), (0, 3, 1, 2))
    else:
        ua, va = ((2, 0, 1, 3), (0, 3, 1, 2)) if fwd else ((1, 2, 0, 3), (0, 2, 3, 1))
    return ops.permute(U, ua), ops.permute(VH, va)


# ---------------------------------------------------------------------------
# leg swapping (:67-235)
# ---------------------------------------------------------------------------
def rsvd_for_ATRG_leg_swapping(Bt, Ct, chi_k, p, q, eps, seed, keep=None):
    """Randomized SVD of M_{(i,g),(j,b)} = sum_t Bt_{b,i,t} Ct_{t,j,g}
    (:67-140).  Returns UM (i,g,k), sM (k,), VMH (k,j,b) as float64.
    `keep`: the number of singular triplets the caller keeps at most (the
    sketch may shrink to it, see below)."""
    nb, ni, nt = Bt.shape
    _, nj, ng = Ct.shape
    l = chi_k + p
    rows_i, rows_j = ni * ng, nj * nb
    ls = min(l, rows_i, rows_j)
    dt = Bt.dtype

    # Both applies contract the small bond first, so the intermediate is
    # (i,t,j,l) / (t,j,i,l) -- chi^3 l elements -- and each costs 4 chi^5 l
    # flop, the optimal path of the reference's three-operand einsums (:99, :110, :117).
    def apply_M(Yjbl):            # (j,b,l) -> (i,g,l)
        W = contract("bit,jbl->itjl", Bt, Yjbl)
        return contract("itjl,tjg->igl", W, Ct)

    def apply_MT(Qigl):           # (i,g,l) -> (j,b,l)
        W = contract("tjg,igl->tjil", Ct, Qigl)
        return contract("tjil,bit->jbl", W, Bt)

    # The small SVDs run in one workgroup's LDS, which caps the sketch block
    # at linalg.SMALL_MAX columns (a documented device limit: above it the
    # block is truncated to SMALL_MAX).
    if ls >= min(rows_i, rows_j) and min(rows_i, rows_j) > linalg.SMALL_MAX:
        ls = linalg.SMALL_MAX
    elif ls > linalg.SMALL_MAX:
        ls = linalg.SMALL_MAX
    if ls >= min(rows_i, rows_j):
        # the sketch spans the whole space: dense SVD of M (small early steps)
        M = contract("bit,tjg->igjb", Bt, Ct).reshape(rows_i, rows_j)
        u, s, vh = linalg.svd_thin(M[None])
        UM, sM, VMH = u[0], s[0], vh[0]
    else:
        def sketch(ls, perturb):
            infos = []

            def orth(X):
                Q, info = linalg.orthonormalize(X, perturb=perturb, return_info=True)
                infos.append(info)
                return Q[0]
            Y = ops.gaussian((nj, nb, l), seed=seed, device=Bt.device)[..., :ls]
            Y = ops.contiguous(Y, dt)
            Q = orth(apply_M(Y).reshape(1, rows_i, ls))
            for _ in range(q):
                Z = apply_MT(ops.cast(Q.view(ni, ng, ls), dt))
                Q2 = orth(_f64(Z).reshape(1, rows_j, ls))
                Y = apply_M(ops.cast(Q2.view(nj, nb, ls), dt))
                Q = orth(_f64(Y).reshape(1, rows_i, ls))
            # X^T = M^T Q : (j,b,l), then the thin SVD of X (l x rows_j)
            XT = _f64(apply_MT(ops.cast(Q.view(ni, ng, ls), dt))).reshape(1, rows_j, ls)
            v, s, uh = linalg.svd_thin(XT)       # X^T = v s uh  ->  X = uh^T s v^T
            bad = int(torch.stack(infos).max().item()) != 0
            return ops.gemm(Q, ops.contiguous(uh[0].transpose(0, 1))), s[0], ops.contiguous(v[0].transpose(0, 1)), bad

        # A sketch wider than the rank of M (small systems, or a bond that
        # is rank deficient) breaks the Cholesky of the range finder down at
        # the rounding level.  Narrow the sketch towards the number of
        # triplets kept; only if that still fails, repair the blocks with a
        # perturbation at 1e-10 of their largest entry.
        floor = max(1, min(ls, keep if keep else ls))
        UM, sM, VMH, bad = sketch(ls, linalg.PERTURB)
        while bad and ls > floor:
            ls = max(floor, ls // 2)
            UM, sM, VMH, bad = sketch(ls, linalg.PERTURB)
        if bad:
            UM, sM, VMH, bad = sketch(ls, 1e-10)
            if bad:
                raise TrgError("leg-swap range finder: Cholesky breakdown on a rank-deficient sketch")
    k = min(linalg.count_above(sM, eps), chi_k)
    k = max(k, 1)
    UM = ops.contiguous(UM[:, :k]).view(ni, ng, k)
    VMH = ops.contiguous(VMH[:k]).view(k, nj, nb)
    return UM, ops.contiguous(sM[:k]), VMH


def _f64(X):
    return X if X.dtype == torch.float64 else ops.cast(X, torch.float64)


def leg_swapping(B, C, Dcut, ispure=True, prec="f64"):
    """:142-235.  B (i,t1,x1,y1), C (T0,X0,Y0,j) -> UM (i,X0,Y0,k), sM (k), VMH (k,j,x1,y1)."""
    ni, nt1, nx1, ny1 = B.shape
    nT0, nX0, nY0, nj = C.shape
    beta_ref = min(nx1 * ny1, ni * nt1)
    gamma_ref = min(nT0 * nj, nX0 * nY0)
    chi_k = min(Dcut, nj * beta_ref, ni * gamma_ref)
    Bt = ops.permute(B, (2, 3, 0, 1)).view(nx1 * ny1, ni, nt1)      # (b, i, t)
    Ct = ops.permute(C, (0, 3, 1, 2)).view(nT0, nj, nX0 * nY0)      # (t, j, g)
    q = Dcut if RSVD["q"] is None else RSVD["q"]
    UM, sM, VMH = rsvd_for_ATRG_leg_swapping(Bt, Ct, RSVD["k_factor"] * chi_k, RSVD["p"], q, truncate_eps,
                                             RSVD["seed"], keep=chi_k)
    if LEG_SWAP_HOOK is not None:
        LEG_SWAP_HOOK(sM, max(Dcut, chi_k), ispure)
    chi, _ = linalg.truncate(sM, chi_k, degeneracy_eps, truncate_eps)
    UM = ops.contiguous(UM[:, :, :chi]).view(ni, nX0, nY0, chi)
    VMH = ops.contiguous(VMH[:chi]).view(chi, nj, nx1, ny1)
    return UM, ops.contiguous(sM[:chi]), VMH


# ---------------------------------------------------------------------------
# squeezer (:246-317)
# ---------------------------------------------------------------------------
def squeezer_environments(A, UM, sM, VMH, D):
    """The four environments (:262-265) as float64 matrices."""
    nX1, nY1 = A.shape[1], A.shape[2]
    nX0, nY0 = UM.shape[1], UM.shape[2]
    nx1, ny1 = VMH.shape[2], VMH.shape[3]
    nx0, ny0 = D.shape[2], D.shape[3]
    dt = A.dtype
    sMd = ops.cast(sM, dt)
    UMs = ops.contiguous(UM)
    ops.scale_rows_(UMs, sMd, UM.shape[3], 1, 0)                      # UM * s (last leg)
    VMHs = ops.contiguous(VMH)
    ops.scale_rows_(VMHs, sMd, VMH.shape[0], VMH.numel() // VMH.shape[0], 0)   # s * VMH (first leg)
    P = contract("aibc,akbf->ickf", A, A)
    Q = contract("cjde,flde->cjfl", UM, UMs)
    XdX = contract("ickf,cjfl->ijkl", P, Q).reshape(nX1 * nX0, nX1 * nX0)
    P = contract("abic,abkf->ickf", A, A)
    Q = contract("cdje,fdle->cjfl", UM, UMs)
    YdY = contract("ickf,cjfl->ijkl", P, Q).reshape(nY1 * nY0, nY1 * nY0)
    P = contract("bejf,delf->bjdl", D, D)
    Q = contract("abic,adkc->bidk", VMHs, VMH)
    xxd = contract("bidk,bjdl->ijkl", Q, P).reshape(nx1 * nx0, nx1 * nx0)
    P = contract("befj,defl->bjdl", D, D)
    Q = contract("abci,adck->bidk", VMHs, VMH)
    yyd = contract("bidk,bjdl->ijkl", Q, P).reshape(ny1 * ny0, ny1 * ny0)
    return [_f64(e) for e in (XdX, YdY, xxd, yyd)]


EIG_Q = 12   # power iterations of the top-k SVD of RX Rx when it is too large for the dense path
EIG_L = 2    # its block: EIG_L x the kept rank (at most linalg.SMALL_MAX columns)
# Run the squeezer's independent decompositions (four Cholesky roots, two SVDs)
# as batched launches.  Every kernel is batch-invariant, so each result is the
# same bits either way; False reproduces the reference's one-at-a-time order.
BATCH_DECOMPOSITIONS = True


def _by_shape(fn, mats):
    """fn applied to each matrix of `mats`; with BATCH_DECOMPOSITIONS the
    matrices of equal shape are stacked into one batched call.  fn maps a
    (B, m, n) stack to a tuple of (B, ...) stacks (or one stack)."""
    if not BATCH_DECOMPOSITIONS:
        res = [fn(m[None]) for m in mats]
        return [tuple(x[0] for x in r) if isinstance(r, tuple) else r[0] for r in res]
    groups = {}
    for i, m in enumerate(mats):
        groups.setdefault(tuple(m.shape), []).append(i)
    out = [None] * len(mats)
    for idx in groups.values():
        if len(idx) == 1:
            stack = mats[idx[0]][None]
        else:
            stack = torch.empty((len(idx),) + tuple(mats[idx[0]].shape), dtype=mats[idx[0]].dtype,
                                device=mats[idx[0]].device)
            for j, i in enumerate(idx):
                ops.copy_into(stack[j], mats[i])
        r = fn(stack)
        for j, i in enumerate(idx):
            out[i] = tuple(x[j] for x in r) if isinstance(r, tuple) else r[j]
    return out


def _oblique_pairs(items, Dcut, dt, deg_eps=None, trunc_eps=None):
    """:278-292 for each (Env_big, Env_small, shape_big, shape_small) of `items`, with
    R^T R = Env_big (RX) and Rx Rx^T = Env_small; returns [(Px, PX), ...].  The
    Cholesky roots, the products RX Rx and their SVDs of all items run as
    batched launches (BATCH_DECOMPOSITIONS).  The rank rule uses this module's
    DGENERACY_EPS / truncate_eps unless deg_eps / trunc_eps are given."""
    deg_eps = degeneracy_eps if deg_eps is None else deg_eps
    trunc_eps = truncate_eps if trunc_eps is None else trunc_eps
    roots = _by_shape(linalg.chol_psd_upper, [e for it in items for e in (it[0], it[1])])
    RXs = roots[0::2]                                                  # R^T R = XdagX
    Rxs = [ops.contiguous(r.transpose(0, 1)) for r in roots[1::2]]    # Rx Rx^T = xxdag
    Ps = [ops.gemm(RX, Rx) for RX, Rx in zip(RXs, Rxs)]
    k1s = [min(RX.shape[0], Rx.shape[1], Dcut) for RX, Rx in zip(RXs, Rxs)]
    small = [linalg._fits_small(P.shape[0], P.shape[0]) for P in Ps]
    svds = [None] * len(Ps)
    dense = [i for i in range(len(Ps)) if small[i]]
    for i, r in zip(dense, _by_shape(linalg.svd_small, [Ps[i] for i in dense])):
        svds[i] = r
    topk = {}
    for i in range(len(Ps)):
        if not small[i]:
            n = Ps[i].shape[0]
            topk.setdefault(min(n, max(EIG_L * k1s[i], k1s[i] + 8), linalg.SMALL_MAX), []).append(i)
    for l, idx in topk.items():
        for i, r in zip(idx, _by_shape(lambda A: linalg.svd_topk(A, l, EIG_Q), [Ps[i] for i in idx])):
            svds[i] = r
    out = []
    for i, (Env_big, Env_small, shape_big, shape_small) in enumerate(items):
        RX, Rx, P, k1 = RXs[i], Rxs[i], Ps[i], k1s[i]
        u, s, vh = svds[i]
        n = P.shape[0]
        if small[i]:
            chi, _ = linalg.truncate(s, k1, deg_eps, trunc_eps)
        else:
            chi, _ = linalg.truncate_with_tail(s, P, k1, deg_eps, trunc_eps)
        chi = min(chi, Dcut)
        sis = ops.contiguous(s[:chi])
        v = ops.contiguous(vh[:chi].transpose(0, 1))                  # (n, chi)
        PX = ops.gemm(Rx, v)
        ops.scale_rows_(PX, sis, chi, 1, 2)                            # * s^-1/2 (columns)
        Px = ops.gemm(ops.contiguous(u[:, :chi].transpose(0, 1)), RX)
        ops.scale_rows_(Px, sis, chi, n, 2)                            # s^-1/2 * (rows)
        out.append((ops.cast(Px, dt).view((chi,) + shape_small), ops.cast(PX, dt).view(shape_big + (chi,))))
    return out


def _oblique_pair(Env_big, Env_small, Dcut, shape_big, shape_small, dt, deg_eps=None, trunc_eps=None):
    """One oblique projector pair (:278-292); see _oblique_pairs."""
    return _oblique_pairs([(Env_big, Env_small, shape_big, shape_small)], Dcut, dt, deg_eps, trunc_eps)[0]


def squeezer(A, UM, sM, VMH, D, Dcut):
    """:246-317 -> Px (x,x1,x0), PX (X1,X0,X), Py, PY.

    The environments are Gram matrices whose smallest kept eigenvalues sit
    near 1e-8 of the largest at chi = 64, and the oblique projectors divide
    by their square roots.  They are therefore always formed in float64 (24
    chi^6 flop, small next to the chi^7 coarse graining): a float32 Gram is
    not positive semi-definite at that level, and its negative noise
    eigenvalues, amplified by s^-1/2, make a float32 run diverge."""
    dt = A.dtype
    XdX, YdY, xxd, yyd = squeezer_environments(_f64(A), _f64(UM), sM, _f64(VMH), _f64(D))
    (Px, PX), (Py, PY) = _oblique_pairs([(XdX, xxd, (A.shape[1], UM.shape[1]), (VMH.shape[2], D.shape[2])),
                                         (YdY, yyd, (A.shape[2], UM.shape[2]), (VMH.shape[3], D.shape[3]))],
                                        Dcut, dt)
    return Px, PX, Py, PY


# ---------------------------------------------------------------------------
# coarse graining (:357-438)
# ---------------------------------------------------------------------------
def contract_G(A, PX, UM, PY, compute=None):
    """G_{T,X,Y,k} = A_{T,a,b,i} PX_{a,c,X} UM_{i,c,d,k} PY_{b,d,Y} (:379-384)."""
    W = contract("Tabi,acX->TbicX", A, PX, compute=compute, tag="atrg_cg")
    W = contract("TbicX,icdk->TbXdk", W, UM, compute=compute, tag="atrg_cg")
    return contract("TbXdk,bdY->TXYk", W, PY, compute=compute, tag="atrg_cg")


def contract_H(VMH, Px, D, Py, compute=None):
    """H_{k,t,x,y} = VMH_{k,j,a,b} Px_{x,a,c} D_{j,t,c,d} Py_{y,b,d} (:400-405)."""
    W = contract("kjab,xac->kjbxc", VMH, Px, compute=compute, tag="atrg_cg")
    W = contract("kjbxc,jtcd->kbxtd", W, D, compute=compute, tag="atrg_cg")
    return contract("kbxtd,ybd->ktxy", W, Py, compute=compute, tag="atrg_cg")


def _trunc_triplet(u, s, vh, k):
    chi, _ = linalg.truncate(s, k, degeneracy_eps, truncate_eps)
    return ops.contiguous(u[:, :chi]), ops.contiguous(s[:chi]), ops.contiguous(vh[:chi])


def _svd_trunc(M2, k):
    u, s, vh = linalg.svd_thin(_f64(M2)[None])
    return _trunc_triplet(u[0], s[0], vh[0], k)


def coarse_graining(A, UM, sM, VMH, D, Px, PX, Py, PY, compute=None):
    nT, nt = A.shape[0], D.shape[1]
    nX, nY = PX.shape[2], PY.shape[2]
    nx, ny = Px.shape[0], Py.shape[0]
    nk = UM.shape[3]
    G = contract_G(A, PX, UM, PY, compute)
    H = contract_H(VMH, Px, D, Py, compute)
    Gm, Hm = G.reshape(nT * nX * nY, nk), H.reshape(nk, nt * nx * ny)
    if BATCH_DECOMPOSITIONS and Gm.shape == Hm.shape[::-1] and Gm.shape[0] >= Gm.shape[1]:
        # the thin SVDs of G and of H^T share one batched launch sequence
        # (svd(H) is the transpose of svd(H^T))
        (uG, sGa, vhG), (uHt, sHa, vhHt) = _by_shape(linalg.svd_thin, [_f64(Gm), _f64(Hm).transpose(0, 1)])
        UG, sG, VGH = _trunc_triplet(uG, sGa, vhG, min(nT * nX * nY, nk))
        UH, sH, VHH = _trunc_triplet(vhHt.transpose(0, 1), sHa, uHt.transpose(0, 1), min(nk, nt * nx * ny))
    else:
        UG, sG, VGH = _svd_trunc(Gm, min(nT * nX * nY, nk))
        UH, sH, VHH = _svd_trunc(Hm, min(nk, nt * nx * ny))
    # K = sG VGH sM UH sH
    Kl = ops.contiguous(VGH)
    ops.scale_rows_(Kl, sG, sG.numel(), Kl.shape[1], 0)
    ops.scale_rows_(Kl, ops.contiguous(sM), sM.numel(), 1, 0)
    Kr = ops.contiguous(UH)
    ops.scale_rows_(Kr, sH, sH.numel(), 1, 0)
    K = ops.gemm(Kl, Kr)
    UK, sK, VKH = _svd_trunc(K, min(K.shape))
    dt = A.dtype
    U = ops.cast(ops.gemm(UG, UK), dt).view(nT, nX, nY, -1)
    VH = ops.cast(ops.gemm(VKH, VHH), dt).view(-1, nt, nx, ny)
    return U, sK, VH


def _pair_operands(U, s, VH, direction, dt):
    """Transposed legs and the pair operands of one tensor: A = U, D = VH,
    B = s_i VH_{itxy}, C = U_{TXYj} s_j (trg/ATRG_3d_new.py:446-450)."""
    U, VH = ops.cast(U, dt), ops.cast(VH, dt)
    U, VH = leg_transposition(U, VH, "transpose", direction)
    sd = ops.cast(s, dt)
    B = ops.contiguous(VH)
    ops.scale_rows_(B, sd, sd.numel(), VH.numel() // VH.shape[0], 0)    # s_i VH_{itxy}
    C = ops.contiguous(U)
    ops.scale_rows_(C, sd, sd.numel(), 1, 0)                             # U_{TXYj} s_j
    return U, VH, B, C


def atrg_pure_tensor(U, s, VH, Dcut, direction, prec="f64"):
    """:441-482 -> (U, s, VH) of the coarse tensor (legs restored)."""
    (U2, s2, VH2), _ = atrg_1point_step(U, s, VH, [], Dcut, direction, prec)
    return U2, s2, VH2


def atrg_1point_step(U, s, VH, imps, Dcut, direction, prec="f64"):
    """One step of the pure tensor and of each impurity
    (trg/ATRG_3d_new.py:669-734; with no impurity it is the pure step
    :441-482).  `imps` holds (U0, s0, VH0, odd): odd is True when the
    impurity's coordinate along `direction` is odd, and the impurity is
    then the upper tensor (A, B) of its pair, else the lower one (C, D).
    Returns ((U, s, VH), [(U0, s0, VH0), ...]) with the legs restored."""
    pairs = [((U0, s0, VH0), None) if odd else (None, (U0, s0, VH0)) for U0, s0, VH0, odd in imps]
    return atrg_pairs_step(U, s, VH, pairs, Dcut, direction, prec)


def atrg_pairs_step(U, s, VH, pairs, Dcut, direction, prec="f64"):
    """One step of the pure tensor and of impurity pairs.  Each pair is
    (upper, lower), each a (U, s, VH) or None for the pure tensor: the upper
    tensor gives A = U and B = s VH, the lower one C = U s and D = VH.  The
    squeezers come from the pure pair; each impurity pair gets its own leg
    swap and the pure pair's coarse graining (trg/ATRG_3d_new.py:485-734:
    a one-point or two-point step pairs each impurity with a pure tensor,
    the absorb step pairs the two impurities).  Returns
    ((U, s, VH), [(U, s, VH) per pair]) with the legs restored."""
    dt = _dt(prec)
    pure = _pair_operands(U, s, VH, direction, dt)
    A, D, B, C = pure
    UM, sM, VMH = leg_swapping(B, C, Dcut, True, prec)
    UM, VMH = ops.cast(UM, dt), ops.cast(VMH, dt)
    Px, PX, Py, PY = squeezer(A, UM, sM, VMH, D, Dcut)
    out = []
    for up, lo in pairs:
        Au, _, Bu, _ = pure if up is None else _pair_operands(*up, direction, dt)
        _, Dl, _, Cl = pure if lo is None else _pair_operands(*lo, direction, dt)
        UMi, sMi, VMHi = leg_swapping(Bu, Cl, Dcut, False, prec)
        UMi, VMHi = ops.cast(UMi, dt), ops.cast(VMHi, dt)
        Ui, si, VHi = coarse_graining(Au, UMi, sMi, VMHi, Dl, Px, PX, Py, PY)
        Ui, VHi = leg_transposition(Ui, VHi, "restore", direction)
        out.append((Ui, si, VHi))
    U2, s2, VH2 = coarse_graining(A, UM, sM, VMH, D, Px, PX, Py, PY)
    U2, VH2 = leg_transposition(U2, VH2, "restore", direction)
    return (U2, s2, VH2), out


# ---------------------------------------------------------------------------
# batched pure steps: several instances of equal bond shapes per launch
# ---------------------------------------------------------------------------
SPLITS = 0          # batched steps re-run per group of equal rank decisions (a counter)


class RankSplit(Exception):
    """Raised inside a batched step when the instances' rank decisions (or the
    range finder's breakdown flags) disagree; `keys` holds one value per
    instance, and the caller re-runs the step for each group of equal keys."""

    def __init__(self, keys):
        super().__init__(f"rank decisions differ across the batch: {keys}")
        self.keys = keys


def _uniform(keys):
    keys = list(keys)
    if len(set(keys)) > 1:
        raise RankSplit(keys)
    return keys[0]


def _truncate_b(s, k, deg_eps, trunc_eps):
    """linalg.truncate for every row of s (B, n): one launch of the rank rule
    per instance, one host read for all of them."""
    Bn, n = s.shape
    if deg_eps == 0 and trunc_eps == 0:
        return [int(k)] * Bn
    if k > n:
        return [n] * Bn
    sc = ops.contiguous(_f64(s))
    chi = torch.empty(Bn, dtype=torch.int32, device=s.device)
    err = torch.empty(Bn, dtype=torch.float64, device=s.device)
    for b in range(Bn):
        call("trg_truncate_rank", n, ops.ptr(sc[b]), int(k), float(deg_eps), float(trunc_eps), ops.ptr(chi[b:b + 1]),
             ops.ptr(err[b:b + 1]), ops.stream_ptr())
    return [int(v) for v in chi.cpu().tolist()]


def _count_above_b(s, eps):
    """linalg.count_above for every row of s (B, n), one host read."""
    Bn, n = s.shape
    flags = ops.contiguous(_f64(s))
    c = ops.maxabs(flags, batch=Bn)
    ops.normalize_(flags, c, batch=Bn)
    vals = flags.cpu()
    return [int((vals[b] > eps).sum().item()) for b in range(Bn)]


def _stack(ts):
    """Stack equal-shape device tensors into one (B, ...) buffer with libtrg copies."""
    out = torch.empty((len(ts),) + tuple(ts[0].shape), dtype=ts[0].dtype, device=ts[0].device)
    for b, t in enumerate(ts):
        ops.copy_into(out[b], t)
    return out


def _leg_transposition_b(U, VH, do_what, direction):
    if direction in ("t", "T"):
        return U, VH
    fwd = do_what == "transpose"
    if direction in ("x", "X"):
        ua, va = ((1, 2, 0, 3), (0, 2, 3, 1)) if fwd else ((2, 0, 1, 3), (0, 3, 1, 2))
    else:
        ua, va = ((2, 0, 1, 3), (0, 3, 1, 2)) if fwd else ((1, 2, 0, 3), (0, 2, 3, 1))
    sh = lambda p: (0,) + tuple(i + 1 for i in p)
    return ops.permute(U, sh(ua)), ops.permute(VH, sh(va))


def _rsvd_leg_swap_b(Bt, Ct, chi_k, p, q, eps, seed, keep):
    """rsvd_for_ATRG_leg_swapping for a batch: Bt (B, b, i, t), Ct (B, t, j, g)."""
    Bn, nb, ni, nt = Bt.shape
    _, _, nj, ng = Ct.shape
    l = chi_k + p
    rows_i, rows_j = ni * ng, nj * nb
    ls = min(l, rows_i, rows_j)
    dt = Bt.dtype

    def apply_M(Y):
        W = contract("Zbit,Zjbl->Zitjl", Bt, Y)
        return contract("Zitjl,Ztjg->Zigl", W, Ct)

    def apply_MT(Q):
        W = contract("Ztjg,Zigl->Ztjil", Ct, Q)
        return contract("Ztjil,Zbit->Zjbl", W, Bt)

    if ls >= min(rows_i, rows_j) and min(rows_i, rows_j) > linalg.SMALL_MAX:
        ls = linalg.SMALL_MAX
    elif ls > linalg.SMALL_MAX:
        ls = linalg.SMALL_MAX
    if ls >= min(rows_i, rows_j):
        M = contract("Zbit,Ztjg->Zigjb", Bt, Ct).reshape(Bn, rows_i, rows_j)
        UM, sM, VMH = linalg.svd_thin(M)
    else:
        def sketch(ls, perturb):
            infos = []

            def orth(X):
                Q, info = linalg.orthonormalize(X, perturb=perturb, return_info=True)
                infos.append(info)
                return Q
            Y = ops.gaussian((nj, nb, l), seed=seed, device=Bt.device, batch=Bn, stream_stride=0)
            Y = ops.contiguous(Y.view(Bn, nj, nb, l)[..., :ls], dt)
            Q = orth(apply_M(Y).reshape(Bn, rows_i, ls))
            for _ in range(q):
                Z = apply_MT(ops.cast(Q.view(Bn, ni, ng, ls), dt))
                Q2 = orth(_f64(Z).reshape(Bn, rows_j, ls))
                Y = apply_M(ops.cast(Q2.view(Bn, nj, nb, ls), dt))
                Q = orth(_f64(Y).reshape(Bn, rows_i, ls))
            XT = _f64(apply_MT(ops.cast(Q.view(Bn, ni, ng, ls), dt))).reshape(Bn, rows_j, ls)
            v, s, uh = linalg.svd_thin(XT)
            bad = [int(x) != 0 for x in torch.stack(infos).max(dim=0).values.cpu().tolist()]
            return ops.gemm(Q, ops.contiguous(uh.transpose(1, 2))), s, ops.contiguous(v.transpose(1, 2)), bad

        floor = max(1, min(ls, keep if keep else ls))
        UM, sM, VMH, bad = sketch(ls, linalg.PERTURB)
        while _uniform(bad) and ls > floor:
            ls = max(floor, ls // 2)
            UM, sM, VMH, bad = sketch(ls, linalg.PERTURB)
        if _uniform(bad):
            UM, sM, VMH, bad = sketch(ls, 1e-10)
            if any(bad):
                raise TrgError("leg-swap range finder: Cholesky breakdown on a rank-deficient sketch")
    k = _uniform(max(min(c, chi_k), 1) for c in _count_above_b(sM, eps))
    UM = ops.contiguous(UM[:, :, :k]).view(Bn, ni, ng, k)
    VMH = ops.contiguous(VMH[:, :k]).view(Bn, k, nj, nb)
    return UM, ops.contiguous(sM[:, :k]), VMH


def _leg_swapping_b(B, C, Dcut):
    """leg_swapping for a batch: B (B, i, t1, x1, y1), C (B, T0, X0, Y0, j)."""
    Bn, ni, nt1, nx1, ny1 = B.shape
    _, nT0, nX0, nY0, nj = C.shape
    beta_ref = min(nx1 * ny1, ni * nt1)
    gamma_ref = min(nT0 * nj, nX0 * nY0)
    chi_k = min(Dcut, nj * beta_ref, ni * gamma_ref)
    Bt = ops.permute(B, (0, 3, 4, 1, 2)).view(Bn, nx1 * ny1, ni, nt1)
    Ct = ops.permute(C, (0, 1, 4, 2, 3)).view(Bn, nT0, nj, nX0 * nY0)
    q = Dcut if RSVD["q"] is None else RSVD["q"]
    UM, sM, VMH = _rsvd_leg_swap_b(Bt, Ct, RSVD["k_factor"] * chi_k, RSVD["p"], q, truncate_eps, RSVD["seed"],
                                   keep=chi_k)
    chi = _uniform(_truncate_b(sM, chi_k, degeneracy_eps, truncate_eps))
    UM = ops.contiguous(UM[:, :, :, :chi]).view(Bn, ni, nX0, nY0, chi)
    VMH = ops.contiguous(VMH[:, :chi]).view(Bn, chi, nj, nx1, ny1)
    return UM, ops.contiguous(sM[:, :chi]), VMH


def _squeezer_b(A, UM, sM, VMH, D, Dcut):
    """squeezer for a batch: the environments in batched contractions, then
    the oblique pairs of every instance through _oblique_pairs (its Cholesky
    roots and SVDs run as launches over all instances)."""
    dt = A.dtype
    A, UM, VMH, D = _f64(A), _f64(UM), _f64(VMH), _f64(D)
    Bn = A.shape[0]
    nX1, nY1 = A.shape[2], A.shape[3]
    nX0, nY0 = UM.shape[2], UM.shape[3]
    nx1, ny1 = VMH.shape[3], VMH.shape[4]
    nx0, ny0 = D.shape[3], D.shape[4]
    k = sM.shape[1]
    sMd = ops.cast(sM, A.dtype)
    UMs = ops.contiguous(UM)
    ops.scale_rows_(UMs, sMd, k, 1, 0, batch=Bn)
    VMHs = ops.contiguous(VMH)
    ops.scale_rows_(VMHs, sMd, k, VMH[0].numel() // k, 0, batch=Bn)
    P = contract("Zaibc,Zakbf->Zickf", A, A)
    Q = contract("Zcjde,Zflde->Zcjfl", UM, UMs)
    XdX = contract("Zickf,Zcjfl->Zijkl", P, Q).reshape(Bn, nX1 * nX0, nX1 * nX0)
    P = contract("Zabic,Zabkf->Zickf", A, A)
    Q = contract("Zcdje,Zfdle->Zcjfl", UM, UMs)
    YdY = contract("Zickf,Zcjfl->Zijkl", P, Q).reshape(Bn, nY1 * nY0, nY1 * nY0)
    P = contract("Zbejf,Zdelf->Zbjdl", D, D)
    Q = contract("Zabic,Zadkc->Zbidk", VMHs, VMH)
    xxd = contract("Zbidk,Zbjdl->Zijkl", Q, P).reshape(Bn, nx1 * nx0, nx1 * nx0)
    P = contract("Zbefj,Zdefl->Zbjdl", D, D)
    Q = contract("Zabci,Zadck->Zbidk", VMHs, VMH)
    yyd = contract("Zbidk,Zbjdl->Zijkl", Q, P).reshape(Bn, ny1 * ny0, ny1 * ny0)
    items = []
    for b in range(Bn):
        items.append((XdX[b], xxd[b], (nX1, nX0), (nx1, nx0)))
        items.append((YdY[b], yyd[b], (nY1, nY0), (ny1, ny0)))
    pairs = _oblique_pairs(items, Dcut, dt)
    _uniform((pairs[2 * b][0].shape[0], pairs[2 * b + 1][0].shape[0]) for b in range(Bn))
    Px = _stack([pairs[2 * b][0] for b in range(Bn)])
    PX = _stack([pairs[2 * b][1] for b in range(Bn)])
    Py = _stack([pairs[2 * b + 1][0] for b in range(Bn)])
    PY = _stack([pairs[2 * b + 1][1] for b in range(Bn)])
    return Px, PX, Py, PY


def _coarse_graining_b(A, UM, sM, VMH, D, Px, PX, Py, PY):
    """coarse_graining for a batch."""
    Bn = A.shape[0]
    nT, nt = A.shape[1], D.shape[2]
    nX, nY = PX.shape[3], PY.shape[3]
    nx, ny = Px.shape[1], Py.shape[1]
    nk = UM.shape[4]
    W = contract("ZTabi,ZacX->ZTbicX", A, PX, tag="atrg_cg")
    W = contract("ZTbicX,Zicdk->ZTbXdk", W, UM, tag="atrg_cg")
    G = contract("ZTbXdk,ZbdY->ZTXYk", W, PY, tag="atrg_cg")
    W = contract("Zkjab,Zxac->Zkjbxc", VMH, Px, tag="atrg_cg")
    W = contract("Zkjbxc,Zjtcd->Zkbxtd", W, D, tag="atrg_cg")
    H = contract("Zkbxtd,Zybd->Zktxy", W, Py, tag="atrg_cg")
    del W
    Gm, Hm = G.reshape(Bn, nT * nX * nY, nk), H.reshape(Bn, nk, nt * nx * ny)
    rG, rH = min(nT * nX * nY, nk), min(nk, nt * nx * ny)
    if BATCH_DECOMPOSITIONS and Gm.shape[1:] == Hm.shape[1:][::-1] and Gm.shape[1] >= Gm.shape[2]:
        res = _by_shape(linalg.svd_thin, [_f64(Gm[b]) for b in range(Bn)] +
                        [_f64(Hm[b]).transpose(0, 1) for b in range(Bn)])
        uG, sGa, vhG = (_stack([r[i] for r in res[:Bn]]) for i in range(3))
        uHt, sHa, vhHt = (_stack([r[i] for r in res[Bn:]]) for i in range(3))
        uH, vhH = vhHt.transpose(1, 2), uHt.transpose(1, 2)
    else:
        uG, sGa, vhG = linalg.svd_thin(_f64(Gm))
        uH, sHa, vhH = linalg.svd_thin(_f64(Hm))
    cG = _uniform(_truncate_b(sGa, rG, degeneracy_eps, truncate_eps))
    cH = _uniform(_truncate_b(sHa, rH, degeneracy_eps, truncate_eps))
    UG, sG, VGH = ops.contiguous(uG[:, :, :cG]), ops.contiguous(sGa[:, :cG]), ops.contiguous(vhG[:, :cG])
    UH, sH, VHH = ops.contiguous(uH[:, :, :cH]), ops.contiguous(sHa[:, :cH]), ops.contiguous(vhH[:, :cH])
    # K = sG VGH sM UH sH
    Kl = ops.contiguous(VGH)
    ops.scale_rows_(Kl, sG, cG, Kl.shape[2], 0, batch=Bn)
    ops.scale_rows_(Kl, ops.contiguous(sM), sM.shape[1], 1, 0, batch=Bn)
    Kr = ops.contiguous(UH)
    ops.scale_rows_(Kr, sH, cH, 1, 0, batch=Bn)
    K = ops.gemm(Kl, Kr)
    uK, sKa, vhK = linalg.svd_thin(_f64(K))
    cK = _uniform(_truncate_b(sKa, min(K.shape[1:]), degeneracy_eps, truncate_eps))
    UK, sK, VKH = ops.contiguous(uK[:, :, :cK]), ops.contiguous(sKa[:, :cK]), ops.contiguous(vhK[:, :cK])
    dt = A.dtype
    U = ops.cast(ops.gemm(UG, UK), dt).view(Bn, nT, nX, nY, -1)
    VH = ops.cast(ops.gemm(VKH, VHH), dt).view(Bn, -1, nt, nx, ny)
    return U, sK, VH


def _pure_step_b(U, s, VH, Dcut, direction, prec):
    dt = _dt(prec)
    Bn = U.shape[0]
    U, VH = ops.cast(U, dt), ops.cast(VH, dt)
    U, VH = _leg_transposition_b(U, VH, "transpose", direction)
    sd = ops.cast(s, dt)
    ni = sd.shape[1]
    Bm = ops.contiguous(VH)
    ops.scale_rows_(Bm, sd, ni, VH[0].numel() // ni, 0, batch=Bn)      # s_i VH_{itxy}
    C = ops.contiguous(U)
    ops.scale_rows_(C, sd, ni, 1, 0, batch=Bn)                          # U_{TXYj} s_j
    UM, sM, VMH = _leg_swapping_b(Bm, C, Dcut)
    UM, VMH = ops.cast(UM, dt), ops.cast(VMH, dt)
    Px, PX, Py, PY = _squeezer_b(U, UM, sM, VMH, VH, Dcut)
    U2, s2, VH2 = _coarse_graining_b(U, UM, sM, VMH, VH, Px, PX, Py, PY)
    return _leg_transposition_b(U2, VH2, "restore", direction) + (s2,)


def atrg_pure_tensor_batch(U, s, VH, Dcut, direction, prec="f64"):
    """:441-482 for a batch of instances of equal bond shapes: U (B,T,X,Y,i),
    s (B,i), VH (B,i,t,x,y).  Returns [(U, s, VH)] per instance.  Every
    kernel is batch-invariant, so each instance's result is the one it gets
    alone; where the instances' ranks part (a truncation, the leg-swap eps
    cut or a range-finder breakdown), the step is re-run for each group of
    equal decisions."""
    Bn = U.shape[0]
    try:
        U2, VH2, s2 = _pure_step_b(U, s, VH, Dcut, direction, prec)
        return [(U2[b], s2[b], VH2[b]) for b in range(Bn)]
    except RankSplit as e:
        global SPLITS
        SPLITS += 1
        groups = {}
        for b, key in enumerate(e.keys):
            groups.setdefault(key, []).append(b)
        out = [None] * Bn
        for idx in groups.values():
            sub = atrg_pure_tensor_batch(_stack([U[b] for b in idx]), _stack([s[b] for b in idx]),
                                         _stack([VH[b] for b in idx]), Dcut, direction, prec)
            for j, b in enumerate(idx):
                out[b] = sub[j]
        return out


def normalize_batch(U, s, VH):
    """normalize() for a batch (B, ...): returns (s, c) with c (B,) float64."""
    Bn = U.shape[0]
    d = contract("Ztxyi,Zitxy->Zi", _f64(U), _f64(VH))
    c = ops.gemm(d.view(Bn, 1, -1), ops.contiguous(_f64(s)).view(Bn, -1, 1)).view(Bn)
    s = ops.contiguous(_f64(s))
    ops.normalize_(s, c, batch=Bn)
    return s, c


def trace(U, s, VH):
    """Tr T = sum_{txy,i} U_{txyi} s_i VH_{itxy} (tensor_class/tensor_class.py:66-86), float64 (1,)."""
    d = contract("txyi,itxy->i", _f64(U), _f64(VH))
    return ops.gemm(d.view(1, -1), _f64(s).view(-1, 1)).view(1)


def normalize(U, s, VH):
    """:954-958: c = |Tr T|; s /= c.  Returns (s, c) with c a float64 device scalar (1,)."""
    c = trace(U, s, VH)
    s = ops.contiguous(_f64(s))
    ops.normalize_(s.view(1, -1), c)
    return s, c
