"""HOTRG 3D with squeezers (trg/HOTRG_3d_QR.py) on the device, batched over instances.

A step (trg/HOTRG_3d_QR.py:408-425) transposes the legs for the direction,
builds projectors for the doubled x and y legs with one of three squeezers,
and contracts T0 T1 Px PX Py PY (:318-391) with the batched HOTRG 3D merge
(`hotrg.merge3d`, five GEMMs chunked over the output leg).

Squeezers (:74-281), all fed by the float64 environments of `hotrg.env3d`
(leg 1 = LXdagLX, 0 = RxRxdag, 2 = LydagLy, 3 = RYRYdag of :212-215):

* "original": isometries from the top eigenvectors of the x-leg and y-leg
  environments, with the truncation rule of utility/truncated_svd.py
  (the reference's side comparison sums empty slices, so these sides always
  win, :164-189);
* "fix": the X-leg and y-leg environments (:74-118);
* "QR": oblique projectors Px = S^-1/2 U^T RL, PX = RR V S^-1/2 from square
  roots RL^T RL, RR RR^T of the two environments at each cut and the
  truncated SVD of RL RR (:195-281).  The roots are PSD Cholesky factors;
  any square root gives the same projectors (the reference takes them from
  full SVDs).

The kept rank can differ between instances (the truncation rule reads each
spectrum), so a batch whose ranks disagree splits into groups that continue
separately; within a group every instance's arithmetic is the same as alone.
"""
import math

import torch

from . import atrg, hotrg, linalg, ops
from ._lib import TrgError

TRUNCATE_EPS = 1e-10                        # trg/HOTRG_3d_QR.py:27
SQUEEZERS = ("original", "QR", "fix")
CYCLE = "ZXY"                               # trg/HOTRG_3d_QR.py:780


def _eig_desc(E, k):
    """Top eigenpairs of the PSD float64 environments E (B, n, n), descending:
    (e (B, kk), U (B, n, kk), tail_sq (B,), total_sq (B,)) with kk > k where
    n > k (the degeneracy rule at position k needs the next value).  Exact
    (full Jacobi) for n <= 256, block subspace iteration above."""
    B, n, _ = E.shape
    if n <= 256:
        w, V = linalg.eigh_small(E)
        return ops.flip(w, 1), ops.flip(V, 2), None, None
    kk = min(n, max(2 * k, k + 8), 256)
    if kk <= k:
        raise TrgError(f"bond dimension {k} above the eigensolver's block limit")
    w, V, eps, _ = ops.eigh_topk(ops.contiguous(E), kk, kk, 12, batch=B)
    return ops.flip(w, 1), ops.flip(V, 2), eps, None


def _isometries(E, Dcut, deg_eps):
    """utility eigh + truncate (:146-152) per instance: returns the
    descending eigenvector blocks U (B, n, kk) and the kept rank of each
    instance (a list of ints)."""
    B, n, _ = E.shape
    k = min(n, Dcut)
    e, U, eps, _ = _eig_desc(E, k)
    e = ops.maxabs(e.reshape(-1, 1), batch=e.numel()).view(B, -1)          # |e|
    chis = []
    for b in range(B):
        if eps is None:
            chi, _ = linalg.truncate(e[b], k, deg_eps, TRUNCATE_EPS)
        else:
            tot = linalg.frobenius_sq(E[b])
            tail = ops.contiguous(tot)
            one = ops.fill_(torch.empty((1, 1), dtype=torch.float64, device=E.device), 1.0)
            ops.gemm(one, linalg.sum_sq(e[b]), tail, alpha=-1.0, beta=1.0)
            chi, _ = linalg.truncate_known(e[b], tail, tot, k, deg_eps, TRUNCATE_EPS)
        chis.append(min(chi, k))
    return U, chis


def _take(T, idx):
    """Instances idx of the batched T as a new contiguous batch (libtrg copies)."""
    if len(idx) == T.shape[0]:
        return T
    out = torch.empty((len(idx),) + tuple(T.shape[1:]), dtype=T.dtype, device=T.device)
    for j, b in enumerate(idx):
        ops.copy_into(out[j], T[b])
    return out


def squeeze(T, Dcut, kind, deg_eps=0.0, prec=hotrg.FP64):
    """Projectors of one step for the batch T (B, x, X, y, Y, z, Z) (legs
    already transposed for the direction).  Returns a list of groups
    (idx, U1a, U1b, U2a, U2b) in the ijklez,mnopzf,ima,jnb,koc,lpd form of
    `hotrg.merge3d` (ima = Px^(1,2,0), jnb = PX, koc = Py, lpd = PY^(1,2,0)),
    one group per distinct pair of kept ranks."""
    if kind not in SQUEEZERS:
        raise TrgError(f"unknown squeezer {kind!r} (choose from {SQUEEZERS})")
    B, Dx, DX, Dy, DY = T.shape[:5]
    c = prec.env_compute()
    dt = T.dtype
    env = lambda leg: hotrg.env3d(T, leg, c)
    if kind in ("original", "fix"):
        Ux, cx = _isometries(env(0 if kind == "original" else 1), Dcut, deg_eps)
        Uy, cy = _isometries(env(2), Dcut, deg_eps)
        groups = {}
        for b in range(B):
            groups.setdefault((cx[b], cy[b]), []).append(b)
        out = []
        for (kx, ky), idx in groups.items():
            U1 = ops.cast(ops.contiguous(_take(Ux, idx)[:, :, :kx]), dt).view(len(idx), Dx, DX, kx)
            U2 = ops.cast(ops.contiguous(_take(Uy, idx)[:, :, :ky]), dt).view(len(idx), Dy, DY, ky)
            out.append((idx, U1, U1, U2, U2))
        return out
    LX, Rx, Ly, RY = env(1), env(0), env(2), env(3)
    per = []
    for b in range(B):
        Px, PX = atrg._oblique_pair(LX[b], Rx[b], Dcut, (DX, DX), (Dx, Dx), dt, deg_eps, TRUNCATE_EPS)
        PY, Py = atrg._oblique_pair(Ly[b], RY[b], Dcut, (Dy, Dy), (DY, DY), dt, deg_eps, TRUNCATE_EPS)
        per.append((Px, PX, Py, PY))
    groups = {}
    for b, (Px, _, Py, _) in enumerate(per):
        groups.setdefault((Px.shape[0], Py.shape[2]), []).append(b)
    out = []
    for (kx, ky), idx in groups.items():
        nb = len(idx)
        U1a = torch.empty((nb, Dx, Dx, kx), dtype=dt, device=T.device)
        U1b = torch.empty((nb, DX, DX, kx), dtype=dt, device=T.device)
        U2a = torch.empty((nb, Dy, Dy, ky), dtype=dt, device=T.device)
        U2b = torch.empty((nb, DY, DY, ky), dtype=dt, device=T.device)
        for j, b in enumerate(idx):
            Px, PX, Py, PY = per[b]
            ops.copy_into(U1a[j], Px.permute(1, 2, 0))
            ops.copy_into(U1b[j], PX)
            ops.copy_into(U2a[j], Py)
            ops.copy_into(U2b[j], PY.permute(1, 2, 0))
        out.append((idx, U1a, U1b, U2a, U2b))
    return out


def step(T, Dcut, direction, kind, deg_eps=0.0, prec=hotrg.FP64):
    """One RG step (trg/HOTRG_3d_QR.py:408-425) on a batch; returns a list of
    (idx, T_new) groups (one group unless the kept ranks differ)."""
    if T.dtype != prec.store:
        T = ops.cast(T, prec.store)
    T = hotrg.rearrange3d(T, direction)
    out = []
    for idx, U1a, U1b, U2a, U2b in squeeze(T, Dcut, kind, deg_eps, prec):
        Tg = _take(T, idx)
        Tn = hotrg.merge3d(Tg, Tg, U1a, U1b, U2a, U2b, compute=prec.merge_compute())
        out.append((idx, hotrg.restore3d(Tn, direction)))
    return out


def schedule(TOT_RGSTEPS):
    """Directions in the order of trg/HOTRG_3d_QR.py:780-786 ("ZXY" cycle,
    each direction until its count is reached)."""
    cnt = {"X": 0, "Y": 0, "Z": 0}
    total = sum(TOT_RGSTEPS.get(d, 0) for d in "XYZ")
    order = []
    i = 0
    while sum(cnt.values()) < total:
        d = CYCLE[i % 3]
        i += 1
        if cnt[d] < TOT_RGSTEPS.get(d, 0):
            cnt[d] += 1
            order.append(d)
    return order


def renorm(T, Dcut, TOT_RGSTEPS, kind="QR", deg_eps=0.0, policy=hotrg.ALL_FP64, sync_each_step=False,
           on_step=None):
    """trg/HOTRG_3d_QR.py:747-795 for a batch T (B, x, X, y, Y, z, Z).

    Returns a list of groups (idx, T, lnf) covering every instance: lnf[:, n]
    = ln c_n / 2^n (c_n = |Tr T| after step n, n = 0 before the first step),
    the values the reference stores under the keys "(X,Y,Z)".  `on_step(n,
    direction, groups)` (optional) sees the normalised groups after the
    initial normalisation (n = 0, direction None) and after every step."""
    order = schedule(TOT_RGSTEPS)
    B = T.shape[0]
    health = ops.Health(T.device)
    with health:
        T = ops.contiguous(T, policy(0).store)
        lnf = torch.empty((B, len(order) + 1), dtype=torch.float64, device=T.device)
        ops.fill_(lnf, 0.0)
        hotrg.normalize_trace(T, lnf, 0)
        groups = [(list(range(B)), T, lnf)]
        if on_step is not None:
            on_step(0, None, groups)
        for n, direction in enumerate(order, start=1):
            nxt = []
            for idx, Tg, lg in groups:
                for sub, Tn in step(Tg, Dcut, direction, kind, deg_eps, policy(n)):
                    ln = _take(lg, sub)
                    Tn = ops.contiguous(Tn)
                    hotrg.normalize_trace(Tn, ln, n)
                    nxt.append(([idx[i] for i in sub], Tn, ln))
            groups = nxt
            if on_step is not None:
                on_step(n, direction, groups)
            if sync_each_step:
                torch.cuda.current_stream().synchronize()
    health.check("HOTRG_3d_QR solve")
    return groups


def factor_keys(TOT_RGSTEPS):
    """The reference's "(X,Y,Z)" keys in the order the factors are stored."""
    cnt = {"X": 0, "Y": 0, "Z": 0}
    keys = ["(0,0,0)"]
    for d in schedule(TOT_RGSTEPS):
        cnt[d] += 1
        keys.append(f"({cnt['X']},{cnt['Y']},{cnt['Z']})")
    return keys


def ln_z_over_v(T, Dcut, TOT_RGSTEPS, kind="QR", deg_eps=0.0, policy=hotrg.ALL_FP64):
    """ln Z / V per instance (measurement/Ising_model_3d.py:101-137: sum of the
    factors + ln Tr T / 2^N), as a float64 device vector (B,)."""
    nsteps = sum(TOT_RGSTEPS.get(d, 0) for d in "XYZ")
    B = T.shape[0]
    res = torch.empty(B, dtype=torch.float64, device=T.device)
    for idx, Tg, lg in renorm(T, Dcut, TOT_RGSTEPS, kind, deg_eps, policy):
        z = hotrg.lnz3d(Tg, lg, nsteps)
        for j, b in enumerate(idx):
            ops.copy_into(res[b:b + 1], z[j:j + 1])
    return res
