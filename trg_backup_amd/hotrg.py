"""Batched HOTRG coarse-graining steps on libtrg (2D and 3D).

A batch is a set of instances (couplings beta, same bond shapes) stacked on a
leading axis; every kernel runs once for the whole batch, and every
instance's arithmetic is identical to what it would be alone (the kernels
never mix instances and their reduction orders do not depend on the batch).

Leg orders follow the reference: 2D T[x, x', y, y'] (trg/HOTRG_2d.py), 3D
T[x, x', y, y', z, z'] (trg/HOTRG_3d.py).  A 2D "Y" step merges two tensors
along y and truncates the doubled x legs; 3D "Z" truncates x and y.

Each contraction is a chain of libtrg GEMMs whose operands are strided
views; the only data movement is one permuted copy per operand where no
stride pattern exists (these copies are HBM-bound and small next to the
chi^7 / chi^11 GEMMs).
"""
import math
from dataclasses import dataclass

import torch

from . import ops
from ._lib import TrgError
from .ops import gemm, permute

# subspace-iteration settings of the isometry eigensolver.  2D (chi^2 x chi^2
# environments, block l = 2 chi): 12 iterations reproduce LAPACK eigh to the
# last bits at chi = 64.  The 3D environments' spectra fall off more slowly
# past the cut: at chi = 16 (n = 256, l = 32) 12 iterations leave ln Z/V
# 1.5e-8 from the exact-eigh value after 24 steps, 30 reproduce it exactly.
EIG = {"l_factor": 2, "q": 12, "seed": 20240}
EIG3 = {"l_factor": 2, "q": 30, "seed": 20240}
# memory budgets (bytes) for the chi^5 / chi^8 merge intermediates: per
# instance (sets the chunk length, a function of the shape only) and per batch
INSTANCE_BUDGET = 1 << 30
MERGE_BUDGET = 32 << 30


SPLIT_COMPUTE = ("bf16", "bf16x3", "bf16x6")   # GEMM compute types on the bf16 MFMA


@dataclass(frozen=True)
class Precision:
    """Storage dtype of T and the GEMM compute type of the environment and
    the merge ("f64", "f32", "bf16x6" = three-piece split-bf16 products with
    the 24 bits of fp32, "bf16x3" = split-bf16 products with about 16
    significant bits per operand, "bf16"; with float64 storage the merge may
    be "f64i8": float64 products on the int8 MFMA from 6 slices of 7 bits per
    operand row, 42 bits of fixed point relative to the row's largest entry,
    for the products with M, N, K >= ops.I8_MIN_K).
    Eigensolvers, traces and log-norms are always float64; eig_q > 0 sets the
    isometry eigensolver's subspace iterations for the step (0: the module
    setting EIG / EIG3)."""
    store: torch.dtype = torch.float64
    env: str = "f64"
    merge: str = "f64"
    eig_q: int = 0

    def env_compute(self):
        return None if self.store == torch.float64 else (self.env if self.env in SPLIT_COMPUTE else None)

    def merge_compute(self):
        if self.store == torch.float64:
            return "f64i8" if self.merge == "f64i8" else None
        return self.merge if self.merge in SPLIT_COMPUTE else None


FP64 = Precision()
FP64_I8 = Precision(torch.float64, "f64", "f64i8")
FP32 = Precision(torch.float32, "f32", "f32")
FP32_BF16ENV = Precision(torch.float32, "bf16", "f32")


class Policy:
    """Precision per RG step.  An error made at step n enters ln Z/V with
    weight 2^-n (lnf[n] = ln c_n / 2^n and T_n only feeds later terms), so
    early steps run in float64 and later ones in lower precision."""

    def __init__(self, fp64_steps=10**9, low=FP32, bf16_env_from=None, bf16_merge_from=None,
                 split_merge_from=None, split6_merge_from=None, split6_env_from=None, eig_q=(), fp64_merge="f64"):
        self.fp64_steps = fp64_steps
        # merges of the float64 steps: "f64" (fp64 MFMA) or "f64i8" (int8 MFMA slices)
        self.fp64_merge = fp64_merge
        # ((after_step, q), ...): fewer eigensolver iterations once a step's error weighs little
        self.eig_q = tuple(eig_q or ())
        self.low = low
        self.split6_env_from = split6_env_from
        self.bf16_env_from = bf16_env_from
        self.bf16_merge_from = bf16_merge_from
        self.split_merge_from = split_merge_from
        self.split6_merge_from = split6_merge_from

    def __call__(self, step):
        if step <= self.fp64_steps:
            return FP64_I8 if self.fp64_merge == "f64i8" else FP64
        env, merge = self.low.env, self.low.merge
        if self.split6_env_from is not None and step > self.split6_env_from:
            env = "bf16x6"
        if self.bf16_env_from is not None and step > self.bf16_env_from:
            env = "bf16"
        if self.split6_merge_from is not None and step > self.split6_merge_from:
            merge = "bf16x6"
        if self.split_merge_from is not None and step > self.split_merge_from:
            merge = "bf16x3"
        if self.bf16_merge_from is not None and step > self.bf16_merge_from:
            merge = "bf16"
        q = 0
        for after, qq in self.eig_q:
            if step > after:
                q = qq
        if (env, merge, q) == (self.low.env, self.low.merge, self.low.eig_q):
            return self.low
        return Precision(self.low.store, env, merge, q)

    def describe(self):
        s = f"fp64 for steps <= {self.fp64_steps}"
        if self.fp64_merge != "f64":
            s += f" (merges {self.fp64_merge})"
        s += f", then {self.low.store} env={self.low.env} merge={self.low.merge}"
        if self.split6_env_from is not None:
            s += f", bf16x6 env after step {self.split6_env_from}"
        if self.bf16_env_from is not None:
            s += f", bf16 env after step {self.bf16_env_from}"
        if self.split6_merge_from is not None:
            s += f", bf16x6 merge after step {self.split6_merge_from}"
        if self.split_merge_from is not None:
            s += f", bf16x3 merge after step {self.split_merge_from}"
        if self.bf16_merge_from is not None:
            s += f", bf16 merge after step {self.bf16_merge_from}"
        for after, q in self.eig_q:
            s += f", {q} eigensolver iterations after step {after}"
        return s


ALL_FP64 = Policy()


# ---------------------------------------------------------------------------
# isometries
# ---------------------------------------------------------------------------
def topk_isometry(MM, k, eig=None):
    """MM: float64 (nb, n, n) environments (every instance and side stacked
    in one batch).  Returns V (nb, n, k), the eigenvectors of the k largest
    eigenvalues in ascending order, and eps (nb,), the discarded eigenvalue
    mass tr(MM) - sum of the kept eigenvalues."""
    nb, n, _ = MM.shape
    eig = EIG if eig is None else eig
    l = min(n, max(eig["l_factor"] * k, k + 8), 256)
    if k > l:
        raise TrgError(f"bond dimension {k} above the eigensolver's block limit")
    q = 0 if l == n else eig["q"]
    _, V, eps, _ = ops.eigh_topk(MM, k, l, q, seed=eig["seed"], batch=nb)
    return V, eps


def _pick(MM, B, Dcut, dtype, eig=None):
    """Eigen-decompose both sides' environments (MM[:B] = side 0, MM[B:] =
    side 1) in one batched call and keep the side with the smaller
    discarded mass, on the device."""
    V, eps = topk_isometry(MM, Dcut, eig)
    return ops.select_isometry(eps[:B], eps[B:], V[:B], V[B:], dtype, batch=B)


def _identity_isometry(B, D, dtype, device):
    n = D * D
    return ops.identity(n, dtype, device, batch=B).view(B, D, D, n)


# ---------------------------------------------------------------------------
# 2D
# ---------------------------------------------------------------------------
def env2d(T, leg, compute=None, out=None):
    """MM_{(ab),(cd)} of trg/HOTRG_2d.py:26 (leg 0) / :27 (leg 1) as float64 (B, Dx^2, Dx^2)
    (written into `out` when given).

    A[a,k,c,n] = sum_ij T[a,i,j,k] T[c,i,j,n]; B[b,k,d,n] = sum_lm T[b,l,k,m] T[d,l,n,m];
    MM[(a,b),(c,d)] = sum_kn A[a,k,c,n] B[b,k,d,n].
    """
    B, Dx, _, Dy, _ = T.shape
    Tl = T if leg == 0 else T.transpose(1, 2)
    P = permute(Tl, (0, 1, 4, 2, 3)).view(B, Dx * Dy, Dx * Dy)          # [(a,k),(i,j)]
    A1 = gemm(P, P.transpose(1, 2), compute=compute)                      # [(a,k),(c,n)]
    Q = permute(Tl, (0, 1, 3, 2, 4)).view(B, Dx * Dy, Dx * Dy)          # [(b,k),(l,m)]
    B1 = gemm(Q, Q.transpose(1, 2), compute=compute)                      # [(b,k),(d,n)]
    A1p = permute(A1.view(B, Dx, Dy, Dx, Dy), (0, 1, 3, 2, 4)).view(B, Dx * Dx, Dy * Dy)  # [(a,c),(k,n)]
    B1p = permute(B1.view(B, Dx, Dy, Dx, Dy), (0, 1, 3, 2, 4)).view(B, Dx * Dx, Dy * Dy)  # [(b,d),(k,n)]
    M = gemm(A1p, B1p.transpose(1, 2), compute=compute)                   # [(a,c),(b,d)]
    src = M.view(B, Dx, Dx, Dx, Dx).permute(0, 1, 3, 2, 4)
    if out is None:
        out = torch.empty((B, Dx * Dx, Dx * Dx), dtype=torch.float64, device=T.device)
    ops.copy_into(out.view(B, Dx, Dx, Dx, Dx), src)
    return out


def projector2d(T, Dcut, prec=FP64):
    """trg/HOTRG_2d.py:22-55 -> (U (B, Dx, Dx, chi') in T's dtype, flag (B,) 0=LD 1=RU)."""
    B, Dx = T.shape[0], T.shape[1]
    n = Dx * Dx
    if n <= Dcut:
        flag = torch.empty(B, dtype=torch.int32, device=T.device)
        ops.fill_(flag, 0)
        return _identity_isometry(B, Dx, T.dtype, T.device), flag
    c = prec.env_compute()
    MM = torch.empty((2 * B, n, n), dtype=torch.float64, device=T.device)
    env2d(T, 0, c, out=MM[:B])
    env2d(T, 1, c, out=MM[B:])
    U, flag = _pick(MM, B, Dcut, T.dtype, dict(EIG, q=prec.eig_q) if prec.eig_q else None)
    return U.view(B, Dx, Dx, Dcut), flag


def merge2d(T1, T2, Ua, Ub, compute=None):
    """acke,bdel,abi,cdj->ijkl (trg/HOTRG_2d.py:65/67) for batched operands.

    Z1[c,k,e,b,i] = sum_a T1[a,(c,k,e)] Ua[a,(b,i)]          2 Dx^3 Dy^2 chi
    Z2[(c,k,i),(l,d)] = sum_(e,b) Z1[e,b,(c,k,i)] T2[(e,b),(l,d)]   2 Dx^3 Dy^3 chi
    out[(k,i,l),j] = sum_(c,d) Z2[k,i,l,(c,d)] Ub[(c,d),j]      2 Dx^2 Dy^2 chi^2
    The c leg is processed in chunks that fit MERGE_BUDGET; the last GEMM
    accumulates over the chunks.
    """
    B, Dx, _, Dy, _ = T1.shape
    chi_a, chi_b = Ua.shape[3], Ub.shape[3]
    es = T1.element_size()
    # The chunk length depends on the per-instance shape only (never on B),
    # so an instance's summation order is the same alone and in any batch;
    # batches too large for the budget are split into sub-batches instead.
    per_c = 2 * Dy * Dy * Dx * max(chi_a, Dy) * es
    cc = max(1, min(Dx, INSTANCE_BUDGET // max(per_c, 1)))
    nsub = max(1, min(B, MERGE_BUDGET // max(per_c * cc, 1)))
    if B > nsub:
        outs = [merge2d(T1[b0:b0 + nsub], T2[b0:b0 + nsub], Ua[b0:b0 + nsub], Ub[b0:b0 + nsub], compute)
                for b0 in range(0, B, nsub)]
        res = torch.empty((B,) + tuple(outs[0].shape[1:]), dtype=T1.dtype, device=T1.device)
        for i, b0 in enumerate(range(0, B, nsub)):
            ops.copy_into(res[b0:b0 + outs[i].shape[0]], outs[i])
        return res
    # fp32 MFMA wants K-outer operands, bf16 MFMA K-contiguous ones (8
    # consecutive k per lane).  The first two GEMMs write their products
    # straight into the layout the next GEMM reads (two-level output index),
    # so no chi^5 intermediate is permuted; the column order of each product
    # keeps its innermost index contiguous in the destination (coalesced
    # epilogue stores).
    kinner = compute in SPLIT_COMPUTE
    # plain bf16: the two operands of the chi^7 GEMM are stored in bf16 (the
    # rounding the MFMA would apply anyway, done once by the producer), which
    # halves the operand traffic of the L2-fed bf16 kernel; same bits
    zdt = torch.bfloat16 if compute == "bf16" and (Dy * Dx) % 8 == 0 else T1.dtype
    if kinner:
        Tp = permute(T2, (0, 4, 2, 3, 1), zdt).view(B, Dy * Dx, Dy * Dx).transpose(1, 2)   # [(e,b),(l,d)]
        Uam = permute(Ua, (0, 1, 3, 2)).view(B, Dx, chi_a * Dx)                     # [a,(i,b)]
    else:
        Tp = permute(T2, (0, 3, 1, 4, 2)).view(B, Dy * Dx, Dy * Dx)      # [(e,b),(l,d)]
        Uam = Ua.reshape(B, Dx, Dx * chi_a)                               # [a,(b,i)]
    T1m = T1.reshape(B, Dx, Dx * Dy * Dy)                                 # [a,(c,k,e)]
    Ubm = Ub.reshape(B, Dx * Dx, chi_b)                                   # [(c,d),j]
    out = torch.empty((B, Dy, chi_a, Dy, chi_b), dtype=T1.dtype, device=T1.device)  # [k,i,l,j]
    outm = out.view(B, Dy * chi_a * Dy, chi_b)
    for c0 in range(0, Dx, cc):
        c1 = min(Dx, c0 + cc)
        nc = c1 - c0
        A = T1m[:, :, c0 * Dy * Dy:c1 * Dy * Dy].transpose(1, 2)          # (B, (c,k,e), a)
        Z1p = torch.empty(B * nc * Dy * Dy * Dx * chi_a, dtype=zdt, device=T1.device)
        # Z1 is computed transposed (Ua^T T1^T): its destination is contiguous
        # along the Ua index, which then runs over the product's rows, so the
        # epilogue stores four consecutive rows as one vector
        UamT, AT = Uam.transpose(1, 2), A.transpose(1, 2)
        if kinner:
            # [c,k,i,e,b]: rows (i | b), columns (c,k | e).  Z1's contraction is short (K = a), so
            # its GEMM is bound by writing the chi^5 product, and its operands come M- and
            # N-contiguous: the split tiers run it on the fp32 MFMA, whose kernel streams that
            # layout as 16-byte vectors (and is at least as accurate as the split products)
            z1c = "f32" if compute in ("bf16x3", "bf16x6") else compute
            gemm(UamT, AT, Z1p, compute=z1c,
                 out_layout=([(chi_a, Dy * Dx), (Dx, 1)], [(nc * Dy, chi_a * Dy * Dx), (Dy, Dx)]))
            Z2a = Z1p.view(B, nc * Dy * chi_a, Dy * Dx)
        else:
            # [e,b,c,k,i]: rows (b | i), columns (c,k | e)
            gemm(UamT, AT, Z1p, compute=compute,
                 out_layout=([(Dx, nc * Dy * chi_a), (chi_a, 1)], [(nc * Dy, chi_a), (Dy, Dx * nc * Dy * chi_a)]))
            Z2a = Z1p.view(B, Dy * Dx, nc * Dy * chi_a).transpose(1, 2)
        # Z2 [(c,k,i),(l,d)] -> [k,i,l,c,d]: rows (c | k,i), columns (l | d)
        Z2p = torch.empty(B * Dy * chi_a * Dy * nc * Dx, dtype=T1.dtype, device=T1.device)
        gemm(Z2a, Tp, Z2p, compute=compute, tag="merge2d_z2",
             out_layout=([(nc, Dx), (Dy * chi_a, Dy * nc * Dx)], [(Dy, nc * Dx), (Dx, 1)]))
        # out^T[j,(k,i,l)] += Ub^T Z2^T: rows j are contiguous in out (vector epilogue).  On the bf16
        # MFMA the small Ub^T slice is copied K-contiguous, so both operands stream as 16-byte
        # vectors through the pipelined kernel (same product bits in any layout)
        UbT = (permute(Ubm[:, c0 * Dx:c1 * Dx, :], (0, 2, 1)) if kinner
               else Ubm[:, c0 * Dx:c1 * Dx, :].transpose(1, 2))
        gemm(UbT, Z2p.view(B, Dy * chi_a * Dy, nc * Dx).transpose(1, 2),
             outm.transpose(1, 2), beta=0.0 if c0 == 0 else 1.0, compute=compute)
        del Z1p, Z2p
    return permute(out, (0, 2, 4, 1, 3))                                   # [i,j,k,l]


def rearrange2d(T, direction):
    """trg/HOTRG_2d.py:8-20 (X: abcd->cdab, an involution) on the batched tensor."""
    if direction in ("Y", "y"):
        return T
    return permute(T, (0, 3, 4, 1, 2))


def step2d(T, Dcut, direction, prec=FP64):
    """One HOTRG 2D step (trg/HOTRG_2d.py:57-72) on a batch; returns the new T
    in prec.store.  (Real tensors: LD and RU use the same U.)"""
    if T.dtype != prec.store:
        T = ops.cast(T, prec.store)
    T = rearrange2d(T, direction)
    U, _ = projector2d(T, Dcut, prec)
    T = merge2d(T, T, U, U, compute=prec.merge_compute())
    return rearrange2d(T, direction)


def normalize_maxabs(T, lnf, step):
    """trg/HOTRG_2d.py:181-185 + lnf[n] = ln c / 2^n, all on the device."""
    B = T.shape[0]
    c = ops.maxabs(T, batch=B)
    ops.normalize_(T, c, lnf, step, batch=B, lnf_stride=lnf.shape[1])
    return c


def renorm2d(T, Dcut, XLOOPS, YLOOPS, policy=ALL_FP64, sync_each_step=False):
    """trg/HOTRG_2d.py:187-216 for a batch: returns (T, lnf (B, X+Y+1))."""
    B = T.shape[0]
    lnf = torch.empty((B, XLOOPS + YLOOPS + 1), dtype=torch.float64, device=T.device)
    ops.fill_(lnf, 0.0)
    health = ops.Health(T.device)
    with health:
        T = ops.contiguous(T, policy(0).store) if T.dtype != policy(0).store else ops.contiguous(T)
        normalize_maxabs(T, lnf, 0)
        nx = ny = n = 0
        while nx < XLOOPS or ny < YLOOPS:
            for direction in ("Y", "X"):
                if direction == "Y" and ny < YLOOPS:
                    ny += 1
                elif direction == "X" and nx < XLOOPS:
                    nx += 1
                else:
                    continue
                n += 1
                T = step2d(T, Dcut, direction, policy(n))
                normalize_maxabs(T, lnf, n)
                if sync_each_step:
                    torch.cuda.current_stream().synchronize()
    health.check("HOTRG 2D solve")
    return T, lnf


def lnz2d(T, lnf, nsteps):
    B = T.shape[0]
    tr = ops.trace_pairs(T, batch=B)
    return ops.lnz(lnf, tr, 2.0 ** nsteps, batch=B)


# ---------------------------------------------------------------------------
# 3D
# ---------------------------------------------------------------------------
def env3d(T, leg, compute=None, out=None):
    """MM_dag1m/1p/2m/2p of trg/HOTRG_3d.py:42-45 for leg 0..3 as float64 (B, D^2, D^2).

    With T1 = T2 = T: A[a,m,c,n] = Gram of T over its 4 legs other than
    (leg, z'), B[b,m,d,n] = Gram over the 4 legs other than (leg, z),
    MM[(a,b),(c,d)] = sum_{m,n} A[a,m,c,n] B[b,m,d,n].
    """
    Bt = T.shape[0]
    shape = list(T.shape[1:])
    D = shape[leg]
    Dz = shape[4]
    others = [i for i in range(4) if i != leg]
    r = math.prod(shape[i] for i in others)
    # P[(a,m),(o1,o2,o3,h)] = T[.., a, .., h, m]: keep leg and z' as rows
    P = permute(T, [0, 1 + leg, 6] + [1 + i for i in others] + [5]).view(Bt, D * Dz, r * Dz)
    A1 = gemm(P, P.transpose(1, 2), compute=compute)                    # [(a,m),(c,n)]
    # Q[(b,m),(o1,o2,o3,l)] = T[.., b, .., m, l]: keep leg and z as rows
    Q = permute(T, [0, 1 + leg, 5] + [1 + i for i in others] + [6]).view(Bt, D * Dz, r * Dz)
    B1 = gemm(Q, Q.transpose(1, 2), compute=compute)                    # [(b,m),(d,n)]
    A1p = permute(A1.view(Bt, D, Dz, D, Dz), (0, 1, 3, 2, 4)).view(Bt, D * D, Dz * Dz)
    B1p = permute(B1.view(Bt, D, Dz, D, Dz), (0, 1, 3, 2, 4)).view(Bt, D * D, Dz * Dz)
    M = gemm(A1p, B1p.transpose(1, 2), compute=compute)                 # [(a,c),(b,d)]
    if out is None:
        out = torch.empty((Bt, D * D, D * D), dtype=torch.float64, device=T.device)
    ops.copy_into(out.view(Bt, D, D, D, D), M.view(Bt, D, D, D, D).permute(0, 1, 3, 2, 4))
    return out


def projector3d(T, Dcut, prec=FP64):
    """trg/HOTRG_3d.py:38-99 -> (U1 (B,Dx,Dx,chi_x), U2 (B,Dy,Dy,chi_y), flag1, flag2)."""
    B, Dx, _, Dy = T.shape[:4]
    c = prec.env_compute()
    res = []
    for legs, D in (((0, 1), Dx), ((2, 3), Dy)):
        n = D * D
        if n <= Dcut:
            flag = torch.empty(B, dtype=torch.int32, device=T.device)
            ops.fill_(flag, 0)
            res.append((_identity_isometry(B, D, T.dtype, T.device), flag))
            continue
        MM = torch.empty((2 * B, n, n), dtype=torch.float64, device=T.device)
        env3d(T, legs[0], c, out=MM[:B])
        env3d(T, legs[1], c, out=MM[B:])
        U, flag = _pick(MM, B, Dcut, T.dtype, EIG3)
        res.append((U.view(B, D, D, Dcut), flag))
    return res[0][0], res[1][0], res[0][1], res[1][1]


def merge3d(T1, T2, U1a, U1b, U2a, U2b, compute=None):
    """ijklez,mnopzf,ima,jnb,koc,lpd->abcdef (trg/HOTRG_3d.py:107-116), batched.

    X1[(j,k,l,e,z),(m,a)] = sum_i T1 U1a                      2 chi^8
    X2[(j,l,e,z,m,a),(o,c)] = sum_k X1 U2a                     2 chi^9
    X3[(j,l,e,a,c),(n,p,f)] = sum_(z,m,o) X2 T2               2 chi^11
    X4 = X3 . U1b over (j,n) ; out = X4 . U2b over (l,p)      2 chi^9 + 2 chi^8
    The output leg a is chunked (INSTANCE_BUDGET, a function of the shape
    only) and the batch split into sub-batches (MERGE_BUDGET) to bound the
    chi^8 intermediates.
    """
    B = T1.shape[0]
    Dx, _, Dy, _, Dz, _ = T1.shape[1:]
    ca, cb = U1a.shape[3], U1b.shape[3]
    cc_, cd = U2a.shape[3], U2b.shape[3]
    es = T1.element_size()
    per_a = 2 * Dx * Dy * Dz * cc_ * Dx * Dy * Dz * es     # X2/X3 bytes per unit of a, per instance
    na = max(1, min(ca, INSTANCE_BUDGET // max(per_a, 1)))   # shape-only chunk length (batch invariant)
    nsub = max(1, min(B, MERGE_BUDGET // max(per_a * na, 1)))
    if B > nsub:
        outs = [merge3d(T1[b0:b0 + nsub], T2[b0:b0 + nsub], U1a[b0:b0 + nsub], U1b[b0:b0 + nsub],
                        U2a[b0:b0 + nsub], U2b[b0:b0 + nsub], compute) for b0 in range(0, B, nsub)]
        res = torch.empty((B,) + tuple(outs[0].shape[1:]), dtype=T1.dtype, device=T1.device)
        for i, b0 in enumerate(range(0, B, nsub)):
            ops.copy_into(res[b0:b0 + outs[i].shape[0]], outs[i])
        return res
    out = torch.empty((B, ca, cb, cc_, cd, Dz, Dz), dtype=T1.dtype, device=T1.device)
    T1m = T1.reshape(B, Dx, Dx * Dy * Dy * Dz * Dz)                     # [i,(j,k,l,e,z)]
    kinner = compute in SPLIT_COMPUTE
    if kinner:
        # The bf16 MFMA wants K-contiguous operands.  The second and third GEMMs are
        # computed transposed and write their products through two-level output
        # indices straight into the layout the next GEMM reads, so neither chi^8
        # intermediate is permuted: X2 lands as [l,e,a,c,j,(z,m,o)] and X3 as
        # [(l,e,a,c,p,f),(j,n)].  T2 is stored [p,f,n,(z,m,o)] (legs m,n,o,p,z,f)
        # and U2a as [(c,o),k].
        T2p = permute(T2, (0, 4, 6, 2, 5, 1, 3)).view(B, Dy * Dz * Dx, Dz * Dx * Dy)    # [(p,f,n),(z,m,o)]
        U2aT = permute(U2a, (0, 3, 2, 1)).view(B, cc_ * Dy, Dy)                         # [(c,o),k]
    else:
        # T2 as [(z,m,o),(n,p,f)]: permute legs (m,n,o,p,z,f) -> (z,m,o,n,p,f)
        T2p = permute(T2, (0, 5, 1, 3, 2, 4, 6)).view(B, Dz * Dx * Dy, Dx * Dy * Dz)
    zmo = Dz * Dx * Dy
    for a0 in range(0, ca, na):
        a1 = min(ca, a0 + na)
        nca = a1 - a0
        # X1[(j,k,l,e,z),(m,a)] = sum_i T1[i,(jklez)] U1a[i,(m,a)]
        U1s = U1a[:, :, :, a0:a1]                                       # [i,m,a]
        X1 = gemm(T1m.transpose(1, 2), ops.contiguous(U1s).view(B, Dx, Dx * nca), compute=compute)
        if kinner:
            # X1 legs (j,k,l,e,z,m,a) -> X1p [(l,e,a,j,z,m),(k)]
            X1p = permute(X1.view(B, Dx, Dy, Dy, Dz, Dz, Dx, nca), (0, 3, 4, 7, 1, 5, 6, 2))
            del X1
            lea, jzm = Dy * Dz * nca, Dx * Dz * Dx
            # X2^T[(c,o),(l,e,a,j,z,m)] = U2a^T . X1p^T, written as [l,e,a,c,j,z,m,o]:
            # rows (c | o) at strides (Dx zmo | 1), columns (l,e,a | j,z,m) at (cc Dx zmo | Dy)
            X2p = torch.empty(B * lea * cc_ * Dx * zmo, dtype=T1.dtype, device=T1.device)
            gemm(U2aT, X1p.view(B, lea * jzm, Dy).transpose(1, 2), X2p, compute=compute,
                 out_layout=([(cc_, Dx * zmo), (Dy, 1)], [(lea, cc_ * Dx * zmo), (jzm, Dy)]))
            del X1p
            # X3^T[(p,f,n),(l,e,a,c,j)] = T2p . X2p^T, written as [(l,e,a,c,p,f),(j,n)]:
            # rows (p,f | n) at strides (Dx^2 | 1), columns (l,e,a,c | j) at (Dy Dz Dx^2 | Dx)
            X3p = torch.empty(B * lea * cc_ * Dy * Dz * Dx * Dx, dtype=T1.dtype, device=T1.device)
            gemm(T2p, X2p.view(B, lea * cc_ * Dx, zmo).transpose(1, 2), X3p, compute=compute, tag="merge3d_x3",
                 out_layout=([(Dy * Dz, Dx * Dx), (Dx, 1)], [(lea * cc_, Dy * Dz * Dx * Dx), (Dx, Dx)]))
            del X2p
            X3p = X3p.view(B, Dy, Dz, nca, cc_, Dy, Dz, Dx, Dx)         # [l,e,a,c,p,f,j,n]
        else:
            # X1 legs (j,k,l,e,z,m,a) -> X1p [(j,l,e,z,m,a),(k)]
            X1p = permute(X1.view(B, Dx, Dy, Dy, Dz, Dz, Dx, nca), (0, 1, 3, 4, 5, 6, 7, 2))
            X2 = gemm(X1p.view(B, -1, Dy), U2a.reshape(B, Dy, Dy * cc_), compute=compute)  # [(j,l,e,z,m,a),(o,c)]
            # X2 legs (j,l,e,z,m,a,o,c) -> [(j,l,e,a,c),(z,m,o)]
            X2p = permute(X2.view(B, Dx, Dy, Dz, Dz, Dx, nca, Dy, cc_), (0, 1, 2, 3, 6, 8, 4, 5, 7))
            del X1, X1p, X2
            X3 = gemm(X2p.view(B, Dx * Dy * Dz * nca * cc_, zmo), T2p, compute=compute,
                      tag="merge3d_x3")                                    # [(j,l,e,a,c),(n,p,f)]
            del X2p
            # X3 legs (j,l,e,a,c,n,p,f) -> [(l,e,a,c,p,f),(j,n)]
            X3p = permute(X3.view(B, Dx, Dy, Dz, nca, cc_, Dx, Dy, Dz), (0, 2, 3, 4, 5, 7, 8, 1, 6))
            del X3
        X4 = gemm(X3p.reshape(B, -1, Dx * Dx), U1b.reshape(B, Dx * Dx, cb), compute=compute)  # [(l,e,a,c,p,f),b]
        del X3p
        # X4 legs (l,e,a,c,p,f,b) -> [(e,a,c,f,b),(l,p)]
        X4p = permute(X4.view(B, Dy, Dz, nca, cc_, Dy, Dz, cb), (0, 2, 3, 4, 6, 7, 1, 5))
        del X4
        X5 = gemm(X4p.view(B, -1, Dy * Dy), U2b.reshape(B, Dy * Dy, cd), compute=compute)  # [(e,a,c,f,b),d]
        del X4p
        # X5 legs (e,a,c,f,b,d) -> out[a,b,c,d,e,f]
        ops.copy_into(out[:, a0:a1], X5.view(B, Dz, nca, cc_, Dz, cb, cd).permute(0, 2, 5, 3, 6, 1, 4))
        del X5
    return out


def rearrange3d(T, direction):
    """trg/HOTRG_3d.py:18-26 on the batched tensor."""
    if direction in ("Z", "z"):
        return T
    if direction in ("X", "x"):
        return permute(T, (0, 3, 4, 5, 6, 1, 2))
    return permute(T, (0, 5, 6, 1, 2, 3, 4))


def restore3d(T, direction):
    """trg/HOTRG_3d.py:28-36."""
    if direction in ("Z", "z"):
        return T
    if direction in ("X", "x"):
        return permute(T, (0, 5, 6, 1, 2, 3, 4))
    return permute(T, (0, 3, 4, 5, 6, 1, 2))


def step3d(T, Dcut, direction, prec=FP64):
    """One HOTRG 3D step (trg/HOTRG_3d.py:101-121) on a batch."""
    if T.dtype != prec.store:
        T = ops.cast(T, prec.store)
    T = rearrange3d(T, direction)
    U1, U2, _, _ = projector3d(T, Dcut, prec)
    T = merge3d(T, T, U1, U1, U2, U2, compute=prec.merge_compute())
    return restore3d(T, direction)


def normalize_trace(T, lnf, step):
    """trg/HOTRG_3d.py:232-238: c = |Tr T| over (xx)(yy)(tt); lnf[n] = ln c / 2^n."""
    B = T.shape[0]
    c = ops.trace_pairs(T, batch=B)
    ops.normalize_(T, c, lnf, step, batch=B, lnf_stride=lnf.shape[1])
    return c


def renorm3d(T, Dcut, XLOOPS, YLOOPS, ZLOOPS, policy=ALL_FP64, sync_each_step=False):
    """trg/HOTRG_3d.py:247-322 for a batch (order Z, Y, X)."""
    B = T.shape[0]
    lnf = torch.empty((B, XLOOPS + YLOOPS + ZLOOPS + 1), dtype=torch.float64, device=T.device)
    ops.fill_(lnf, 0.0)
    health = ops.Health(T.device)
    with health:
        T = ops.contiguous(T, policy(0).store)
        normalize_trace(T, lnf, 0)
        cnt = {"X": 0, "Y": 0, "Z": 0}
        tot = {"X": XLOOPS, "Y": YLOOPS, "Z": ZLOOPS}
        n = 0
        while any(cnt[d] < tot[d] for d in "XYZ"):
            for direction in ("Z", "Y", "X"):
                if cnt[direction] < tot[direction]:
                    cnt[direction] += 1
                    n += 1
                    T = step3d(T, Dcut, direction, policy(n))
                    normalize_trace(T, lnf, n)
                    if sync_each_step:
                        torch.cuda.current_stream().synchronize()
    health.check("HOTRG 3D solve")
    return T, lnf


def lnz3d(T, lnf, nsteps):
    B = T.shape[0]
    tr = ops.trace_pairs(T, batch=B)
    return ops.lnz(lnf, tr, 2.0 ** nsteps, batch=B)
