"""Drop-in for the reference's trg/ATRG_3d_new.py pure-tensor path, on the GPU.

Same names as trg/ATRG_3d_new.py:18-482 and 954-1039.  ``pure_tensor``
takes an ATRG_Tensor (U, s, VH as NumPy arrays or GPU tensors) and returns it
coarse grained, with normal_const filled as in :967-972.  Environment:
DGENERACY_EPS (default 0.0, the reference requires it, :15) and OUTPUT_DIR:
when it is set, pure_tensor writes the reference's per-step files
(tensor_sv/*.dat, X.dat, leg_exchange_pure/*.dat; :197-212, :736-766,
:991-993, :1023-1025); without it nothing is written.  The one-point
(:614-734, :1043-1189, and the single-impurity form :1191-1331) and two-point
(:485-612, :1333-1489) impurity flows run on the device with the pure tensor,
as do the diagnostics find_max, find_1stto3rdmax, cal_scaling_dimensions and
cal_coorelation (:768-951).
``set_precision("f32")`` runs the contractions in float32 on MFMA;
decompositions stay float64.
"""
import itertools
import math
import os

import torch

from .. import atrg, linalg, ops
from ..contract import contract
from .._io import from_device, to_device
from .._lib import TrgError

atrg.degeneracy_eps = float(os.environ.get("DGENERACY_EPS", "0.0"))
degeneracy_eps = atrg.degeneracy_eps
truncate_eps = atrg.truncate_eps
PRECISION = {"contract": "f64", "fp64_steps": 0}


def set_precision(contract_dtype="f64", fp64_steps=0):
    """Contraction dtype of the coarse-graining GEMMs ("f64" or "f32"); the
    first `fp64_steps` RG steps run in float64 regardless (an error made at
    step n enters ln Z/V with weight 2^-n).  Decompositions, environments of
    the squeezer and the normalisation are always float64."""
    PRECISION["contract"] = contract_dtype
    PRECISION["fp64_steps"] = int(fp64_steps)


def _step_precision(n):
    return "f64" if n <= PRECISION["fp64_steps"] else PRECISION["contract"]


def set_rsvd_params(k_factor=3, p=0, q=None, seed=12345):
    """Leg-swap rSVD parameters (reference: 3, 0, Dcut, 12345 at :193)."""
    atrg.set_rsvd_params(k_factor, p, q, seed)


def _dev3(T):
    (U, kind), (s, _), (VH, _) = to_device(T.U), to_device(T.s), to_device(T.VH)
    return U, s, VH, kind


def leg_transposition(A, do_what="transpose", direction="t"):
    """:18-65 on an ATRG_Tensor (in place, like the reference)."""
    U, s, VH, kind = _dev3(A)
    U, VH = atrg.leg_transposition(U, VH, do_what, direction)
    A.U, A.VH = from_device(U, kind), from_device(VH, kind)
    return A


def rsvd_for_ATRG_leg_swapping(B, C, chi_k, p=32, q=64, eps=1e-10, seed=12345):
    """:67-140 -> UM (i,g,k), sM (k), VMH (k,j,b)."""
    (Bd, kind), (Cd, _) = to_device(B), to_device(C)
    out = atrg.rsvd_for_ATRG_leg_swapping(Bd, Cd, chi_k, p, q, eps, seed)
    return tuple(from_device(o, kind) for o in out)


def leg_swapping(B, C, Dcut, ispure=True):
    """:142-235."""
    (Bd, kind), (Cd, _) = to_device(B), to_device(C)
    out = atrg.leg_swapping(Bd, Cd, Dcut, ispure, PRECISION["contract"])
    return tuple(from_device(o, kind) for o in out)


def squeezer(A, UM, sM, VMH, D, Dcut):
    """:246-317 -> Px, PX, Py, PY."""
    ts = [to_device(x) for x in (A, UM, sM, VMH, D)]
    kind = ts[0][1]
    out = atrg.squeezer(*[t for t, _ in ts], Dcut)
    return tuple(from_device(o, kind) for o in out)


def leg_slicing(leg_size, slicing):
    """:319-355 (host-side slice list).  A zero-size leg raises ValueError
    (the reference prints and calls sys.exit(0), :330-334)."""
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


def coarse_graining(A, UM, sM, VMH, D, Px, PX, Py, PY, slicing=100):
    """:357-438 (the slice loops of the reference only bound memory; the GPU
    evaluates the sums as GEMM chains)."""
    ts = [to_device(x) for x in (A, UM, sM, VMH, D, Px, PX, Py, PY)]
    kind = ts[0][1]
    out = atrg.coarse_graining(*[t for t, _ in ts])
    return tuple(from_device(o, kind) for o in out)


def atrg_pure_tensor(T, Dcut, direction):
    """:441-482 (updates T in place and returns it)."""
    U, s, VH, kind = _dev3(T)
    U, s, VH = atrg.atrg_pure_tensor(U, s, VH, Dcut, direction, PRECISION["contract"])
    T.update(from_device(U, kind), from_device(s, kind), from_device(VH, kind))
    return T


def normalization(T):
    """:954-958: c = |Tr T|, s /= c."""
    U, s, VH, kind = _dev3(T)
    s2, c = atrg.normalize(U, s, VH)
    T.s = from_device(s2, kind)
    cv = abs(float(c[0].item()))
    return T, cv


def exec_pure_tensor_renorm(T, Dcut, direction, rgstep):
    """:967-972."""
    T = atrg_pure_tensor(T, Dcut, direction)
    key = f"({rgstep['X']},{rgstep['Y']},{rgstep['T']})"
    T, c = normalization(T)
    T.normal_const[key] = math.log(c) / 2 ** (sum(rgstep.values()))
    return T


def _reset_steps():
    """Zero the module's step counters and return them (the flows below
    advance them in place, as the reference's globals, :975-982)."""
    global rgstep
    rgstep.update({"X": 0, "Y": 0, "T": 0})
    return rgstep


def _log_norm(c, nsteps):
    """ln|c| / 2^nsteps of a device normalisation constant; a zero or non-finite
    norm (a tensor that vanished or overflowed) raises instead of returning inf."""
    v = abs(float(c[0].item()))
    if not (v > 0.0 and math.isfinite(v)):
        raise TrgError(f"ATRG normalisation constant is {v} after {nsteps} RG steps")
    return math.log(v) / 2 ** nsteps


def _leg_exchange_writer(out_dir):
    """The leg-swap spectrum file of trg/ATRG_3d_new.py:197-212: sM (the
    reference divides by 1), zero-padded to 3 max(Dcut, chi_k), to
    leg_exchange_{pure,impure}/leg_exchange_sv_lx{X}_ly{Y}_lt{T}.dat."""
    def hook(sM, k, ispure):
        d = os.path.join(out_dir, "leg_exchange_pure" if ispure else "leg_exchange_impure")
        os.makedirs(d, exist_ok=True)
        vals = [float(v) for v in sM.double().cpu().tolist()]
        vals += [0.0] * max(0, 3 * k - len(vals))
        with open(os.path.join(d, f"leg_exchange_sv_lx{rgstep['X']}_ly{rgstep['Y']}_lt{rgstep['T']}.dat"), "w") as f:
            f.write(f"#s1={1.0:.12e}\n")
            f.writelines(f"{v:.12e}\n" for v in vals)
    return hook


def _step_outputs(U, s, VH, Dcut):
    """The per-step files of :991-993 and :1023-1025 (tensor_sv/*.dat and
    X.dat) for the current device tensors; only with OUTPUT_DIR set."""
    from ..tensor_class.tensor_class import ATRG_Tensor
    t = ATRG_Tensor(U, s, VH, Dcut, 3, False, {})
    save_singularvalues(t, Dcut)
    cal_X(t, save=True)


def pure_tensor(T, Dcut, TOT_RGSTEPS):
    """:974-1039: cycle T, X, Y until the totals are reached.  The tensor stays
    on the GPU between steps and is returned in the caller's array kind."""
    U, s, VH, kind = _dev3(T)
    rgstep = _reset_steps()
    out_dir = os.environ.get("OUTPUT_DIR")
    old_hook = atrg.LEG_SWAP_HOOK
    if out_dir:
        atrg.LEG_SWAP_HOOK = _leg_exchange_writer(out_dir)
    health = ops.Health(U.device)
    try:
        with health:
            s, c = atrg.normalize(U, s, VH)
            T.normal_const[f"({rgstep['X']},{rgstep['Y']},{rgstep['T']})"] = _log_norm(c, 0)
            if out_dir:
                _step_outputs(U, s, VH, Dcut)
            total = sum(TOT_RGSTEPS.values())
            for direction in itertools.cycle("TXY"):
                if sum(rgstep.values()) >= total:
                    break
                if rgstep[direction] < TOT_RGSTEPS[direction]:
                    rgstep[direction] += 1
                    U, s, VH = atrg.atrg_pure_tensor(U, s, VH, Dcut, direction, _step_precision(sum(rgstep.values())))
                    s, c = atrg.normalize(U, s, VH)
                    key = f"({rgstep['X']},{rgstep['Y']},{rgstep['T']})"
                    T.normal_const[key] = _log_norm(c, sum(rgstep.values()))
                    if out_dir:
                        _step_outputs(U, s, VH, Dcut)
    finally:
        atrg.LEG_SWAP_HOOK = old_hook
    health.check("ATRG solve")
    T.update(from_device(U, kind), from_device(s, kind), from_device(VH, kind))
    return T


def ln_Z_over_V(T, Dcut, TOT_RGSTEPS):
    """measurement/Potts3_model_3d.py:32-54 on the device."""
    T = pure_tensor(T, Dcut, TOT_RGSTEPS)
    U, s, VH, _ = _dev3(T)
    tr = abs(float(atrg.trace(U, s, VH)[0].item()))
    V = 2 ** sum(TOT_RGSTEPS.values())
    return sum(T.normal_const.values()) + math.log(tr) / V, T


# ---------------------------------------------------------------------------
# one-point function (impurity) flow, :614-734 and :1043-1189
# ---------------------------------------------------------------------------
def _imp_args(T0, direction):
    out = []
    for t in T0:
        U0, s0, VH0, _ = _dev3(t)
        out.append((U0, s0, VH0, t.loc[direction] % 2 == 1))
    return out


def new_impuer_tensor_1point_MultipleImpureTensors(T, T0, Dcut, direction):
    """:669-734: one step of the pure tensor T and of every impurity in the
    list T0 (each ATRG_Tensor with ``loc``), all updated in place."""
    for i, t in enumerate(T0):
        assert hasattr(t, "U") and hasattr(t, "loc"), f"T0[{i}] must be ATRG_Tensor type"
    U, s, VH, kind = _dev3(T)
    (U, s, VH), outs = atrg.atrg_1point_step(U, s, VH, _imp_args(T0, direction), Dcut, direction,
                                             PRECISION["contract"])
    T.update(from_device(U, kind), from_device(s, kind), from_device(VH, kind))
    for t, (U0, s0, VH0) in zip(T0, outs):
        t.update(from_device(U0, kind), from_device(s0, kind), from_device(VH0, kind))
    return T, T0


def new_impuer_tensor_1point(T, T0, Dcut, direction):
    """:614-667: the one-impurity form of the step above."""
    T, (T0,) = new_impuer_tensor_1point_MultipleImpureTensors(T, [T0], Dcut, direction)
    return T, T0


def exec_rgstep_1point_func(T, T0, Dcut, direction, rgstep):
    """:1079-1101: step, normalise T (factor ln c / 2^n) and every impurity
    (factor ln c0 - ln c), halve the impurity coordinates along `direction`."""
    key = f"({rgstep['X']},{rgstep['Y']},{rgstep['T']})"
    T, T0 = new_impuer_tensor_1point_MultipleImpureTensors(T, T0, Dcut, direction)
    T, c = normalization(T)
    T.normal_const[key] = math.log(c) / 2 ** (sum(rgstep.values()))
    for t in T0:
        t, c0 = normalization(t)
        t.normal_const[key] = math.log(c0) - math.log(c)
        t.loc[direction] //= 2
    return T, T0


def cal_field_expected_value(T, T0, rgstep, save=False):
    """:1043-1076: ln Z/V and P_i = exp(sum of T0[i]'s factors) Tr T0[i] / Tr T.
    Returns (lnZ/V, [P_i]); with ``save`` the line "nsteps lnZ/V P_0 P_1" is
    written to OUTPUT_DIR/field_expected_value_rgsteps_dependence.dat like the
    reference (new file at step 0, appended after)."""
    U, s, VH, _ = _dev3(T)
    trT = float(atrg.trace(U, s, VH)[0].item())
    n = sum(rgstep.values())
    lnz = float(sum(T.normal_const.values())) + math.log(trT) / 2 ** n
    P = []
    for t in T0:
        U0, s0, VH0, _ = _dev3(t)
        P.append(math.exp(float(sum(t.normal_const.values()))) * float(atrg.trace(U0, s0, VH0)[0].item()) / trT)
    if save:
        out_dir = os.environ.get("OUTPUT_DIR")
        if out_dir:
            with open(os.path.join(out_dir, "field_expected_value_rgsteps_dependence.dat"),
                      "w" if n == 0 else "a") as f:
                f.write(f"{n} {lnz:.12e} " + " ".join(f"{p:.12e}" for p in P[:2]) + "\n")
    return lnz, P


def one_point_function(T, T0, Dcut, TOT_RGSTEPS):
    """:1103-1189: cycle T, X, Y until the totals are reached, carrying the
    impurities T0 (a list of ATRG_Tensor with ``loc``) with the pure tensor.
    The tensors stay on the GPU between steps and are returned in the
    caller's array kind; the per-step singular-value files (:736-748) are not
    written."""
    for i, t in enumerate(T0):
        assert hasattr(t, "U") and hasattr(t, "loc"), f"T0[{i}] must be ATRG_Tensor type"
    U, s, VH, kind = _dev3(T)
    imps = [list(_dev3(t)[:3]) for t in T0]
    rgstep = _reset_steps()

    def norm_all(n_total):
        nonlocal s
        key = f"({rgstep['X']},{rgstep['Y']},{rgstep['T']})"
        s, c = atrg.normalize(U, s, VH)
        lc = math.log(abs(float(c[0].item())))
        T.normal_const[key] = lc / 2 ** n_total
        for t, imp in zip(T0, imps):
            imp[1], c0 = atrg.normalize(*imp)
            t.normal_const[key] = math.log(abs(float(c0[0].item()))) - lc

    norm_all(0)
    total = sum(TOT_RGSTEPS.values())
    for direction in itertools.cycle("TXY"):
        if sum(rgstep.values()) >= total:
            break
        if rgstep[direction] < TOT_RGSTEPS[direction]:
            rgstep[direction] += 1
            args = [(u, sv, vh, t.loc[direction] % 2 == 1) for t, (u, sv, vh) in zip(T0, imps)]
            (U, s, VH), outs = atrg.atrg_1point_step(U, s, VH, args, Dcut, direction,
                                                     _step_precision(sum(rgstep.values())))
            imps = [list(o) for o in outs]
            norm_all(sum(rgstep.values()))
            for t in T0:
                t.loc[direction] //= 2
    T.update(from_device(U, kind), from_device(s, kind), from_device(VH, kind))
    for t, (U0, s0, VH0) in zip(T0, imps):
        t.update(from_device(U0, kind), from_device(s0, kind), from_device(VH0, kind))
    return T, T0


# ---------------------------------------------------------------------------
# small helpers and output files (:736-766, :960-965)
# ---------------------------------------------------------------------------
# step counters of the running flow, like the reference's module globals
# (:975-982); save_singularvalues and cal_X name their output by them
rgstep = {"X": 0, "Y": 0, "T": 0}


def sqr_distance(loc1, loc2):
    """:960-965: squared distance of two location dicts."""
    return sum((a - b) * (a - b) for a, b in zip(loc1.values(), loc2.values()))


def save_singularvalues(T, k):
    """:736-748: s / max(s), zero-padded to max(k, len(s)), to
    OUTPUT_DIR/tensor_sv/tensor_sv_lx{X}_ly{Y}_lt{T}.dat (no-op without OUTPUT_DIR)."""
    out_dir = os.environ.get("OUTPUT_DIR")
    if not out_dir:
        return
    s, _ = to_device(T.s)
    s = s.double().cpu().numpy() if hasattr(s, "cpu") else s
    d = os.path.join(out_dir, "tensor_sv")
    os.makedirs(d, exist_ok=True)
    smax = float(max(s))
    vals = list(s / smax) + [0.0] * (max(k, len(s)) - len(s))
    with open(os.path.join(d, f"tensor_sv_lx{rgstep['X']}_ly{rgstep['Y']}_lt{rgstep['T']}.dat"), "w") as f:
        f.write(f"#s1={smax:.12e}\n")
        f.writelines(f"{v:.12e}\n" for v in vals)


def cal_X(T, save=False):
    """:750-766: the X ratios of tensor_class.ATRG_Tensor.cal_X; with ``save``
    one line "Xt Xx Xy" goes to OUTPUT_DIR/X.dat (new file at step 0)."""
    Xt, Xx, Xy = T.cal_X()
    out_dir = os.environ.get("OUTPUT_DIR")
    if save and out_dir:
        with open(os.path.join(out_dir, "X.dat"), "w" if sum(rgstep.values()) == 0 else "a") as f:
            f.write(f"{float(Xt):.12e} {float(Xx):.12e} {float(Xy):.12e}\n")
    return Xt, Xx, Xy


def _scaled_u(T):
    """(U s (T,X,Y,i) float64, VH float64) of an ATRG tensor on the device."""
    U, s, VH, _ = _dev3(T)
    Us = ops.cast(U, torch.float64)
    ops.scale_rows_(Us, ops.cast(s, torch.float64).reshape(-1), Us.shape[3], 1, 0)
    return Us, ops.cast(VH, torch.float64)


def _density_matrix(T, leg):
    """Partial traces of the tensor T_{TXYtxy} = U s VH (:856, :884-923):
    leg "t" -> [T,t] (x, y traced), "x" -> [X,x], "y" -> [Y,y], on the device."""
    Us, VH = _scaled_u(T)
    spec = {"t": "TXYi,itXY->Tt", "x": "TXYi,iTxY->Xx", "y": "TXYi,iTXy->Yy"}[leg]
    return contract(spec, Us, VH)


def _append_mode():
    return "w" if sum(rgstep.values()) == 0 else "a"


def find_max(T, max_nth):
    """:768-816: the max_nth entries of the full tensor T_{TXYtxy} = U s VH with the
    largest |value| and their 6-leg indices.  The slices Tsub[T,X,Y,t] of each
    (x, y) are contracted on the device and the running top entries are kept
    on the host (the reference sorts each merged slice the same way).  The
    reference prints and calls sys.exit(0) at :813-814; this returns
    (values, indices) instead."""
    import numpy as np
    Us, VH = _scaled_u(T)
    chiT, chiX, chiY, _ = Us.shape
    shape6 = (chiT, chiX, chiY, VH.shape[1], chiX, chiY)
    best_v = np.empty(0)
    best_i = np.empty(0, dtype=np.int64)
    for ix in range(chiX):
        for iy in range(chiY):
            sub = contract("TXYi,it->TXYt", Us, VH[:, :, ix, iy]).cpu().numpy().ravel()
            k = np.arange(sub.size)
            TT, XX, YY, tt = np.unravel_index(k, (chiT, chiX, chiY, VH.shape[1]))
            flat = np.ravel_multi_index((TT, XX, YY, tt, np.full_like(k, ix), np.full_like(k, iy)), shape6)
            v = np.concatenate((best_v, sub))
            i = np.concatenate((best_i, flat))
            order = np.argsort(-np.abs(v), kind="stable")[:max_nth]
            best_v, best_i = v[order], i[order]
    return best_v, np.array(np.unravel_index(best_i, shape6)).T


def find_1stto3rdmax(T, save=False):
    """:818-852: the three diagonal entries T[i,j,k,i,j,k] of largest |value|
    (diagonal = sum_a U[i,j,k,a] s_a VH[a,i,j,k], contracted on the device);
    with ``save`` one line goes to OUTPUT_DIR/max_1to3-th_values.dat (new file
    at step 0).  Returns the three values."""
    import numpy as np
    Us, VH = _scaled_u(T)
    d = contract("ijka,aijk->ijk", Us, VH).cpu().numpy().ravel()
    top = d[np.argsort(-np.abs(d), kind="stable")[:3]]
    top = np.pad(top, (0, 3 - len(top)))
    out_dir = os.environ.get("OUTPUT_DIR")
    if save and out_dir:
        with open(os.path.join(out_dir, "max_1to3-th_values.dat"), _append_mode()) as f:
            f.write(f"{top[0]:40.12e} {top[1]:40.12e} {top[2]:40.12e}\n")
    return top


def _lower_symmetric(A):
    """The symmetric matrix of A's lower triangle (what cupy.linalg.eigh reads,
    UPLO="L"), built with device copies."""
    n = A.shape[0]
    S = ops.contiguous(A)
    for i in range(n - 1):
        ops.copy_into(S[i, i + 1:], A[i + 1:, i])
    return S


def cal_scaling_dimensions(T, save=False):
    """:854-877: eigenvalues of the t-density matrix (device Jacobi on its lower
    triangle), descending, floored at 1e-16, padded to Dcut with 1e-16, and
    log(e0 / e) / (2 pi); with ``save`` one line goes to
    OUTPUT_DIR/scaling_dimensions.dat.  Returns the scaling dimensions."""
    import numpy as np
    M = _lower_symmetric(_density_matrix(T, "t"))
    w, _ = linalg.eigh_small(M[None])
    e = w[0].cpu().numpy()[::-1].copy()
    e[e < 1e-16] = 1e-16
    e = np.pad(e, (0, max(0, T.Dcut - len(e))), constant_values=1e-16)
    sd = np.log(np.max(e) / e) / (2 * np.pi)
    out_dir = os.environ.get("OUTPUT_DIR")
    if save and out_dir:
        with open(os.path.join(out_dir, "scaling_dimensions.dat"), _append_mode()) as f:
            f.write(" ".join(f"{v:.12e}" for v in sd) + "\n")
    return sd


def cal_coorelation(T, save=False):
    """:879-951: singular values of the t, x and y density matrices (device
    Jacobi SVD) normalised by their largest, and ln xi = n ln 2 - ln(-1 / ln e_1)
    per leg (n = this leg's RG steps); the y branch tests the x leg's e_1
    against 1e-200, as the reference does (:944).  With ``save`` the files
    lnξt.dat, lnξx.dat, lnξy.dat and density_matrix_t/DentMatT_*.dat are written
    under OUTPUT_DIR.  Returns ({leg: e}, {leg: ln xi})."""
    import numpy as np
    ev = {}
    for leg in ("t", "x", "y"):
        M = _density_matrix(T, leg)
        _, sv, _ = linalg.svd_small(M[None])
        e = sv[0].cpu().numpy()
        if len(e) < T.Dcut:
            e = np.pad(e, (0, T.Dcut - len(e)))
        ev[leg] = e / np.max(e)
        if leg == "t":
            etmax = float(np.max(e))
    lnxi = {}
    for leg, key in (("t", "T"), ("x", "X"), ("y", "Y")):
        cond = ev["x"][1] if leg == "y" else ev[leg][1]
        e1 = ev[leg][1] if cond > 1e-200 else 1e-200
        lnxi[leg] = rgstep[key] * math.log(2) - math.log(-1 / math.log(e1))
    out_dir = os.environ.get("OUTPUT_DIR")
    if save and out_dir:
        os.makedirs(out_dir, exist_ok=True)
        mode = _append_mode()
        for leg in ("t", "x", "y"):
            with open(os.path.join(out_dir, f"lnξ{leg}.dat"), mode) as f:
                f.write(f"{lnxi[leg]:.12e}\n")
        dm = os.path.join(out_dir, "density_matrix_t")
        os.makedirs(dm, exist_ok=True)
        with open(os.path.join(dm, f"DentMatT_lx{rgstep['X']}_ly{rgstep['Y']}_lt{rgstep['T']}.dat"), mode) as f:
            f.write(f"λ1={etmax:.12e}\n")
            f.writelines(f"{v:.12e}\n" for v in ev["t"])
    return ev, lnxi


# ---------------------------------------------------------------------------
# one-point flow with a single impurity tensor (:1191-1331)
# ---------------------------------------------------------------------------
def cal_field_expected_value1(T, T0, rgstep, save=False):
    """:1191-1222: ln Z/V and P = exp(sum of T0's factors) Tr T0 / Tr T for one
    impurity tensor; with ``save`` the line "nsteps lnZ/V P" goes to
    OUTPUT_DIR/field_expected_value_rgsteps_dependence.dat.  Returns (lnZ/V, P)."""
    lnz, P = cal_field_expected_value(T, [T0], rgstep, False)
    n = sum(rgstep.values())
    out_dir = os.environ.get("OUTPUT_DIR")
    if save and out_dir:
        with open(os.path.join(out_dir, "field_expected_value_rgsteps_dependence.dat"),
                  "w" if n == 0 else "a") as f:
            f.write(f"{n} {lnz:.12e} {P[0]:.12e}\n")
    return lnz, P[0]


def exec_rgstep_1point_func1(T, T0, Dcut, direction, rgstep):
    """:1224-1244: one step of T and the single impurity T0, their
    normalisations (factors ln c / 2^n and ln c0 - ln c) and T0's location."""
    T, (T0,) = exec_rgstep_1point_func(T, [T0], Dcut, direction, rgstep)
    return T, T0


def one_point_function1(T, T0, Dcut, TOT_RGSTEPS):
    """:1246-1331: the single-impurity flow, cycling T, X, Y; after the initial
    normalisation and after every step it writes the singular values
    (save_singularvalues) and the field expectation line
    (cal_field_expected_value1, save=True) when OUTPUT_DIR is set, as the
    reference does.  Returns (T, T0)."""
    assert hasattr(T0, "U") and hasattr(T0, "loc"), "T0 must be ATRG_Tensor type"
    step = _reset_steps()
    key = f"({step['X']},{step['Y']},{step['T']})"
    T, c = normalization(T)
    T.normal_const[key] = math.log(c)
    T0, c0 = normalization(T0)
    T0.normal_const[key] = math.log(c0) - math.log(c)
    save_singularvalues(T, Dcut)
    cal_field_expected_value1(T, T0, step, True)
    total = sum(TOT_RGSTEPS.values())
    for direction in itertools.cycle("TXY"):
        if sum(step.values()) >= total:
            break
        if step[direction] < TOT_RGSTEPS[direction]:
            step[direction] += 1
            T, T0 = exec_rgstep_1point_func1(T, T0, Dcut, direction, step)
            save_singularvalues(T, Dcut)
            cal_field_expected_value1(T, T0, step, True)
    return T, T0


# ---------------------------------------------------------------------------
# two-point function (impurity) flow, :485-612 and :1333-1489
# ---------------------------------------------------------------------------
def _dev_imp(t):
    U0, s0, VH0, _ = _dev3(t)
    return (U0, s0, VH0)


def new_impuer_tensor_2points(T, T0, Tn, Dcut, direction):
    """:485-555: each impurity pairs with a pure tensor, as the upper tensor
    when its coordinate along `direction` is odd, the lower one when even
    (the one-point step's pairing; the reference's even branch, :519-521,
    contracts U with s and s with VH, and :545 passes an extra first argument
    to coarse_graining, which both fail on the shapes).  Updates T, T0, Tn
    in place and returns them."""
    U, s, VH, kind = _dev3(T)
    pairs = [((_dev_imp(t), None) if t.loc[direction] % 2 == 1 else (None, _dev_imp(t))) for t in (T0, Tn)]
    (U, s, VH), outs = atrg.atrg_pairs_step(U, s, VH, pairs, Dcut, direction, PRECISION["contract"])
    T.update(from_device(U, kind), from_device(s, kind), from_device(VH, kind))
    for t, (U0, s0, VH0) in zip((T0, Tn), outs):
        t.update(from_device(U0, kind), from_device(s0, kind), from_device(VH0, kind))
    return T, T0, Tn


def new_impuer_tensor_absorb(T, T0, Tn, Dcut, direction):
    """:557-612: the two impurities form one pair (the one at the odd
    coordinate is the upper tensor) and T0 becomes their merged impurity;
    returns (T, T0)."""
    U, s, VH, kind = _dev3(T)
    up, lo = (Tn, T0) if T0.loc[direction] % 2 == 0 else (T0, Tn)
    (U, s, VH), ((U0, s0, VH0),) = atrg.atrg_pairs_step(U, s, VH, [(_dev_imp(up), _dev_imp(lo))], Dcut, direction,
                                                       PRECISION["contract"])
    T.update(from_device(U, kind), from_device(s, kind), from_device(VH, kind))
    T0.update(from_device(U0, kind), from_device(s0, kind), from_device(VH0, kind))
    return T, T0


def exec_rgstep_2point_func(T, T0, Tn, Dcut, direction, rgstep):
    """:1333-1417: while the impurities stay apart after the step, the
    two-point step (T0's factor ln c0 + ln cn - 2 ln c); when they meet, the
    absorb step (factor ln c0 - ln c, Tn becomes None); with Tn None, the
    one-point step.  Returns (T, T0, Tn)."""
    key = f"({rgstep['X']},{rgstep['Y']},{rgstep['T']})"
    n = sum(rgstep.values())
    if Tn is not None:
        after0 = dict(T0.loc, **{direction: T0.loc[direction] // 2})
        aftern = dict(Tn.loc, **{direction: Tn.loc[direction] // 2})
        if sqr_distance(after0, aftern) > 0:
            T, T0, Tn = new_impuer_tensor_2points(T, T0, Tn, Dcut, direction)
            T, c = normalization(T)
            T0, c0 = normalization(T0)
            Tn, c1 = normalization(Tn)
            T.normal_const[key] = math.log(c) / 2 ** n
            T0.normal_const[key] = math.log(c0) + math.log(c1) - 2 * math.log(c)
            T0.loc[direction] //= 2
            Tn.loc[direction] //= 2
            return T, T0, Tn
        T, T0 = new_impuer_tensor_absorb(T, T0, Tn, Dcut, direction)
    else:
        T, T0 = new_impuer_tensor_1point(T, T0, Dcut, direction)
    T, c = normalization(T)
    T0, c0 = normalization(T0)
    T.normal_const[key] = math.log(c) / 2 ** n
    T0.normal_const[key] = math.log(c0) - math.log(c)
    T0.loc[direction] //= 2
    return T, T0, None


def two_point_function(T, T0, Tn, Dcut, TOT_RGSTEPS):
    """:1419-1489: normalise the three tensors (T0's factor ln c0 + ln cn -
    2 ln c), then cycle T, X, Y with exec_rgstep_2point_func.  Returns
    (T, T0); the correlator is cal_field_expected_value(T, [T0], steps)[1][0]."""
    rgstep = _reset_steps()
    T, c = normalization(T)
    T0, c0 = normalization(T0)
    Tn, c1 = normalization(Tn)
    T.normal_const["(0,0,0)"] = math.log(c)
    T0.normal_const["(0,0,0)"] = math.log(c0) + math.log(c1) - 2 * math.log(c)
    total = sum(TOT_RGSTEPS.values())
    for direction in itertools.cycle("TXY"):
        if sum(rgstep.values()) >= total:
            break
        if rgstep[direction] < TOT_RGSTEPS[direction]:
            rgstep[direction] += 1
            T, T0, Tn = exec_rgstep_2point_func(T, T0, Tn, Dcut, direction, rgstep)
    return T, T0
