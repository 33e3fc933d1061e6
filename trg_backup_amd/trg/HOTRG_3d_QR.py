"""Drop-in for the reference's trg/HOTRG_3d_QR.py, computed on the GPU by libtrg.

Same names and arguments as trg/HOTRG_3d_QR.py:29-795 for the pure-tensor
flow.  Configuration the reference reads at import (:19-27) is kept as module
globals with defaults instead of required inputs: `select_squeezer` from
HOTRG_SQUEEZER (default "QR"), `degeneracy_eps` from DGENERACY_EPS (default
0), `truncate_eps` = 1e-10, and `slicing` (sys.argv[4] * 5 in the reference),
which only splits the reference's contraction loop and does not change the
result.  The per-step diagnostics (cal_X, cal_coorelation,
cal_scaling_dimensions) run on the device and write their files when
OUTPUT_DIR is set; prints and timings are not emitted.  The squeezers take
the pure-tensor form (T0 = T1 = T2 = T3); the one-point (:520-557, :798-908)
and two-point (:428-518, :910-1033) impurity flows are on the device.
"""
import math
import os

import torch

from .. import diagnostics, hotrg, hotrg_qr, ops
from .._io import from_device, to_device
from .._lib import TrgError

select_squeezer = os.environ.get("HOTRG_SQUEEZER", "QR")
degeneracy_eps = float(os.environ.get("DGENERACY_EPS", "0") or 0)
truncate_eps = hotrg_qr.TRUNCATE_EPS
slicing = None
PRECISION = hotrg.ALL_FP64


def set_precision(policy):
    """Per-step precision of the device flow (default: float64 throughout)."""
    global PRECISION
    PRECISION = policy


def leg_transposition(T, do_what="transpose", direction="Z"):
    """trg/HOTRG_3d_QR.py:29-72."""
    Td, kind = to_device(T)
    f = hotrg.rearrange3d if do_what == "transpose" else hotrg.restore3d
    return from_device(f(Td[None], direction)[0], kind)


def _pure(T0, T1, T2, T3):
    Td, kind = to_device(T0)
    for other in (T1, T2, T3):
        if other is not T0:
            Od, _ = to_device(other)
            if Od.shape != Td.shape or not torch.equal(Od, Td):
                raise TrgError("the device squeezers take the pure-tensor form T0 = T1 = T2 = T3 "
                               "(trg/HOTRG_3d_QR.py:412); impurity flows are not on the device")
    return ops.contiguous(Td[None]), kind


def _squeezer(kind_sq, T0, T1, T2, T3, Dcut):
    Tb, kind = _pure(T0, T1, T2, T3)
    groups = hotrg_qr.squeeze(Tb, Dcut, kind_sq, degeneracy_eps, hotrg.FP64 if Tb.dtype == torch.float64
                              else hotrg.FP32)
    (_, U1a, U1b, U2a, U2b), = groups
    LX = hotrg.env3d(Tb, 1)
    TdagT = ops.trace_pairs(LX[0])
    Px = ops.contiguous(U1a[0].permute(2, 0, 1))
    PY = ops.contiguous(U2b[0].permute(2, 0, 1))
    tdt = float(TdagT[0].item()) if kind == "numpy" else TdagT[0]
    return (from_device(Px, kind), from_device(U1b[0], kind), from_device(U2a[0], kind), from_device(PY, kind), tdt)


def squeezer_fix(T0, T1, T2, T3, Dcut):
    """trg/HOTRG_3d_QR.py:74-118 -> (Px, PX, Py, PY, TdagT)."""
    return _squeezer("fix", T0, T1, T2, T3, Dcut)


def squeezer_original(T0, T1, T2, T3, Dcut):
    """trg/HOTRG_3d_QR.py:120-193 -> (Px, PX, Py, PY, TdagT)."""
    return _squeezer("original", T0, T1, T2, T3, Dcut)


def squeezer_QR(T0, T1, T2, T3, Dcut):
    """trg/HOTRG_3d_QR.py:195-281 -> (Px, PX, Py, PY, TdagT)."""
    return _squeezer("QR", T0, T1, T2, T3, Dcut)


def squeezer(T0, T1, T2, T3, Dcut, squeezer_type):
    """trg/HOTRG_3d_QR.py:394-406."""
    return _squeezer(squeezer_type, T0, T1, T2, T3, Dcut)


def slice_legs(leg_size, slicing):
    """trg/HOTRG_3d_QR.py:284-316 (host logic): `slicing` contiguous slices of
    range(leg_size), one per index when leg_size < slicing."""
    leg_size, slicing = int(leg_size), int(slicing)
    if leg_size == 0:
        raise ValueError("leg_size or slicing is zero!")
    if leg_size < slicing:
        return [slice(i, i + 1) for i in range(leg_size)], leg_size
    if slicing <= 1:
        return [slice(0, leg_size)], 1
    short = leg_size // slicing
    n_short = slicing - leg_size % slicing
    out, start = [], 0
    for j in range(slicing):
        size = short if j < n_short else short + 1
        out.append(slice(start, start + size))
        start += size
    return out, len(out)


def coarse_graining(T0, T1, Px, PX, Py, PY, slicing=None, Dcut=None):
    """adhezi,bcgfiZ,xab,dcX,hgy,Yef->xXyYzZ (trg/HOTRG_3d_QR.py:318-391) with
    the batched 3D merge; `slicing` does not change the result."""
    (A, kind), (Bt, _) = to_device(T0), to_device(T1)
    P = [to_device(p)[0] for p in (Px, PX, Py, PY)]
    dt = A.dtype
    U1a = ops.contiguous(ops.cast(P[0], dt).permute(1, 2, 0))[None]
    U1b = ops.contiguous(ops.cast(P[1], dt))[None]
    U2a = ops.contiguous(ops.cast(P[2], dt))[None]
    U2b = ops.contiguous(ops.cast(P[3], dt).permute(1, 2, 0))[None]
    out = hotrg.merge3d(ops.contiguous(A[None]), ops.contiguous(ops.cast(Bt, dt)[None]), U1a, U1b, U2a, U2b)
    return from_device(out[0], kind)


def normalization(T):
    """trg/HOTRG_3d_QR.py:715-719 -> (T / c, c), c = |Tr T| over (xx)(yy)(zz)."""
    Td, kind = to_device(T)
    Tb = ops.contiguous(Td[None])
    c = ops.trace_pairs(Tb)
    ops.normalize_(Tb, c)
    return from_device(Tb[0], kind), abs(float(c[0].item()))


def new_pure_tensor(T, Dcut, direction):
    """trg/HOTRG_3d_QR.py:408-425 on the dict {'Tensor', ...}."""
    Td, kind = to_device(T["Tensor"])
    groups = hotrg_qr.step(ops.contiguous(Td[None]), Dcut, direction, select_squeezer, degeneracy_eps,
                           hotrg.FP64 if Td.dtype == torch.float64 else hotrg.FP32)
    T["Tensor"] = from_device(groups[0][1][0], kind)
    return T


def exec_rgstep_pure_tensor(T, Dcut, direction, rgstep):
    """trg/HOTRG_3d_QR.py:734-745."""
    key = f"({rgstep['X']},{rgstep['Y']},{rgstep['Z']})"
    T = new_pure_tensor(T, Dcut, direction)
    T["Tensor"], c = normalization(T["Tensor"])
    T["factor"][key] = math.log(c) / 2 ** sum(rgstep.values())
    return T


def pure_tensor(T, Dcut, TOT_RGSTEPS):
    """trg/HOTRG_3d_QR.py:747-795: the whole flow on the device (one instance,
    no host round trip per step); fills T['factor'] with the reference's
    "(X,Y,Z)" keys and returns T.  With OUTPUT_DIR set, the X ratios,
    correlation lengths and scaling dimensions of the initial and of every
    coarse-grained tensor go to the reference's files (:757-759, :781-783)."""
    Td, kind = to_device(T["Tensor"])
    rgstep.update({"X": 0, "Y": 0, "Z": 0})
    hook = _step_diagnostics(Dcut) if os.environ.get("OUTPUT_DIR") else None
    groups = hotrg_qr.renorm(Td[None], Dcut, TOT_RGSTEPS, select_squeezer, degeneracy_eps, PRECISION, on_step=hook)
    (_, Tn, lnf), = groups
    for key, v in zip(hotrg_qr.factor_keys(TOT_RGSTEPS), lnf[0].tolist()):
        T["factor"][key] = v
    T["Tensor"] = from_device(Tn[0], kind)
    return T


def ln_Z_over_V(T, Dcut, TOT_RGSTEPS):
    """ln Z / V as measurement/Ising_model_3d.py:101-137 assembles it from
    pure_tensor (sum of the factors + ln Tr T / 2^N), on the device."""
    Td, _ = to_device(T)
    return float(hotrg_qr.ln_z_over_v(Td[None], Dcut, TOT_RGSTEPS, select_squeezer, degeneracy_eps,
                                      PRECISION)[0].item())


def _prec_of(Td):
    return hotrg.FP64 if Td.dtype == torch.float64 else hotrg.FP32


def _trace_norm(Td):
    """(T / |Tr T|, |Tr T|) on the device, Tr over (xx)(yy)(zz)."""
    Tb = ops.contiguous(Td[None])
    c = ops.trace_pairs(Tb)
    ops.normalize_(Tb, c)
    return Tb[0], abs(float(c[0].item()))


def _impure_step(Td, T0d, odd, Dcut, direction):
    """One step of the pure tensor and one impurity (trg/HOTRG_3d_QR.py:520-557):
    projectors from the pure tensor; the impurity is the upper tensor of its
    pair when `odd`, the lower one otherwise."""
    prec = _prec_of(Td)
    Tb = hotrg.rearrange3d(ops.contiguous(Td[None]), direction)
    T0b = hotrg.rearrange3d(ops.contiguous(T0d[None]), direction)
    (_, U1a, U1b, U2a, U2b), = hotrg_qr.squeeze(Tb, Dcut, select_squeezer, degeneracy_eps, prec)
    lo, up = (Tb, T0b) if odd else (T0b, Tb)
    T0n = hotrg.merge3d(lo, up, U1a, U1b, U2a, U2b, compute=prec.merge_compute())
    Tn = hotrg.merge3d(Tb, Tb, U1a, U1b, U2a, U2b, compute=prec.merge_compute())
    return hotrg.restore3d(Tn, direction)[0], hotrg.restore3d(T0n, direction)[0]


def new_impuer_tensor_1point(T, T0, Dcut, direction):
    """trg/HOTRG_3d_QR.py:520-557 on the dicts (T0['loc'] gives the impurity's coordinates)."""
    (Td, kind), (T0d, _) = to_device(T["Tensor"]), to_device(T0["Tensor"])
    Tn, T0n = _impure_step(Td, T0d, T0["loc"][direction] % 2 == 1, Dcut, direction)
    T["Tensor"], T0["Tensor"] = from_device(Tn, kind), from_device(T0n, kind)
    return T, T0


def exec_rgstep_1point_func(T, T0, Dcut, direction, rgstep):
    """trg/HOTRG_3d_QR.py:833-857: step, normalise both tensors, store
    ln c / 2^n for T and ln c0 - ln c for T0, halve the impurity coordinate."""
    key = f"({rgstep['X']},{rgstep['Y']},{rgstep['Z']})"
    (Td, kind), (T0d, _) = to_device(T["Tensor"]), to_device(T0["Tensor"])
    Tn, T0n = _impure_step(Td, T0d, T0["loc"][direction] % 2 == 1, Dcut, direction)
    Tn, c = _trace_norm(Tn)
    T0n, c0 = _trace_norm(T0n)
    T["Tensor"], T0["Tensor"] = from_device(Tn, kind), from_device(T0n, kind)
    T["factor"][key] = math.log(c) / 2 ** sum(rgstep.values())
    T0["factor"][key] = math.log(c0) - math.log(c)
    T0["loc"][direction] //= 2
    return T, T0


def cal_field_expected_value(T, T0, rgstep, save):
    """trg/HOTRG_3d_QR.py:798-831: returns (ln Z/V, [P_i]) with
    P_i = exp(sum T0[i] factors) Tr T0[i] / Tr T; with save and OUTPUT_DIR set,
    appends the reference's line to field_expected_value_rgsteps_dependence.dat."""
    def tr(x):
        Xd, _ = to_device(x)
        return float(ops.trace_pairs(ops.contiguous(Xd[None]))[0].item())
    TrT = tr(T["Tensor"])
    P = [math.exp(sum(t["factor"].values())) * tr(t["Tensor"]) / TrT for t in T0]
    n = sum(rgstep.values())
    lnz = sum(T["factor"].values()) + math.log(TrT) / 2 ** n
    out = os.environ.get("OUTPUT_DIR")
    if save and out:
        with open(os.path.join(out, "field_expected_value_rgsteps_dependence.dat"), "w" if n == 0 else "a") as f:
            f.write(f"{n} {lnz:.12e} {P[0]:.12e} {0.0}\n")
    return lnz, P


def one_point_function(T, T0, Dcut, TOT_RGSTEPS):
    """trg/HOTRG_3d_QR.py:859-908: the pure tensor and one impurity through
    the "ZXY" cycle; returns (T, T0) with their factor dicts filled.  The
    one-point value is cal_field_expected_value(T, [T0], ...)[1][0]."""
    rgstep = {"X": 0, "Y": 0, "Z": 0}
    (Td, kind), (T0d, _) = to_device(T["Tensor"]), to_device(T0["Tensor"])
    Td, c = _trace_norm(Td)
    T0d, c0 = _trace_norm(T0d)
    T["Tensor"], T0["Tensor"] = from_device(Td, kind), from_device(T0d, kind)
    T["factor"]["(0,0,0)"] = math.log(c)
    T0["factor"]["(0,0,0)"] = math.log(c0) - math.log(c)
    for direction in hotrg_qr.schedule(TOT_RGSTEPS):
        rgstep[direction] += 1
        T, T0 = exec_rgstep_1point_func(T, T0, Dcut, direction, rgstep)
    return T, T0


# ---------------------------------------------------------------------------
# diagnostics (trg/HOTRG_3d_QR.py:559-713) and their per-step files
# ---------------------------------------------------------------------------
# step counters of the running flow (the reference's global, :747-750); the
# diagnostics name their files and ln xi by them
rgstep = {"X": 0, "Y": 0, "Z": 0}


def _mode():
    return "w" if sum(rgstep.values()) == 0 else "a"


def _out_dir(save):
    d = os.environ.get("OUTPUT_DIR")
    return d if save and d else None


def cal_X(Tensor, save=False):
    """:559-583: X_t, X_x, X_y = (Tr T)^2 / TT with TT the trace of two tensors
    stacked along the leg, on the device; with ``save`` one line
    "Xt Xx Xy" goes to OUTPUT_DIR/X.dat (new file at step 0)."""
    Td, _ = to_device(Tensor["Tensor"])
    Xt, Xx, Xy = diagnostics.x_ratios(Td)
    d = _out_dir(save)
    if d:
        with open(os.path.join(d, "X.dat"), _mode()) as f:
            f.write(f"{Xt:.12e} {Xx:.12e} {Xy:.12e}\n")
    return Xt, Xx, Xy


def find_1stto3rdmax(Tensor, save=False):
    """:585-611: the three diagonal entries T[x,x,y,y,t,t] of largest |value|
    (the diagonal is gathered on the device, the three picked on the host);
    with ``save`` one line goes to OUTPUT_DIR/max_1to3-th_values.dat.  The
    reference returns nothing; this returns the three values."""
    import numpy as np
    Td, _ = to_device(Tensor["Tensor"])
    dg = diagnostics.diagonal(Td).cpu().numpy()
    top = np.pad(dg[np.argsort(-np.abs(dg), kind="stable")[:3]], (0, max(0, 3 - dg.size)))
    d = _out_dir(save)
    if d:
        with open(os.path.join(d, "max_1to3-th_values.dat"), _mode()) as f:
            f.write(f"{top[0]:40.12e} {top[1]:40.12e} {top[2]:40.12e}\n")
    return top


def cal_scaling_dimensions(Tensor, Dcut, save=False):
    """:613-637: eigenvalues of the t-density matrix "xxyytT->Tt" (device
    Jacobi on its lower triangle), descending, floored at 1e-16, padded to
    Dcut with 1e-16, log(e0 / e) / (2 pi); with ``save`` one line goes to
    OUTPUT_DIR/scaling_dimensions.dat.  Returns the scaling dimensions."""
    import numpy as np
    Td, _ = to_device(Tensor["Tensor"])
    e = diagnostics.eigvals_lower(diagnostics.density_matrix(Td, "t")).cpu().numpy()[::-1].copy()
    e[e < 1e-16] = 1e-16
    e = np.pad(e, (0, max(0, Dcut - len(e))), constant_values=1e-16)
    sd = np.log(np.max(e) / e) / (2 * np.pi)
    d = _out_dir(save)
    if d:
        with open(os.path.join(d, "scaling_dimensions.dat"), _mode()) as f:
            f.write(" ".join(f"{v:.12e}" for v in sd) + "\n")
    return sd


def cal_coorelation(Tensor, Dcut, save=False):
    """:639-713: singular values (device Jacobi) of the t, x and y density
    matrices, zero-padded to Dcut^2 and normalised by their largest, and
    ln xi = n ln 2 - ln(-1 / ln e_1) with n = rgstep["Z"], ["X"], ["Y"]; the
    y branch tests the x leg's e_1 against 1e-200, as the reference does
    (:700).  With ``save`` the files lnξt.dat, lnξx.dat, lnξy.dat and
    density_matrix_t/DentMatT_lx*_ly*_lt*.dat are written under OUTPUT_DIR.
    Returns ({leg: e}, {leg: ln xi}) (the reference prints them)."""
    import numpy as np
    Td, _ = to_device(Tensor["Tensor"])
    ev, vmax = {}, {}
    for leg in "txy":
        e = diagnostics.singular_values(diagnostics.density_matrix(Td, leg)).cpu().numpy()
        e = np.pad(e, (0, max(0, Dcut * Dcut - len(e))))
        vmax[leg] = float(np.max(e))
        ev[leg] = e / vmax[leg]
    lnxi = {}
    for leg, key in (("t", "Z"), ("x", "X"), ("y", "Y")):
        cond = ev["x"][1] if leg == "y" else ev[leg][1]
        e1 = ev[leg][1] if cond > 1e-200 else 1e-200
        lnxi[leg] = rgstep[key] * math.log(2) - math.log(-1 / math.log(e1))
    d = _out_dir(save)
    if d:
        mode = _mode()
        for leg in "txy":
            with open(os.path.join(d, f"lnξ{leg}.dat"), mode) as f:
                f.write(f"{lnxi[leg]:.12e}\n")
        dm = os.path.join(d, "density_matrix_t")
        os.makedirs(dm, exist_ok=True)
        with open(os.path.join(dm, f"DentMatT_lx{rgstep['X']}_ly{rgstep['Y']}_lt{rgstep['Z']}.dat"), mode) as f:
            f.write(f"λ1={vmax['t']:.12e}\n")
            f.writelines(f"{v:.12e}\n" for v in ev["t"])
    return ev, lnxi


def __get_tensor_spectrum(T, Dcut):
    """:726-730: the leading 2 Dcut singular values of T as the (x y z) x
    (x' y' z') matrix over the largest (the reference prints them); returns
    (s / s_max, s_max)."""
    Td, _ = to_device(T)
    return diagnostics.tensor_spectrum(Td, 2 * Dcut)


def _step_diagnostics(Dcut):
    """The per-step calls of :757-759 and :781-783 as a renorm hook."""
    def hook(n, direction, groups):
        if direction is not None:
            rgstep[direction] += 1
        T = {"Tensor": groups[0][1][0]}
        cal_X(T, save=True)
        cal_coorelation(T, Dcut, save=True)
        cal_scaling_dimensions(T, Dcut, save=True)
    return hook


# ---------------------------------------------------------------------------
# two-point function (trg/HOTRG_3d_QR.py:428-518, 910-1033)
# ---------------------------------------------------------------------------
def sqr_distance(loc1, loc2):
    """Squared distance of two location dicts."""
    return sum((loc1[k] - loc2[k]) ** 2 for k in loc1)


def _squeeze_pure(Tb, Dcut, direction):
    (_, U1a, U1b, U2a, U2b), = hotrg_qr.squeeze(Tb, Dcut, select_squeezer, degeneracy_eps, _prec_of(Tb))
    return U1a, U1b, U2a, U2b


def _merge(lo, up, P, prec):
    return hotrg.merge3d(lo, up, *P, compute=prec.merge_compute())


def new_impuer_tensor_2points(T, T0, Tn, Dcut, direction):
    """trg/HOTRG_3d_QR.py:428-475 on the dicts: projectors from the pure
    tensor; each impurity is merged with a pure partner, as the upper tensor
    of its pair when its coordinate along `direction` is odd."""
    (Td, kind), (T0d, _), (Tnd, _) = to_device(T["Tensor"]), to_device(T0["Tensor"]), to_device(Tn["Tensor"])
    prec = _prec_of(Td)
    Tb, T0b, Tnb = (hotrg.rearrange3d(ops.contiguous(x[None]), direction) for x in (Td, T0d, Tnd))
    P = _squeeze_pure(Tb, Dcut, direction)
    pair = lambda Xb, odd: _merge(Tb, Xb, P, prec) if odd else _merge(Xb, Tb, P, prec)
    T0n = pair(T0b, T0["loc"][direction] % 2 == 1)
    Tnn = pair(Tnb, Tn["loc"][direction] % 2 == 1)
    Tn_pure = _merge(Tb, Tb, P, prec)
    T["Tensor"] = from_device(hotrg.restore3d(Tn_pure, direction)[0], kind)
    T0["Tensor"] = from_device(hotrg.restore3d(T0n, direction)[0], kind)
    Tn["Tensor"] = from_device(hotrg.restore3d(Tnn, direction)[0], kind)
    return T, T0, Tn


def new_impuer_tensor_absorb(T, T0, Tn, Dcut, direction):
    """trg/HOTRG_3d_QR.py:477-518 on the dicts: the two impurities form one
    pair (the one at the even coordinate is the lower tensor) and T0 becomes
    their merged impurity; returns (T, T0)."""
    (Td, kind), (T0d, _), (Tnd, _) = to_device(T["Tensor"]), to_device(T0["Tensor"]), to_device(Tn["Tensor"])
    prec = _prec_of(Td)
    Tb, T0b, Tnb = (hotrg.rearrange3d(ops.contiguous(x[None]), direction) for x in (Td, T0d, Tnd))
    P = _squeeze_pure(Tb, Dcut, direction)
    p0, pn = T0["loc"][direction] % 2, Tn["loc"][direction] % 2
    if p0 == 0 and pn == 1:
        T0["Tensor"] = from_device(hotrg.restore3d(_merge(T0b, Tnb, P, prec), direction)[0], kind)
    elif p0 == 1 and pn == 0:
        T0["Tensor"] = from_device(hotrg.restore3d(_merge(Tnb, T0b, P, prec), direction)[0], kind)
    T["Tensor"] = from_device(hotrg.restore3d(_merge(Tb, Tb, P, prec), direction)[0], kind)
    return T, T0


def _norm_dict(X):
    Xd, kind = to_device(X["Tensor"])
    Xn, c = _trace_norm(Xd)
    X["Tensor"] = from_device(Xn, kind)
    return c


def exec_rgstep_2point_func(T, T0, Tn, Dcut, direction, rgstep):
    """trg/HOTRG_3d_QR.py:910-995: while both impurities exist and their
    coordinates stay apart after the step, the two-point step with factor
    ln c0 + ln cn - 2 ln c for T0; when they meet, the absorb step (factor
    ln c0 - ln c, Tn becomes {}); with Tn already gone, the one-point step.
    Returns (T, T0, Tn)."""
    key = f"({rgstep['X']},{rgstep['Y']},{rgstep['Z']})"
    n = sum(rgstep.values())
    if len(Tn) > 0:
        after0 = dict(T0["loc"], **{direction: T0["loc"][direction] // 2})
        aftern = dict(Tn["loc"], **{direction: Tn["loc"][direction] // 2})
        if sqr_distance(after0, aftern) > 0:
            T, T0, Tn = new_impuer_tensor_2points(T, T0, Tn, Dcut, direction)
            c, c0, c1 = _norm_dict(T), _norm_dict(T0), _norm_dict(Tn)
            T["factor"][key] = math.log(c) / 2 ** n
            T0["factor"][key] = math.log(c0) + math.log(c1) - 2 * math.log(c)
            T0["loc"][direction] //= 2
            Tn["loc"][direction] //= 2
            return T, T0, Tn
        T, T0 = new_impuer_tensor_absorb(T, T0, Tn, Dcut, direction)
    else:
        T, T0 = new_impuer_tensor_1point(T, T0, Dcut, direction)
    c, c0 = _norm_dict(T), _norm_dict(T0)
    T["factor"][key] = math.log(c) / 2 ** n
    T0["factor"][key] = math.log(c0) - math.log(c)
    T0["loc"][direction] //= 2
    return T, T0, {}


def two_point_function(T, T0, Tn, Dcut, TOT_RGSTEPS):
    """trg/HOTRG_3d_QR.py:997-1033: normalise the three tensors (T0's factor
    ln c0 + ln cn - 2 ln c), then the "ZXY" cycle of exec_rgstep_2point_func.
    Returns (T, T0); the correlator is cal_field_expected_value(T, [T0], ...)[1][0]."""
    rgstep = {"X": 0, "Y": 0, "Z": 0}
    c, c0, c1 = _norm_dict(T), _norm_dict(T0), _norm_dict(Tn)
    T["factor"]["(0,0,0)"] = math.log(c)
    T0["factor"]["(0,0,0)"] = math.log(c0) + math.log(c1) - 2 * math.log(c)
    for direction in hotrg_qr.schedule(TOT_RGSTEPS):
        rgstep[direction] += 1
        T, T0, Tn = exec_rgstep_2point_func(T, T0, Tn, Dcut, direction, rgstep)
    return T, T0
