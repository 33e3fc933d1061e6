"""Drop-in for the reference's trg/HOTRG_2d.py, computed on the GPU by libtrg.

Same function names, arguments and return kinds as trg/HOTRG_2d.py:8-348.
Differences (documented): the bond may grow up to Dcut instead of having to
equal Dcut (:53; zero padding gives the same ln Z/V); the untruncated case
returns LDRU = "LD" (:39-40 leaves it unassigned); the per-step timing
prints (:204,214) are not emitted.  ``set_precision`` selects the
per-step precision policy (default: float64 throughout).
"""
import torch

from .. import hotrg, ops
from .._io import from_device, to_device

PRECISION = hotrg.ALL_FP64


def set_precision(policy):
    """policy: a hotrg.Policy (or a callable step -> hotrg.Precision)."""
    global PRECISION
    PRECISION = policy


def __tensor_legs_rearrange__(T, direction):
    """trg/HOTRG_2d.py:8-13."""
    Td, kind = to_device(T)
    return from_device(hotrg.rearrange2d(Td[None], direction)[0], kind)


def __tensor_legs_restore__(T, direction):
    """trg/HOTRG_2d.py:15-20."""
    return __tensor_legs_rearrange__(T, direction)


def __contraction_projector__(T, Dcut):
    """trg/HOTRG_2d.py:22-55 -> (U (D, D, chi'), "LD" | "RU")."""
    Td, kind = to_device(T)
    U, flag = hotrg.projector2d(Td[None].contiguous(), Dcut, PRECISION(1) if Td.dtype == torch.float32 else hotrg.FP64)
    return from_device(U[0], kind), ("LD" if int(flag[0].item()) == 0 else "RU")


def new_pure_tensor(T, Dcut, direction):
    """trg/HOTRG_2d.py:57-72."""
    Td, kind = to_device(T)
    prec = hotrg.FP64 if Td.dtype == torch.float64 else hotrg.FP32
    return from_device(hotrg.step2d(Td[None].contiguous(), Dcut, direction, prec)[0], kind)


def _impure_step(T, imps, pairs, Dcut, direction):
    """Shared body of the impurity updates: one projector from the pure tensor,
    merges of the listed (first, second) operand pairs (indices into
    [T] + imps), legs restored afterwards."""
    prec = hotrg.FP64 if T.dtype == torch.float64 else hotrg.FP32
    tensors = [hotrg.rearrange2d(t[None].contiguous(), direction) for t in [T] + imps]
    U, _ = hotrg.projector2d(tensors[0], Dcut, prec)
    outs = [hotrg.merge2d(tensors[a], tensors[b], U, U, compute=prec.merge_compute()) for a, b in pairs]
    return [hotrg.rearrange2d(o, direction)[0] for o in outs]


def new_impuer_tensor_2imp(T, Timp0, Timp1, nx, ny, Dcut, direction):
    """trg/HOTRG_2d.py:75-128."""
    (Td, kind), (I0, _), (I1, _) = to_device(T), to_device(Timp0), to_device(Timp1)
    isy = direction in ("Y", "y")
    odd = (ny % 2 == 1) if isy else (nx % 2 == 1)
    T2, A, B = _impure_step(Td, [I0, I1], [(0, 0), (1, 0), (0, 2) if odd else (2, 0)], Dcut, direction)
    return from_device(T2, kind), from_device(A, kind), from_device(B, kind)


def new_impuer_tensor_2to1imp(T, Timp0, Timp1, Dcut, direction):
    """trg/HOTRG_2d.py:130-153."""
    (Td, kind), (I0, _), (I1, _) = to_device(T), to_device(Timp0), to_device(Timp1)
    T2, A = _impure_step(Td, [I0, I1], [(0, 0), (1, 2)], Dcut, direction)
    return from_device(T2, kind), from_device(A, kind)


def new_impuer_tensor_1imp(T, Timp0, Dcut, direction):
    """trg/HOTRG_2d.py:155-177."""
    (Td, kind), (I0, _) = to_device(T), to_device(Timp0)
    T2, A = _impure_step(Td, [I0], [(0, 0), (1, 0)], Dcut, direction)
    return from_device(T2, kind), from_device(A, kind)


def __tensor_normalization__(T):
    """trg/HOTRG_2d.py:181-185 -> (T / c, c) with c = max |T| (float64)."""
    Td, kind = to_device(T)
    Tb = ops.contiguous(Td[None])
    c = ops.maxabs(Tb)
    ops.normalize_(Tb, c)
    cv = float(c[0].item()) if kind == "numpy" else c[0]
    return from_device(Tb[0], kind), cv


def pure_tensor_renorm(T, Dcut, XLOOPS, YLOOPS):
    """trg/HOTRG_2d.py:187-216 -> (T, ln_normalized_factor) with the precision policy."""
    Td, kind = to_device(T)
    Tb, lnf = hotrg.renorm2d(Td[None], Dcut, XLOOPS, YLOOPS, PRECISION)
    return from_device(Tb[0], kind), from_device(lnf[0], kind)


def two_point_func_renorm(T, Timp0, Timp1, nx, ny, Dcut, XLOOPS, YLOOPS):
    """trg/HOTRG_2d.py:219-297 (impurity pair separated by (nx, ny))."""
    from . import _impurity
    return _impurity.two_point_2d(T, Timp0, Timp1, nx, ny, Dcut, XLOOPS, YLOOPS)


def ynearest_two_point_func_renorm(T, Timp0, Timp1, Dcut, XLOOPS, YLOOPS):
    """trg/HOTRG_2d.py:300-348."""
    from . import _impurity
    return _impurity.ynearest_2d(T, Timp0, Timp1, Dcut, XLOOPS, YLOOPS)


def ln_Z_over_V(T, Dcut, XLOOPS, YLOOPS):
    """Caller-side assembly (measurement/SU2_pcm_measurement.py:30-36) on the device."""
    Td, _ = to_device(T)
    Tb, lnf = hotrg.renorm2d(Td[None], Dcut, XLOOPS, YLOOPS, PRECISION)
    return float(hotrg.lnz2d(Tb, lnf, XLOOPS + YLOOPS)[0].item())
