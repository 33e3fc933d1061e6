"""Drop-in for the reference's trg/HOTRG_3d.py, computed on the GPU by libtrg.

Same names and arguments as trg/HOTRG_3d.py:18-453.  Differences
(documented): no MPI / device binding at import (:9-14); the diagnostic
prints and density counts (:61-64, :267-276) are not emitted and the flow
does not call the spectrum diagnostic, which is a function returning its
values (``__get_tensor_spectrum``, :240-245); bonds may grow up to Dcut; untruncated steps return PM = "minus";
``new_impuer_tensor_2to1imp`` restores the legs (the reference re-applies the
rearrangement, :191-192); ``new_impuer_tensor_2imp`` uses the 3D merge (the
reference's 2D merge on 6-leg tensors, :133-147, cannot run).
"""
import torch

from .. import diagnostics, hotrg, ops
from .._io import from_device, to_device

PRECISION = hotrg.ALL_FP64


def set_precision(policy):
    global PRECISION
    PRECISION = policy


def _prec(Td):
    return hotrg.FP64 if Td.dtype == torch.float64 else hotrg.FP32


def __tensor_legs_rearrange__(T, direction):
    """trg/HOTRG_3d.py:18-26."""
    Td, kind = to_device(T)
    return from_device(hotrg.rearrange3d(Td[None], direction)[0], kind)


def __tensor_legs_restore__(T, direction):
    """trg/HOTRG_3d.py:28-36."""
    Td, kind = to_device(T)
    return from_device(hotrg.restore3d(Td[None], direction)[0], kind)


def __contraction_projector__(T, Dcut):
    """trg/HOTRG_3d.py:38-99 -> (U1, U2, PM1, PM2)."""
    Td, kind = to_device(T)
    U1, U2, f1, f2 = hotrg.projector3d(Td[None].contiguous(), Dcut, _prec(Td))
    pm = lambda f: "minus" if int(f[0].item()) == 0 else "plus"
    return from_device(U1[0], kind), from_device(U2[0], kind), pm(f1), pm(f2)


def new_pure_tensor(T, Dcut, direction):
    """trg/HOTRG_3d.py:101-121."""
    Td, kind = to_device(T)
    return from_device(hotrg.step3d(Td[None].contiguous(), Dcut, direction, _prec(Td))[0], kind)


def _impure(T, imps, pairs, Dcut, direction):
    prec = _prec(T)
    ts = [hotrg.rearrange3d(t[None].contiguous(), direction) for t in [T] + imps]
    U1, U2, _, _ = hotrg.projector3d(ts[0], Dcut, prec)
    return [hotrg.restore3d(hotrg.merge3d(ts[a], ts[b], U1, U1, U2, U2, compute=prec.merge_compute()),
                            direction)[0] for a, b in pairs]


def new_impuer_tensor_2imp(T, Timp0, Timp1, nx, ny, Dcut, direction):
    """trg/HOTRG_3d.py:124-157 (3D form)."""
    (Td, kind), (I0, _), (I1, _) = to_device(T), to_device(Timp0), to_device(Timp1)
    odd = (ny % 2 == 1) if direction in ("Y", "y") else (nx % 2 == 1)
    out = _impure(Td, [I0, I1], [(0, 0), (1, 0), (0, 2) if odd else (2, 0)], Dcut, direction)
    return tuple(from_device(o, kind) for o in out)


def new_impuer_tensor_2to1imp(T, Timp0, Timp1, Dcut, direction):
    """trg/HOTRG_3d.py:159-194."""
    (Td, kind), (I0, _), (I1, _) = to_device(T), to_device(Timp0), to_device(Timp1)
    out = _impure(Td, [I0, I1], [(0, 0), (1, 2)], Dcut, direction)
    return tuple(from_device(o, kind) for o in out)


def new_impuer_tensor_1imp(T, Timp0, Dcut, direction):
    """trg/HOTRG_3d.py:196-228."""
    (Td, kind), (I0, _) = to_device(T), to_device(Timp0)
    out = _impure(Td, [I0], [(0, 0), (1, 0)], Dcut, direction)
    return tuple(from_device(o, kind) for o in out)


def __tensor_normalization__(T):
    """trg/HOTRG_3d.py:232-238 -> (T / c, c), c = |Tr T| over (xx)(yy)(tt)."""
    Td, kind = to_device(T)
    Tb = ops.contiguous(Td[None])
    c = ops.trace_pairs(Tb)
    ops.normalize_(Tb, c)
    cv = abs(float(c[0].item())) if kind == "numpy" else c[0].abs()
    return from_device(Tb[0], kind), cv


def __get_tensor_spectrum(T, Dcut):
    """:240-245: the leading 2 Dcut singular values of T as the (x y z) x
    (x' y' z') matrix over the largest, and the largest (the reference prints
    both); returns (s / s_max, s_max)."""
    Td, _ = to_device(T)
    return diagnostics.tensor_spectrum(Td, 2 * Dcut)


def pure_tensor_renorm(T, Dcut, XLOOPS, YLOOPS, ZLOOPS):
    """trg/HOTRG_3d.py:247-322."""
    Td, kind = to_device(T)
    Tb, lnf = hotrg.renorm3d(Td[None], Dcut, XLOOPS, YLOOPS, ZLOOPS, PRECISION)
    return from_device(Tb[0], kind), from_device(lnf[0], kind)


def two_point_func_renorm(T, Timp0, Timp1, nx, ny, Dcut, XLOOPS, YLOOPS):
    """trg/HOTRG_3d.py:325-395 is a 2D leftover that cannot run on 6-leg tensors."""
    raise NotImplementedError("HOTRG_3d.two_point_func_renorm is a 2D leftover in the reference "
                              "(trg/HOTRG_3d.py:325-395); use ynearest_two_point_func_renorm")


def ynearest_two_point_func_renorm(T, Timp0, Timp1, Dcut, XLOOPS, YLOOPS, ZLOOPS):
    """trg/HOTRG_3d.py:398-453."""
    from . import _impurity
    return _impurity.ynearest_3d(T, Timp0, Timp1, Dcut, XLOOPS, YLOOPS, ZLOOPS)


def ln_Z_over_V(T, Dcut, XLOOPS, YLOOPS, ZLOOPS):
    """measurement/Ising_model_3d.py:134-137 on the device."""
    Td, _ = to_device(T)
    Tb, lnf = hotrg.renorm3d(Td[None], Dcut, XLOOPS, YLOOPS, ZLOOPS, PRECISION)
    return float(hotrg.lnz3d(Tb, lnf, XLOOPS + YLOOPS + ZLOOPS)[0].item())
