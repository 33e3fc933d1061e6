"""Drop-in for the reference's trg/ATRG_3d_new.py pure-tensor path, on the GPU.

Same names as trg/ATRG_3d_new.py:18-482 and 954-1039.  ``pure_tensor``
takes an ATRG_Tensor (U, s, VH as NumPy arrays or GPU tensors) and returns it
coarse grained, with normal_const filled as in :967-972.  Environment:
DGENERACY_EPS (default 0.0, the reference requires it, :15); the per-step
files under OUTPUT_DIR (:197-212, :736-766) are not written.  The impurity
and scaling-dimension flows (:485-951, :1043-1489) are outside this module
(see the package documentation).  ``set_precision("f32")`` runs the
contractions in float32 on MFMA; decompositions stay float64.
"""
import itertools
import math
import os

import torch

from .. import atrg, linalg, ops
from .._io import from_device, to_device

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


def pure_tensor(T, Dcut, TOT_RGSTEPS):
    """:974-1039: cycle T, X, Y until the totals are reached.  The tensor stays
    on the GPU between steps and is returned in the caller's array kind."""
    U, s, VH, kind = _dev3(T)
    rgstep = {"X": 0, "Y": 0, "T": 0}
    s, c = atrg.normalize(U, s, VH)
    T.normal_const[f"({rgstep['X']},{rgstep['Y']},{rgstep['T']})"] = math.log(abs(float(c[0].item())))
    total = sum(TOT_RGSTEPS.values())
    for direction in itertools.cycle("TXY"):
        if sum(rgstep.values()) >= total:
            break
        if rgstep[direction] < TOT_RGSTEPS[direction]:
            rgstep[direction] += 1
            U, s, VH = atrg.atrg_pure_tensor(U, s, VH, Dcut, direction, _step_precision(sum(rgstep.values())))
            s, c = atrg.normalize(U, s, VH)
            key = f"({rgstep['X']},{rgstep['Y']},{rgstep['T']})"
            T.normal_const[key] = math.log(abs(float(c[0].item()))) / 2 ** (sum(rgstep.values()))
    T.update(from_device(U, kind), from_device(s, kind), from_device(VH, kind))
    return T


def ln_Z_over_V(T, Dcut, TOT_RGSTEPS):
    """measurement/Potts3_model_3d.py:32-54 on the device."""
    T = pure_tensor(T, Dcut, TOT_RGSTEPS)
    U, s, VH, _ = _dev3(T)
    tr = abs(float(atrg.trace(U, s, VH)[0].item()))
    V = 2 ** sum(TOT_RGSTEPS.values())
    return sum(T.normal_const.values()) + math.log(tr) / V, T
