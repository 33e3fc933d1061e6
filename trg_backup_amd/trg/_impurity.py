"""Impurity (two-point) flows of trg/HOTRG_2d.py:219-348 and trg/HOTRG_3d.py:398-453.

The tensor updates run on the GPU (hotrg merges with the pure tensor's
isometry); the per-step normalisation constants are read back as scalars to
fill the two log vectors exactly as the reference does.
"""
import math

import numpy as np
import torch

from .. import hotrg, ops
from .._io import from_device, to_device


def _norm(t):
    c = ops.maxabs(t[None].contiguous())
    tb = t[None].contiguous()
    ops.normalize_(tb, c)
    return tb[0], float(c[0].item())


def _norm_tr(t):
    tb = t[None].contiguous()
    c = ops.trace_pairs(tb)
    ops.normalize_(tb, c)
    return tb[0], abs(float(c[0].item()))


def _step2d(T, imps, pairs, Dcut, direction):
    prec = hotrg.FP64 if T.dtype == torch.float64 else hotrg.FP32
    ts = [hotrg.rearrange2d(t[None].contiguous(), direction) for t in [T] + imps]
    U, _ = hotrg.projector2d(ts[0], Dcut, prec)
    return [hotrg.rearrange2d(hotrg.merge2d(ts[a], ts[b], U, U, compute=prec.merge_compute()), direction)[0]
            for a, b in pairs]


def two_point_2d(T, Timp0, Timp1, nx, ny, Dcut, XLOOPS, YLOOPS):
    (T, kind), (I0, _), (I1, _) = to_device(T), to_device(Timp0), to_device(Timp1)
    lnf = np.zeros(XLOOPS + YLOOPS + 1)
    lnf_imp = np.zeros(XLOOPS + YLOOPS + 1)
    T, c = _norm(T)
    I0, c0 = _norm(I0)
    I1, c1 = _norm(I1)
    lnf[0], lnf_imp[0] = math.log(c), math.log(c0) + math.log(c1) - 2 * math.log(c)
    cx = cy = n = 0
    while cx < XLOOPS or cy < YLOOPS:
        for d in ("Y", "X"):
            if d == "Y" and cy >= YLOOPS or d == "X" and cx >= XLOOPS:
                continue
            if d == "Y":
                cy += 1
            else:
                cx += 1
            n += 1
            two = int(nx ** 2 + ny ** 2) > 1 or ((nx == 1 and ny == 0) if d == "Y" else (nx == 0 and ny == 1))
            to_one = (nx == 0 and ny == 1) if d == "Y" else (nx == 1 and ny == 0)
            if two:
                odd = (ny % 2 == 1) if d == "Y" else (nx % 2 == 1)
                T, I0, I1 = _step2d(T, [I0, I1], [(0, 0), (1, 0), (0, 2) if odd else (2, 0)], Dcut, d)
                T, c = _norm(T)
                I0, c0 = _norm(I0)
                I1, c1 = _norm(I1)
                lnf[n], lnf_imp[n] = math.log(c) / 2 ** n, math.log(c0) + math.log(c1) - 2 * math.log(c)
            if to_one:
                T, I0 = _step2d(T, [I0, I1], [(0, 0), (1, 2)], Dcut, d)
                T, c = _norm(T)
                I0, c0 = _norm(I0)
                lnf[n], lnf_imp[n] = math.log(c) / 2 ** n, math.log(c0) - math.log(c)
            if nx == 0 and ny == 0:
                T, I0 = _step2d(T, [I0], [(0, 0), (1, 0)], Dcut, d)
                T, c = _norm(T)
                I0, c0 = _norm(I0)
                lnf[n], lnf_imp[n] = math.log(c) / 2 ** n, math.log(c0) - math.log(c)
            if d == "Y":
                ny >>= 1
            else:
                nx >>= 1
    return from_device(T, kind), from_device(I0, kind), lnf, lnf_imp


def ynearest_2d(T, Timp0, Timp1, Dcut, XLOOPS, YLOOPS):
    (T, kind), (I0, _), (I1, _) = to_device(T), to_device(Timp0), to_device(Timp1)
    lnf = np.zeros(XLOOPS + YLOOPS + 1)
    lnf_imp = np.zeros(XLOOPS + YLOOPS + 1)
    T, c = _norm(T)
    I0, c0 = _norm(I0)
    I1, c1 = _norm(I1)
    lnf[0], lnf_imp[0] = math.log(c), math.log(c0) + math.log(c1) - 2 * math.log(c)
    cx = cy = n = 0
    while cx < XLOOPS or cy < YLOOPS:
        for d in ("Y", "X"):
            if d == "Y" and cy >= YLOOPS or d == "X" and cx >= XLOOPS:
                continue
            if d == "Y":
                cy += 1
            else:
                cx += 1
            n += 1
            if n == 1:
                T, I0 = _step2d(T, [I0, I1], [(0, 0), (1, 2)], Dcut, d)
            else:
                T, I0 = _step2d(T, [I0], [(0, 0), (1, 0)], Dcut, d)
            T, c = _norm(T)
            I0, c0 = _norm(I0)
            lnf[n], lnf_imp[n] = math.log(c) / 2 ** n, math.log(c0) - math.log(c)
    return from_device(T, kind), from_device(I0, kind), lnf, lnf_imp


def _step3d(T, imps, pairs, Dcut, direction):
    prec = hotrg.FP64 if T.dtype == torch.float64 else hotrg.FP32
    ts = [hotrg.rearrange3d(t[None].contiguous(), direction) for t in [T] + imps]
    U1, U2, _, _ = hotrg.projector3d(ts[0], Dcut, prec)
    return [hotrg.restore3d(hotrg.merge3d(ts[a], ts[b], U1, U1, U2, U2, compute=prec.merge_compute()), direction)[0]
            for a, b in pairs]


def ynearest_3d(T, Timp0, Timp1, Dcut, XLOOPS, YLOOPS, ZLOOPS):
    (T, kind), (I0, _), (I1, _) = to_device(T), to_device(Timp0), to_device(Timp1)
    N = XLOOPS + YLOOPS + ZLOOPS
    lnf = np.zeros(N + 1)
    lnf_imp = np.zeros(N + 1)
    T, c = _norm_tr(T)
    I0, c0 = _norm_tr(I0)
    I1, c1 = _norm_tr(I1)
    lnf[0], lnf_imp[0] = math.log(c), math.log(c0) + math.log(c1) - 2 * math.log(c)
    cnt = {"X": 0, "Y": 0, "Z": 0}
    tot = {"X": XLOOPS, "Y": YLOOPS, "Z": ZLOOPS}
    n = 0
    while any(cnt[d] < tot[d] for d in "XYZ"):
        for d in ("Z", "Y", "X"):
            if cnt[d] >= tot[d]:
                continue
            cnt[d] += 1
            n += 1
            if n == 1 and d == "Z":
                T, I0 = _step3d(T, [I0, I1], [(0, 0), (1, 2)], Dcut, d)
            else:
                T, I0 = _step3d(T, [I0], [(0, 0), (1, 0)], Dcut, d)
            T, c = _norm_tr(T)
            I0, c0 = _norm_tr(I0)
            lnf[n], lnf_imp[n] = math.log(c) / 2 ** n, math.log(c0) - math.log(c)
    return from_device(T, kind), from_device(I0, kind), lnf, lnf_imp
