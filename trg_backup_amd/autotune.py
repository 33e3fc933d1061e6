"""GEMM tile autotuner.

Every contraction on the path is a batched strided GEMM (trg_gemm).  The
library offers several tile configurations per (storage, compute) pair; all
of them issue the same MFMA instruction and walk K in the same order, so the
choice changes speed only, never the result bits.  This module times each
configuration once per GEMM problem (dtype, compute, batch, M, N, K, operand
layout and 16-byte alignment) on the current stream and remembers the
fastest -- the role cuTENSOR's plan selection plays behind the reference's
cuquantum.contract calls (trg/HOTRG_2d.py:26-27, trg/HOTRG_3d.py:42-45,
trg/ATRG_3d_new.py:232-233).

    autotune.enable()            # tune unseen problems on first use
    autotune.save(path) / load(path)   # JSON table, portable across runs

Tuning never runs while a stream is being captured into a HIP graph; a
problem first met under capture uses the table entry if there is one and the
shape heuristic otherwise.
"""
import json
import os
import threading

import torch

from . import _lib

_STATE = {"enabled": os.environ.get("TRG_GEMM_AUTOTUNE", "0") == "1", "reps": 3}
_TABLE = {}
_LOCK = threading.Lock()


def enable(flag=True, reps=3):
    _STATE["enabled"] = bool(flag)
    _STATE["reps"] = int(reps)


def enabled():
    return _STATE["enabled"]


def active():
    """True when GEMM calls should consult the table (tuning on, or entries loaded)."""
    return _STATE["enabled"] or bool(_TABLE)


def clear():
    with _LOCK:
        _TABLE.clear()


def table():
    with _LOCK:
        return dict(_TABLE)


def config_count(dt, cc, cdt=None):
    return _lib.load().trg_gemm_config_count(dt, dt if cdt is None else cdt, cc)


def config_name(dt, cc, config, cdt=None):
    if config < 0:
        return "heuristic"
    return _lib.load().trg_gemm_config_name(dt, dt if cdt is None else cdt, cc, config).decode()


def _key(dt, cdt, cc, batch, M, N, K, amc, bnc, aligned, layout=""):
    k = f"{dt}:{cdt}:{cc}:{batch}:{M}:{N}:{K}:{int(amc)}{int(bnc)}:{int(aligned)}"
    return k + (f":{layout}" if layout else "")


def save(path):
    with _LOCK:
        data = dict(_TABLE)
    with open(path, "w") as f:
        json.dump(data, f, indent=0, sort_keys=True)


def load(path):
    with open(path) as f:
        data = json.load(f)
    with _LOCK:
        for k, v in data.items():
            _TABLE[str(k)] = int(v)


def _aligned(args):
    A, sA, rsA, csA, B, sB, rsB, csB = args
    vec = 4
    ok = lambda p, *s: p % 16 == 0 and all(x % vec == 0 for x in s)
    return ok(A, sA, rsA if rsA != 1 else csA) and ok(B, sB, rsB if rsB != 1 else csB)


def choose(dt, cc, batch, M, N, K, operands, launch, out=None, cdt=None, layout="", like=None):
    """Config for this problem (-1 = heuristic).  `launch(config, C)` runs
    the GEMM with beta = 0 into a buffer C laid out like the caller's
    destination (same strides and output index, so the epilogue under test is
    the one the call will use); it is called only while tuning, into `out`
    when the caller's output may be overwritten (beta = 0), else into a
    scratch buffer shaped like `like` (default: a (batch, M, N) contiguous
    buffer).  `layout` names the destination layout in the table key."""
    A, sA, rsA, csA, B, sB, rsB, csB = operands
    cdt = dt if cdt is None else cdt
    key = _key(dt, cdt, cc, batch, M, N, K, rsA == 1, csB == 1, _aligned(operands), layout)
    with _LOCK:
        hit = _TABLE.get(key)
    if hit is not None or not _STATE["enabled"]:
        return -1 if hit is None else hit
    if torch.cuda.is_current_stream_capturing():
        return -1
    best = _tune(dt, cdt, cc, batch, M, N, K, launch, out, like)
    with _LOCK:
        _TABLE[key] = best
    return best


def _tune(dt, cdt, cc, batch, M, N, K, launch, out, like=None):
    n = config_count(dt, cc, cdt)
    if n <= 1:
        return -1
    dtype = {_lib.F64: torch.float64, _lib.BF16: torch.bfloat16}.get(cdt, torch.float32)
    if out is not None:
        scratch = out
    elif like is not None:
        scratch = torch.empty_strided(tuple(like.size()), tuple(like.stride()), dtype=dtype, device=like.device)
    else:
        scratch = torch.empty((batch, M, N), dtype=dtype, device=torch.cuda.current_device())
    best, best_ms = -1, float("inf")
    for cfg in range(n):
        launch(cfg, scratch)                       # warm-up (and code object load)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        for _ in range(_STATE["reps"]):
            launch(cfg, scratch)
        ev[1].record()
        ev[1].synchronize()
        ms = ev[0].elapsed_time(ev[1]) / _STATE["reps"]
        if ms < best_ms:
            best, best_ms = cfg, ms
    del scratch
    return best
