"""GPU tests: HIP-graph replays of steady-state batched HOTRG steps equal eager steps."""
import time

import numpy as np
import pytest
import torch

from oracle import models

pytestmark = pytest.mark.gpu


def _steady(B, chi, gpu, prec):
    from trg_backup_amd import hotrg, ops
    T = torch.stack([torch.as_tensor(models.ising2d_tensor(0.40 + 0.01 * b)) for b in range(B)]).to(gpu)
    T = ops.contiguous(T)
    lnf = ops.fill_(torch.empty((B, 24), dtype=torch.float64, device=gpu), 0.0)
    hotrg.normalize_maxabs(T, lnf, 0)
    n, d = 0, "Y"
    while T.shape[1] < chi or T.shape[3] < chi:
        n += 1
        T = hotrg.step2d(T, chi, d, hotrg.FP64)
        hotrg.normalize_maxabs(T, lnf, n)
        d = "X" if d == "Y" else "Y"
    return ops.cast(T, prec.store), lnf, n, d


@pytest.mark.parametrize("precname,shared", [("FP64", False), ("FP32", False), ("FP32", True)])
def test_graph_replays_equal_eager_steps(gpu, precname, shared):
    """Replays equal eager steps, also when the Y and X graphs share one
    workspace pool."""
    from trg_backup_amd import graphs, hotrg, ops
    prec = getattr(hotrg, precname)
    B, chi = 3, 16
    T0, lnf0, n0, d0 = _steady(B, chi, gpu, prec)
    lnf_e = ops.contiguous(lnf0)
    lnf_g = ops.contiguous(lnf0)
    counter = graphs.step_counter(gpu)
    pool = torch.cuda.graph_pool_handle() if shared else None
    gY = graphs.GraphedStep2D(B, chi, "Y", prec, lnf_g, counter, pool=pool)
    gX = graphs.GraphedStep2D(B, chi, "X", prec, lnf_g, counter, pool=pool)
    graphs.set_step(counter, n0 + 1)
    Te, Tg, d = ops.contiguous(T0), ops.contiguous(T0), d0
    for n in range(n0 + 1, n0 + 7):
        Te = hotrg.step2d(Te, chi, d, prec)
        hotrg.normalize_maxabs(Te, lnf_e, n)
        Tg = (gY if d == "Y" else gX)(Tg)
        Tg = ops.contiguous(Tg)                       # eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(Te, Tg), f"replay of step {n} differs"
        assert torch.equal(lnf_e[:, n], lnf_g[:, n])
        d = "X" if d == "Y" else "Y"


def test_graph_replay_not_slower_than_eager(gpu):
    from trg_backup_amd import graphs, hotrg, ops
    B, chi = 1, 16
    T0, lnf, n0, d0 = _steady(B, chi, gpu, hotrg.FP32)
    counter = graphs.step_counter(gpu)
    g = {d: graphs.GraphedStep2D(B, chi, d, hotrg.FP32, lnf, counter) for d in "YX"}
    T = ops.contiguous(T0)
    torch.cuda.synchronize()
    t = time.perf_counter()
    d = d0
    for n in range(n0 + 1, n0 + 11):
        T = hotrg.step2d(T, chi, d, hotrg.FP32)
        hotrg.normalize_maxabs(T, lnf, n)
        d = "X" if d == "Y" else "Y"
    torch.cuda.synchronize()
    te = (time.perf_counter() - t) / 10
    T, d = ops.contiguous(T0), d0
    graphs.set_step(counter, n0 + 1)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for n in range(n0 + 1, n0 + 11):
        T = g[d](T)
        d = "X" if d == "Y" else "Y"
    torch.cuda.synchronize()
    tg = (time.perf_counter() - t) / 10
    print(f"\n[graph] chi=16 B=1 step: eager {te * 1e3:.2f} ms, graph {tg * 1e3:.2f} ms ({te / tg:.1f}x)", flush=True)
    assert tg < 1.25 * te


def test_graph_replays_equal_eager_steps_3d(gpu):
    """Captured steady-state batched 3D HOTRG steps (Z, Y, X) replay bit for bit like eager steps."""
    from trg_backup_amd import graphs, hotrg, ops
    B, chi = 2, 6
    T = torch.stack([torch.as_tensor(models.ising3d_tensor(b)) for b in (0.21, 0.23)]).to(gpu)
    lnf = ops.fill_(torch.empty((B, 16), dtype=torch.float64, device=gpu), 0.0)
    T = ops.contiguous(T)
    hotrg.normalize_trace(T, lnf, 0)
    for n, d in enumerate("ZYX", 1):                    # grow to the steady shape in float64
        T = hotrg.step3d(T, chi, d, hotrg.FP64)
        hotrg.normalize_trace(T, lnf, n)
    T = ops.cast(T, torch.float32)
    lnf_e, lnf_g = ops.contiguous(lnf), ops.contiguous(lnf)
    counter = graphs.step_counter(gpu)
    g = {d: graphs.GraphedStep3D(B, chi, d, hotrg.FP32, lnf_g, counter) for d in "ZYX"}
    graphs.set_step(counter, 4)
    Te, Tg = ops.contiguous(T), ops.contiguous(T)
    for n in range(4, 10):
        d = "ZYX"[(n - 1) % 3]
        Te = hotrg.step3d(Te, chi, d, hotrg.FP32)
        hotrg.normalize_trace(Te, lnf_e, n)
        Tg = ops.contiguous(g[d](Tg))                   # eager work between replays
        torch.cuda.synchronize()
        assert torch.equal(Te, Tg), f"replay of step {n} differs"
        assert torch.equal(lnf_e[:, n], lnf_g[:, n])
