"""GPU test of the allocation-free top-k eigensolver (trg_eigh_topk_f64_noalloc).

It must return what trg_eigh_topk_f64 returns with a caller workspace, bit for
bit, on two calls in a row where the second is smaller, and refuse a problem
whose workspace exceeds its fixed scratch array before launching anything.
"""
import ctypes

import numpy as np
import pytest
import torch

from trg_backup_amd import _lib, ops

pytestmark = pytest.mark.gpu


def _psd(batch, n, seed, dev):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(batch):
        X = rng.standard_normal((n, n // 2))
        d = np.exp(-np.arange(n // 2) / 7.0)
        out.append((X * d) @ X.T)
    return torch.as_tensor(np.stack(out)).to(dev)


def _noalloc(A, k, l, q, seed):
    batch, n, _ = A.shape
    w = torch.empty((batch, k), dtype=torch.float64, device=A.device)
    V = torch.empty((batch, n, k), dtype=torch.float64, device=A.device)
    eps = torch.empty(batch, dtype=torch.float64, device=A.device)
    info = torch.empty(batch, dtype=torch.int32, device=A.device)
    _lib.call("trg_eigh_topk_f64_noalloc", batch, n, k, l, q, A.data_ptr(), n * n, w.data_ptr(), V.data_ptr(),
              eps.data_ptr(), seed, info.data_ptr(), torch.cuda.current_stream().cuda_stream)
    return w, V, eps, info


@pytest.mark.parametrize("first,second", [((6, 512, 32, 64, 8), (2, 256, 16, 32, 5))])
def test_noalloc_eigh_topk_matches_workspace_form(gpu, first, second):
    lib = _lib.load()
    cap = lib.trg_eigh_topk_noalloc_max_bytes()
    assert cap == 64 << 20
    for i, (batch, n, k, l, q) in enumerate((first, second)):
        print(f"noalloc eigh_topk call {i}: batch={batch} n={n} k={k} l={l} q={q}", flush=True)
        assert lib.trg_eigh_topk_workspace(batch, n, l) <= cap
        A = _psd(batch, n, 100 + i, gpu)
        ref = ops.eigh_topk(A, k, l, q, seed=20240 + i, batch=batch)
        got = _noalloc(A, k, l, q, 20240 + i)
        torch.cuda.synchronize()
        for a, b in zip(ref, got):
            assert torch.equal(a, b)
        assert int(got[3].abs().sum()) == 0
        # and against LAPACK, so the shared result is also right
        e = np.linalg.eigvalsh(A.cpu().numpy())[:, -k:]
        assert np.allclose(got[0].cpu().numpy(), e, rtol=1e-10, atol=1e-12 * e.max())


def test_noalloc_eigh_topk_refuses_oversize(gpu):
    lib = _lib.load()
    cap = lib.trg_eigh_topk_noalloc_max_bytes()
    batch, n, l = 64, 4096, 128
    assert lib.trg_eigh_topk_workspace(batch, n, l) > cap
    dummy = torch.empty(1, dtype=torch.float64, device=gpu)
    p = dummy.data_ptr()
    st = lib.trg_eigh_topk_f64_noalloc(batch, n, 64, l, 4, ctypes.c_void_p(p), n * n, p, p, p, 1, None,
                                       ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert st == 3   # TRG_OOM, nothing launched
    with pytest.raises(_lib.TrgError, match="workspace"):
        _lib.call("trg_eigh_topk_f64_noalloc", batch, n, 64, l, 4, p, n * n, p, p, p, 1, None,
                  torch.cuda.current_stream().cuda_stream)
