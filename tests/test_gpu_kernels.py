"""GPU tests of the individual libtrg kernels (through the C ABI via ops)."""
import numpy as np
import pytest
import torch

from oracle import philox

pytestmark = pytest.mark.gpu


def _dev(x, gpu):
    return torch.as_tensor(np.ascontiguousarray(x)).to(gpu)


def test_device_info(gpu):
    import ctypes
    from trg_backup_amd import _lib
    lib = _lib.load()
    cus, lds, ws = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    assert lib.trg_device_info(0, ctypes.byref(cus), ctypes.byref(lds), ctypes.byref(ws)) == 0
    print(f"\n[device] CUs={cus.value} LDS/block={lds.value} wave={ws.value}", flush=True)
    assert ws.value == 64


@pytest.mark.parametrize("dtype,compute,rtol", [(torch.float64, None, 1e-13), (torch.float32, None, 2e-6),
                                                (torch.float32, "bf16", 2e-2)])
@pytest.mark.parametrize("shape", [(1, 1, 1), (17, 33, 9), (128, 128, 16), (300, 257, 130), (64, 1000, 200)])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, False), (False, True), (True, True)])
def test_gemm_layouts(gpu, dtype, compute, rtol, shape, ta, tb):
    from trg_backup_amd import ops
    M, N, K = shape
    rng = np.random.default_rng(M * 7 + N * 3 + K)
    A = rng.standard_normal((K, M) if ta else (M, K))
    B = rng.standard_normal((N, K) if tb else (K, N))
    Ad, Bd = _dev(A, gpu).to(dtype), _dev(B, gpu).to(dtype)
    Av = Ad.t() if ta else Ad
    Bv = Bd.t() if tb else Bd
    C = ops.gemm(Av, Bv, compute=compute)
    ref = (A.T if ta else A) @ (B.T if tb else B)
    err = np.abs(C.double().cpu().numpy() - ref).max() / (np.abs(A).max() * np.abs(B).max() * K)
    assert err < rtol, err


def test_gemm_batched_alpha_beta(gpu):
    from trg_backup_amd import ops
    rng = np.random.default_rng(0)
    A = rng.standard_normal((5, 70, 40))
    B = rng.standard_normal((40, 90))
    C0 = rng.standard_normal((5, 70, 90))
    Cd = _dev(C0, gpu)
    ops.gemm(_dev(A, gpu), _dev(B, gpu)[None], Cd, alpha=0.5, beta=-2.0)
    ref = 0.5 * A @ B - 2.0 * C0
    assert np.allclose(Cd.cpu().numpy(), ref, atol=1e-11)


def test_copy_permute_cast(gpu):
    from trg_backup_amd import ops
    x = np.random.default_rng(1).standard_normal((3, 4, 5, 6, 7))
    xd = _dev(x, gpu)
    y = ops.permute(xd, (4, 2, 0, 3, 1))
    assert np.array_equal(y.cpu().numpy(), np.transpose(x, (4, 2, 0, 3, 1)))
    z = ops.cast(xd, torch.float32)
    assert np.array_equal(z.cpu().numpy(), x.astype(np.float32))
    w = ops.permute(z, (1, 0, 2, 4, 3), dtype=torch.float64)
    assert np.array_equal(w.cpu().numpy(), np.transpose(x.astype(np.float32), (1, 0, 2, 4, 3)).astype(np.float64))


def test_maxabs_trace_normalize_lnz(gpu):
    from trg_backup_amd import ops
    rng = np.random.default_rng(2)
    T = rng.standard_normal((3, 6, 6, 6, 6))
    T[1, 2, 3, 4, 5] = -50.0
    Td = _dev(T, gpu)
    c = ops.maxabs(Td, batch=3).cpu().numpy()
    assert np.array_equal(c, np.abs(T).reshape(3, -1).max(axis=1))
    tr = ops.trace_pairs(Td, batch=3).cpu().numpy()
    assert np.allclose(tr, np.einsum("baabb->b", T.reshape(3, 6, 6, 6, 6)) if False else
                       np.array([np.einsum("aabb", T[b]) for b in range(3)]), atol=1e-12)
    lnf = torch.zeros((3, 4), dtype=torch.float64, device=gpu)
    ops.normalize_(Td, ops.maxabs(Td, batch=3), lnf, step=2, batch=3, lnf_stride=4)
    assert np.allclose(Td.cpu().numpy(), T / c[:, None, None, None, None], rtol=1e-15)
    assert np.allclose(lnf.cpu().numpy()[:, 2], np.log(c) / 4)
    out = ops.lnz(lnf, _dev(np.array([2.0, 3.0, 4.0]), gpu), 8.0, batch=3).cpu().numpy()
    assert np.allclose(out, np.log(c) / 4 + np.log([2.0, 3.0, 4.0]) / 8)


@pytest.mark.parametrize("n", [1, 2, 7, 32, 64, 128, 200])
def test_syevj_matches_lapack(gpu, n):
    from trg_backup_amd import ops
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, n))
    A = X @ X.T * np.exp(-np.arange(n) / 6.0)[None, :]
    A = 0.5 * (A + A.T)
    w, V, info = ops.syevj(_dev(np.stack([A, 2 * A]), gpu), batch=2)
    w, V = w.cpu().numpy(), V.cpu().numpy()
    we = np.linalg.eigvalsh(A)
    assert np.allclose(w[0], we, atol=1e-13 * np.abs(we).max())
    assert np.allclose(w[1], 2 * we, atol=2e-13 * np.abs(we).max())
    R = A @ V[0] - V[0] * w[0][None, :]
    assert np.abs(R).max() < 1e-12 * np.abs(we).max()
    assert np.allclose(V[0].T @ V[0], np.eye(n), atol=1e-13)
    assert int(info.cpu().numpy().max()) == 0


def test_chol_inv(gpu):
    from trg_backup_amd import ops
    rng = np.random.default_rng(5)
    X = rng.standard_normal((300, 96))
    G = X.T @ X
    R, info = ops.chol_inv(_dev(G, gpu))
    R = R[0].cpu().numpy()
    Q = X @ R
    assert np.allclose(Q.T @ Q, np.eye(96), atol=1e-12)
    assert np.allclose(np.tril(R, -1), 0)
    assert int(info.cpu().numpy()[0]) == 0


def test_gaussian_matches_oracle_philox(gpu):
    from trg_backup_amd import ops
    g = ops.gaussian((4, 1000), seed=1, device=gpu).cpu().numpy()
    ref = philox.standard_normal((4, 1000), seed=1)
    assert np.allclose(g, ref, rtol=1e-13, atol=1e-13)
    g2 = ops.gaussian((100,), seed=123456789012345, dtype=torch.float32, device=gpu, stream_id=3).cpu().numpy()
    assert np.allclose(g2, philox.gaussian(100, 123456789012345, stream=3), rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("n,k", [(64, 16), (256, 16), (1024, 32), (4096, 64)])
def test_eigh_topk(gpu, n, k):
    from trg_backup_amd import ops
    rng = np.random.default_rng(n + k)
    lam = np.exp(-np.arange(n) / (k / 3.0))
    Qr, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = (Qr * lam[None, :]) @ Qr.T
    A = 0.5 * (A + A.T)
    l = min(n, 2 * k)
    w, V, eps, info = ops.eigh_topk(_dev(A, gpu), k, l, 0 if l == n else 12)
    w, V, eps = w[0].cpu().numpy(), V[0].cpu().numpy(), float(eps[0].cpu().numpy())
    assert np.allclose(w, lam[:k][::-1], rtol=1e-10, atol=1e-14)
    P = V @ V.T
    P0 = Qr[:, :k] @ Qr[:, :k].T
    assert np.linalg.norm(P - P0) < 1e-9
    assert abs(eps - lam[k:].sum()) < 1e-10 * lam.sum()
    assert int(info.cpu().numpy()[0]) == 0


@pytest.mark.parametrize("perm", [(0, 1, 2, 3), (0, 2, 1, 3), (3, 1, 2, 0), (0, 1, 3, 2), (2, 3, 0, 1), (1, 3, 0, 2)])
def test_permute_fast_paths(gpu, perm):
    """Row-copy and tiled-transpose paths of trg_copy_strided (odd sizes, tails, casts)."""
    from trg_backup_amd import ops
    x = np.random.default_rng(7).standard_normal((5, 33, 40, 17 * 3))
    xd = _dev(x, gpu)
    y = ops.permute(xd, perm)
    assert np.array_equal(y.cpu().numpy(), np.transpose(x, perm))
    y32 = ops.permute(xd, perm, dtype=torch.float32)
    assert np.array_equal(y32.cpu().numpy(), np.transpose(x, perm).astype(np.float32))
    # strided (sliced) input view
    xs = xd[:, 3:30, ::2, :]
    ys = ops.permute(xs, perm)
    assert np.array_equal(ys.cpu().numpy(), np.transpose(x[:, 3:30, ::2, :], perm))


@pytest.mark.parametrize("dtype,compute", [(torch.float32, None), (torch.float32, "bf16"), (torch.float32, "bf16x3"),
                                           (torch.float32, "bf16x6"), (torch.float64, None)])
@pytest.mark.parametrize("shape,ta,tb", [((300, 260, 100), False, False), ((257, 129, 67), True, False),
                                         ((96, 200, 48), False, True), ((64, 4096, 64), True, True),
                                         ((300, 260, 100), False, True), ((130, 70, 67), False, True),
                                         ((520, 264, 1024), False, True)])
def test_gemm_tile_configs_bitwise_equal(gpu, dtype, compute, shape, ta, tb):
    """Every tile configuration gives the same bits as the default dispatch
    (same MFMA instruction, same K order) and agrees with a float64 product."""
    from trg_backup_amd import _lib, autotune, ops
    M, N, K = shape
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((3, K, M) if ta else (3, M, K))
    B = rng.standard_normal((3, N, K) if tb else (3, K, N))
    Ad, Bd = _dev(A, gpu).to(dtype), _dev(B, gpu).to(dtype)
    Av = Ad.transpose(1, 2) if ta else Ad
    Bv = Bd.transpose(1, 2) if tb else Bd
    C0 = ops.gemm(Av, Bv, compute=compute)
    ref = np.matmul(A.transpose(0, 2, 1) if ta else A, B.transpose(0, 2, 1) if tb else B)
    err = np.abs(C0.double().cpu().numpy() - ref).max() / (np.abs(A).max() * np.abs(B).max() * K)
    tol = {"bf16": 2e-3, "bf16x3": 2e-5}.get(compute, 1e-6 if dtype == torch.float32 else 1e-14)
    assert err < tol, err
    dt = ops.dtype_code(Ad)
    cc = {"bf16": _lib.BF16, "bf16x3": _lib.BF16X3, "bf16x6": _lib.BF16X6}.get(compute, dt)
    n = autotune.config_count(dt, cc)
    assert n >= 2
    for cfg in range(n):
        C = torch.empty_like(C0)
        _lib.call("trg_gemm_ex", cfg, dt, dt, cc, 3, M, N, K, 1.0, Av.data_ptr(), Av.stride(0), Av.stride(1),
                  Av.stride(2), Bv.data_ptr(), Bv.stride(0), Bv.stride(1), Bv.stride(2), 0.0, C.data_ptr(),
                  M * N, N, 1, 0, 0, 0, 0, ops.stream_ptr())
        assert torch.equal(C, C0), (cfg, autotune.config_name(dt, cc, cfg))


def test_gemm_autotuner_picks_a_config_and_keeps_bits(gpu, tmp_path):
    from trg_backup_amd import autotune, ops
    rng = np.random.default_rng(5)
    A = _dev(rng.standard_normal((4, 512, 256)), gpu).float()
    B = _dev(rng.standard_normal((4, 256, 384)), gpu).float()
    C0 = ops.gemm(A, B)
    autotune.clear()
    autotune.enable(True, reps=2)
    try:
        C1 = ops.gemm(A, B)
        C2 = ops.gemm(A, B)                      # table hit
    finally:
        autotune.enable(False)
    tab = autotune.table()
    assert len(tab) == 1 and 0 <= next(iter(tab.values())) < autotune.config_count(0, 0)
    assert torch.equal(C1, C0) and torch.equal(C2, C0)
    p = tmp_path / "tune.json"
    autotune.save(str(p))
    autotune.clear()
    autotune.load(str(p))
    assert autotune.table() == tab and autotune.active()
    assert torch.equal(ops.gemm(A, B), C0)
    autotune.clear()


def test_bf16x3_accuracy_between_bf16_and_fp32(gpu):
    """Split-bf16 products: error far below plain bf16 and within ~2^-16 of the float64 product."""
    from trg_backup_amd import ops
    rng = np.random.default_rng(11)
    A = rng.standard_normal((2, 256, 512))
    B = rng.standard_normal((2, 512, 192))
    Ad, Bd = _dev(A, gpu).float(), _dev(B, gpu).float()
    Bt = ops.contiguous(Bd.transpose(1, 2)).transpose(1, 2)        # K-contiguous B
    ref = A @ B
    scale = np.abs(ref).max()
    e = {c: np.abs(ops.gemm(Ad, Bt, compute=c).double().cpu().numpy() - ref).max() / scale
         for c in (None, "bf16x3", "bf16")}
    print(f"\n[bf16x3] rel err fp32 {e[None]:.2e} bf16x3 {e['bf16x3']:.2e} bf16 {e['bf16']:.2e}", flush=True)
    assert e["bf16x3"] < 5e-5 and e["bf16x3"] < e["bf16"] / 50 and e[None] < e["bf16x3"]


@pytest.mark.parametrize("compute", [None, "bf16", "bf16x3", "bf16x6"])
def test_gemm_two_level_output_index(gpu, compute):
    """A product written through a two-level output index equals the plain product
    permuted into the same layout, bit for bit (same kernel, only the epilogue
    addresses differ); beta accumulation reads through the same index."""
    from trg_backup_amd import ops
    rng = np.random.default_rng(21)
    B, M1, M2, N1, N2, K = 3, 6, 16, 8, 12, 40          # rows (M1 | M2), columns (N1 | N2)
    A = _dev(rng.standard_normal((B, M1 * M2, K)), gpu).float()
    Bm = _dev(rng.standard_normal((B, K, N1 * N2)), gpu).float()
    plain = ops.gemm(A, Bm, compute=compute)                           # [(m1,m2),(n1,n2)]
    want = ops.contiguous(plain.view(B, M1, M2, N1, N2).permute(0, 2, 3, 1, 4))   # [m2,n1,m1,n2]
    out = torch.empty(B * M2 * N1 * M1 * N2, dtype=torch.float32, device=gpu)
    lay = ([(M1, N2), (M2, N1 * M1 * N2)], [(N1, M1 * N2), (N2, 1)])
    ops.gemm(A, Bm, out, compute=compute, out_layout=lay)
    assert torch.equal(out.view(B, M2, N1, M1, N2), want)
    ops.gemm(A, Bm, out, beta=1.0, alpha=-1.0, compute=compute, out_layout=lay)
    assert float(out.abs().max()) == 0.0


def test_bf16_storage_gemms_match_fp32_storage_bits(gpu):
    """bf16 operands in memory give the bits of the fp32-storage bf16 GEMM (same rounding,
    once, by the producer); a bf16 product is the fp32-accumulated product rounded once;
    fp32 <-> bf16 copies round to nearest even like torch."""
    from trg_backup_amd import autotune, ops
    rng = np.random.default_rng(33)
    B, M, N, K = 3, 200, 136, 96
    A = _dev(rng.standard_normal((B, M, K)), gpu).float()                   # K-contiguous
    Bt = _dev(rng.standard_normal((B, N, K)), gpu).float()                  # B(k,n) at n*K + k
    Ab, Bb = ops.cast(A, torch.bfloat16), ops.cast(Bt, torch.bfloat16)
    assert torch.equal(Ab, A.to(torch.bfloat16)) and torch.equal(ops.cast(Ab, torch.float32), Ab.float())
    ref = ops.gemm(A, Bt.transpose(1, 2), compute="bf16")
    for cfg in range(autotune.config_count(_lib_code("BF16"), _lib_code("BF16"), _lib_code("F32"))):
        autotune.clear()
        C = torch.empty_like(ref)
        _gemm_cfg(cfg, Ab, Bb.transpose(1, 2), C)
        assert torch.equal(C, ref), cfg
    Cb = torch.empty((B, M, N), dtype=torch.bfloat16, device=gpu)
    ops.gemm(A, Bt.transpose(1, 2), Cb, compute="bf16")
    assert torch.equal(Cb, ref.to(torch.bfloat16))


def _lib_code(name):
    from trg_backup_amd import _lib
    return getattr(_lib, name)


def _gemm_cfg(cfg, A, B, C):
    from trg_backup_amd import _lib, ops
    bt, M, K = A.shape
    N = B.shape[2]
    _lib.call("trg_gemm_ex", cfg, ops.dtype_code(A), ops.dtype_code(C), _lib.BF16, bt, M, N, K, 1.0,
              A.data_ptr(), A.stride(0), A.stride(1), A.stride(2), B.data_ptr(), B.stride(0), B.stride(1),
              B.stride(2), 0.0, C.data_ptr(), M * N, N, 1, 0, 0, 0, 0, ops.stream_ptr())


@pytest.mark.parametrize("cc,cdt", [(None, torch.float32), ("bf16", torch.float32), ("bf16x3", torch.float32),
                                    ("bf16x6", torch.float32), ("bf16", torch.bfloat16)])
def test_gemm_row_contiguous_output_vector_epilogue(gpu, cc, cdt):
    """A destination whose rows are contiguous (rsC = 1) takes the 4-row vector
    epilogue; it equals, bit for bit, the same product written row-major
    (scalar epilogue), with an M that is not a multiple of 4 (edge groups fall
    back to scalars) and with beta != 0 (read-modify-write)."""
    from trg_backup_amd import ops
    g = torch.Generator(device="cpu").manual_seed(11)
    B, M, N, K = 3, 70, 45, 37
    A = torch.randn(B, M, K, generator=g).to(gpu)
    Bm = torch.randn(B, K, N, generator=g).to(gpu)
    C0 = torch.randn(B, M, N, generator=g).to(cdt).to(gpu)
    Cv_buf = torch.empty(B, N, 72, dtype=cdt, device=gpu)      # rows padded to a multiple of 4
    Cv = Cv_buf[:, :, :M].transpose(1, 2)                        # (B, M, N) view, row stride 1
    Cv.copy_(C0)
    Cs = C0.clone()                                              # row-major: scalar epilogue
    assert Cv.stride(1) == 1 and Cv.stride(2) % 4 == 0
    ops.gemm(A, Bm, Cv, alpha=1.5, beta=0.5, compute=cc)
    ops.gemm(A, Bm, Cs, alpha=1.5, beta=0.5, compute=cc)
    torch.cuda.synchronize()
    assert torch.equal(Cv.contiguous(), Cs)
    ref = 1.5 * torch.matmul(A.double(), Bm.double()) + 0.5 * C0.double()
    tol = 1e-4 if (cc is None and cdt == torch.float32) else 5e-2
    assert (Cs.double() - ref).abs().max().item() < tol * ref.abs().max().item()


def test_bf16x6_accuracy_matches_fp32(gpu):
    """Three-piece split-bf16 products carry the 24 bits of fp32: the error
    against the float64 product is within a small factor of the fp32 GEMM's
    and far below the two-piece split.  K is short (32), so the operands'
    representation, not the fp32 accumulation, sets the error."""
    from trg_backup_amd import ops
    rng = np.random.default_rng(12)
    A = rng.standard_normal((2, 256, 32))
    B = rng.standard_normal((2, 32, 192))
    Ad, Bd = _dev(A, gpu).float(), _dev(B, gpu).float()
    Bt = ops.contiguous(Bd.transpose(1, 2)).transpose(1, 2)        # K-contiguous B
    ref = A.astype(np.float32).astype(np.float64) @ B.astype(np.float32).astype(np.float64)
    scale = np.abs(ref).max()
    e = {c: np.abs(ops.gemm(Ad, Bt, compute=c).double().cpu().numpy() - ref).max() / scale
         for c in (None, "bf16x6", "bf16x3")}
    print(f"\n[bf16x6] rel err fp32 {e[None]:.2e} bf16x6 {e['bf16x6']:.2e} bf16x3 {e['bf16x3']:.2e}", flush=True)
    assert e["bf16x6"] < 4 * e[None] and e["bf16x6"] < e["bf16x3"] / 10
