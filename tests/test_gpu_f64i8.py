"""Float64 GEMM on the int8 MFMA (trg_gemm_f64i8): the CPU model against the
mathematics, the device against the model bit for bit, and the 3D HOTRG
merges of the float64 steps against the fp64 MFMA."""
import numpy as np
import pytest
import torch

from f64i8_model import EXP_NONFINITE, gemm_f64i8, k_chunk, slices


def _heavy(rng, shape):
    """Entries spanning many orders of magnitude within each row."""
    return rng.standard_normal(shape) * np.exp(3.0 * rng.standard_normal(shape))


def _bound(A, B, S):
    """Worst case of the scheme per entry: the slices' remainders (below 2^-7S
    of the row / column maximum) and the dropped pairs i + j > S + 1 (together
    below (S + 1) 2^-7S), times K and the scales 2^(e_m + f_n)."""
    _, ea = slices(A, S)
    _, eb = slices(np.ascontiguousarray(B.T), S)
    K = A.shape[1]
    return (S + 3) * K * 2.0 ** (-7 * S) * np.ldexp(1.0, ea[:, None] + eb[None, :])


# ---------------------------------------------------------------------------
# CPU: the model
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("S", [4, 5, 6, 7])
def test_model_within_the_scheme_bound(S):
    rng = np.random.default_rng(S)
    A, B = _heavy(rng, (40, 300)), _heavy(rng, (300, 30))
    err = np.abs(gemm_f64i8(A, B, S) - A @ B)
    assert np.all(err <= _bound(A, B, S) + 1e-15 * np.abs(A @ B))
    # the worst case is loose by about K, but not by 2^14: two slices fewer miss it
    assert np.any(np.abs(gemm_f64i8(A, B, S - 2) - A @ B) > _bound(A, B, S))


def _exact_operands(rng, K, a_slices, b_slices):
    """Integers of 7 a_slices bits (rows of A) and 7 b_slices bits (columns of
    B), each row / column reaching the top bit, so that a_slices and b_slices
    slices hold them exactly."""
    ta, tb = 7 * a_slices, 7 * b_slices
    A = rng.integers(-(2 ** ta - 1), 2 ** ta, (20, K))
    B = rng.integers(-(2 ** tb - 1), 2 ** tb, (K, 15))
    A[:, 0], B[0, :] = 2 ** (ta - 1) + 3, -(2 ** (tb - 1) + 5)
    return A, B


def test_model_exact_when_the_kept_pairs_cover_the_operands():
    """With 3-slice rows of A and 4-slice columns of B every nonzero pair has
    i + j <= 7 = S + 1 for S = 6, so nothing is dropped: the model returns the
    integer product exactly (K = 8 keeps it inside the 53 bits of a double).
    The same check fails when the model drops its last diagonal."""
    rng = np.random.default_rng(1)
    A, B = _exact_operands(rng, 8, 3, 4)
    exact = (A @ B).astype(float)
    assert np.array_equal(gemm_f64i8(A.astype(float), B.astype(float), 6), exact)
    # one more bit in B's columns needs a fifth slice, whose pairs i + j = 8 are dropped
    A, B = _exact_operands(rng, 8, 3, 5)
    assert not np.array_equal(gemm_f64i8(A.astype(float), B.astype(float), 6), (A @ B).astype(float))


def test_model_exact_on_seven_bit_integers():
    """Integers below 128 whose rows and columns reach 64 or more are one
    slice each (e = 7): the products are then exact integer sums."""
    rng = np.random.default_rng(1)
    A = rng.integers(-127, 128, (33, 200)).astype(float)
    B = rng.integers(-127, 128, (200, 21)).astype(float)
    A[:, 0], B[0, :] = 100.0, -90.0
    assert np.array_equal(gemm_f64i8(A, B, 6), A @ B)


def test_model_slices_reassemble_the_rows():
    rng = np.random.default_rng(2)
    X = _heavy(rng, (7, 50))
    q, e = slices(X, 7)
    back = sum(np.ldexp(qi.astype(float), -7 * (i + 1)) for i, qi in enumerate(q))
    assert np.all(np.abs(np.ldexp(back, e[:, None]) - X) <= np.ldexp(1.0, e - 49)[:, None])
    assert all(np.abs(qi).max() <= 127 for qi in q)


def test_model_nonfinite_rows_give_nan():
    A = np.ones((3, 4))
    B = np.ones((4, 2))
    A[1, 2] = np.inf
    B[3, 0] = np.nan
    C = gemm_f64i8(A, B, 6)
    assert np.isnan(C[1]).all() and np.isnan(C[:, 0]).all()
    assert C[0, 1] == 4.0 and C[2, 1] == 4.0
    assert slices(A, 6)[1][1] == EXP_NONFINITE


# ---------------------------------------------------------------------------
# GPU: the device against the model
# ---------------------------------------------------------------------------
def _dev(x, gpu):
    return torch.as_tensor(np.ascontiguousarray(x)).to(gpu)


@pytest.fixture
def any_size(monkeypatch):
    """Route every float64 product with compute="f64i8" to the int8 kernels."""
    from trg_backup_amd import ops
    monkeypatch.setattr(ops, "I8_MIN_K", 1)
    return ops


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(1, 1, 1), (37, 70, 19), (256, 256, 256), (300, 129, 513), (640, 384, 4096)])
@pytest.mark.parametrize("ta,tb", [(False, False), (True, True), (False, True)])
def test_device_equals_model_bitwise(gpu, any_size, shape, ta, tb):
    M, N, K = shape
    print(f"\n[f64i8] shape {shape} ta={ta} tb={tb}", flush=True)
    rng = np.random.default_rng(M + 3 * N + 7 * K)
    A, B = _heavy(rng, (M, K)), _heavy(rng, (K, N))
    Ad = _dev(A.T, gpu).t() if ta else _dev(A, gpu)
    Bd = _dev(B.T, gpu).t() if tb else _dev(B, gpu)
    C = any_size.gemm(Ad, Bd, compute="f64i8").cpu().numpy()
    ref = gemm_f64i8(A, B, any_size.I8_SLICES)
    assert np.array_equal(C, ref), np.abs(C - ref).max()
    assert np.all(np.abs(C - A @ B) <= _bound(A, B, any_size.I8_SLICES) + 1e-15 * np.abs(A @ B))


@pytest.mark.gpu
def test_device_exact_when_the_kept_pairs_cover_the_operands(gpu, any_size):
    rng = np.random.default_rng(3)
    A, B = _exact_operands(rng, 8, 3, 4)
    C = any_size.gemm(_dev(A.astype(float), gpu), _dev(B.astype(float), gpu), compute="f64i8")
    assert np.array_equal(C.cpu().numpy(), (A @ B).astype(float))


@pytest.mark.gpu
def test_device_batched_shared_operand_alpha_beta(gpu, any_size):
    rng = np.random.default_rng(5)
    A, B, C0 = _heavy(rng, (3, 200, 300)), _heavy(rng, (300, 150)), rng.standard_normal((3, 200, 150))
    Cd = _dev(C0, gpu)
    any_size.gemm(_dev(A, gpu), _dev(B, gpu).unsqueeze(0).expand(3, -1, -1), Cd, alpha=0.75, beta=-1.5,
                  compute="f64i8")
    for b in range(3):
        ref = gemm_f64i8(A[b], B, 6, alpha=0.75) - 1.5 * C0[b]
        assert np.allclose(Cd[b].cpu().numpy(), ref, rtol=1e-15, atol=1e-15 * np.abs(ref).max())


@pytest.mark.gpu
def test_device_two_level_output_index(gpu, any_size):
    """The product written through out_layout equals the plain product, moved."""
    rng = np.random.default_rng(6)
    M, N, K = 6 * 64, 4 * 80, 260
    A, B = _heavy(rng, (M, K)), _heavy(rng, (K, N))
    plain = any_size.gemm(_dev(A, gpu), _dev(B, gpu), compute="f64i8").cpu().numpy()
    out = torch.empty(M * N, dtype=torch.float64, device=gpu)
    # rows (m1 | m0) = (6 | 64), columns (n1 | n0) = (4 | 80), stored [n1, m1, n0, m0]
    any_size.gemm(_dev(A, gpu), _dev(B, gpu), out, compute="f64i8",
                  out_layout=([(6, 80 * 64), (64, 1)], [(4, 6 * 80 * 64), (80, 64)]))
    moved = out.cpu().numpy().reshape(4, 6, 80, 64).transpose(1, 3, 0, 2).reshape(M, N)
    assert np.array_equal(moved, plain)


@pytest.mark.gpu
def test_device_nonfinite_and_empty(gpu, any_size):
    A = np.ones((70, 90))
    B = np.ones((90, 65))
    A[3, 5] = np.nan
    B[7, 64] = np.inf
    C = any_size.gemm(_dev(A, gpu), _dev(B, gpu), compute="f64i8").cpu().numpy()
    assert np.isnan(C[3]).all() and np.isnan(C[:, 64]).all()
    ok = np.ones_like(C, bool)
    ok[3], ok[:, 64] = False, False
    assert np.all(C[ok] == 90.0)
    # K = 0: C = beta C, through the int8 path's empty pass
    any_size.I8_MIN_K = 0
    C0 = _dev(np.full((5, 6), 2.0), gpu)
    any_size.gemm(torch.empty((5, 0), dtype=torch.float64, device=gpu),
                  torch.empty((0, 6), dtype=torch.float64, device=gpu), C0, beta=0.5, compute="f64i8")
    assert np.all(C0.cpu().numpy() == 1.0)


@pytest.mark.gpu
def test_device_k_beyond_one_chunk(gpu, any_size):
    """K past the int32 range of one pass runs in chunks added in fp64."""
    S = any_size.I8_SLICES
    K = k_chunk(S) + 200
    rng = np.random.default_rng(7)
    A, B = _heavy(rng, (64, K)), _heavy(rng, (K, 70))
    C = any_size.gemm(_dev(A, gpu), _dev(B, gpu), compute="f64i8").cpu().numpy()
    kc = k_chunk(S)
    bound = _bound(A[:, :kc], B[:kc], S) + _bound(A[:, kc:], B[kc:], S)
    assert np.all(np.abs(C - A @ B) <= bound + 1e-14 * np.abs(A @ B))


@pytest.mark.gpu
def test_hotrg3d_int8_merges_match_fp64_merges(gpu):
    """3D Ising chi = 8, nine steps: every float64 merge on the int8 MFMA
    against the fp64 MFMA (the oracle's chi = 8 solve with this product in
    the merges of steps 4-6, run in NumPy, lands within 5e-12 of float64)."""
    from oracle import models
    from trg_backup_amd import hotrg
    T0 = _dev(models.ising3d_tensor(0.2387), gpu).unsqueeze(0)
    res = []
    for pol in (hotrg.ALL_FP64, hotrg.Policy(fp64_merge="f64i8")):
        T, lnf = hotrg.renorm3d(T0.clone(), 8, 3, 3, 3, pol)
        res.append(float(hotrg.lnz3d(T, lnf, 9)[0]))
    print(f"\n[f64i8] chi=8 9 steps: fp64 {res[0]:.15f} int8 {res[1]:.15f} diff {res[1] - res[0]:.2e}", flush=True)
    assert abs(res[1] - res[0]) <= 1e-10
