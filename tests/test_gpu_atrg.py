"""GPU tests: decompositions, utility drop-ins and ATRG 3D parity."""
import warnings

import numpy as np
import pytest
import torch

from oracle import ATRG_3d_new as oat
from oracle import exact, models
from oracle import truncated_svd as otsvd

pytestmark = pytest.mark.gpu
warnings.filterwarnings("ignore")


def _dev(x, gpu):
    return torch.as_tensor(np.ascontiguousarray(x)).to(gpu)


@pytest.mark.parametrize("m,n", [(5, 3), (64, 64), (128, 128), (200, 40), (256, 64)])
def test_gesvj_matches_lapack(gpu, m, n):
    from trg_backup_amd import linalg
    rng = np.random.default_rng(m + n)
    A = rng.standard_normal((m, n)) * np.exp(-np.arange(n) / 8.0)[None, :]
    u, s, vh = linalg.svd_small(_dev(A, gpu)[None])
    u, s, vh = u[0].cpu().numpy(), s[0].cpu().numpy(), vh[0].cpu().numpy()
    ref = np.linalg.svd(A, compute_uv=False)
    assert np.allclose(s, ref, rtol=1e-12, atol=1e-14 * ref[0])
    assert np.allclose((u * s) @ vh, A, atol=1e-12 * ref[0])
    assert np.allclose(vh @ vh.T, np.eye(n), atol=1e-12)


def test_truncate_device_matches_oracle(gpu):
    from trg_backup_amd import linalg
    rng = np.random.default_rng(0)
    cases = [np.array([1.0, 0.5, 0.5, 0.1, 1e-12]), np.sort(rng.random(40))[::-1],
             np.concatenate([np.ones(3), 0.3 * np.ones(4), np.geomspace(0.1, 1e-14, 20)])]
    for s in cases:
        for k in (1, 2, 3, 5, 7, len(s), len(s) + 3):
            for de, te in ((0.0, 1e-10), (1e-3, 1e-10), (0.0, 0.0), (5e-2, 0.0), (1e-3, 1e-3)):
                ref = otsvd.truncate(s, k, de, te)
                got = linalg.truncate(_dev(s, gpu), k, de, te)
                assert got[0] == ref[0], (s[:6], k, de, te, got, ref)
                assert abs(got[1] - ref[1]) <= 1e-12 + 1e-12 * abs(ref[1])


@pytest.mark.parametrize("n,r", [(100, 100), (300, 40), (600, 250)])
def test_chol_psd_upper(gpu, n, r):
    from trg_backup_amd import linalg
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, r)) * np.exp(-np.arange(r) / 10.0)[None, :]
    A = X @ X.T
    R = linalg.chol_psd_upper(_dev(A, gpu)[None])[0].cpu().numpy()
    assert np.allclose(np.tril(R, -1), 0)
    assert np.abs(R.T @ R - A).max() < 1e-12 * np.abs(A).max()


def test_svd_thin_and_topk(gpu):
    from trg_backup_amd import linalg
    rng = np.random.default_rng(3)
    A = rng.standard_normal((5000, 40)) @ np.diag(np.geomspace(1, 1e-6, 40)) @ rng.standard_normal((40, 40))
    u, s, vh = linalg.svd_thin(_dev(A, gpu)[None])
    s = s[0].cpu().numpy()
    assert np.allclose(s, np.linalg.svd(A, compute_uv=False), rtol=1e-10)
    rec = (u[0].cpu().numpy() * s) @ vh[0].cpu().numpy()
    assert np.abs(rec - A).max() < 1e-12 * np.abs(A).max()
    lam = np.geomspace(1, 1e-12, 300)
    Q1, _ = np.linalg.qr(rng.standard_normal((400, 300)))
    Q2, _ = np.linalg.qr(rng.standard_normal((350, 300)))
    B = (Q1 * lam) @ Q2.T
    u, s, vh = linalg.svd_topk(_dev(B, gpu)[None], 64, 6)
    assert np.allclose(s[0].cpu().numpy()[:32], lam[:32], rtol=1e-9)


def test_utility_svd_eigh_dropins(gpu):
    from trg_backup_amd.utility import truncated_svd as tsvd
    rng = np.random.default_rng(4)
    A = rng.standard_normal((4, 5, 6))
    u, s, vh = tsvd.svd(A, shape=[[0, 2], [1]], k=5, truncate_eps=1e-14, degeneracy_eps=0)
    uo, so, vo = otsvd.svd(A, shape=[[0, 2], [1]], k=5, truncate_eps=1e-14, degeneracy_eps=0)
    assert u.shape == uo.shape and np.allclose(s, so, rtol=1e-12)
    assert np.allclose(np.einsum("aci,i,ib->abc", u, s, vh), A, atol=1e-12)
    H = rng.standard_normal((6, 6))
    H = H @ H.T
    e, v = tsvd.eigh(H.reshape(2, 3, 2, 3), shape=[[0, 1], [2, 3]], k=6, degeneracy_eps=0, truncate_eps=1e-15)
    eo, vo = otsvd.eigh(H.reshape(2, 3, 2, 3), shape=[[0, 1], [2, 3]], k=6, degeneracy_eps=0, truncate_eps=1e-15)
    assert np.allclose(e, eo, rtol=1e-12)


def test_rsvd_dropin_exact_when_sketch_covers_rank(gpu):
    from trg_backup_amd.utility import randomized_svd as rs
    rng = np.random.default_rng(5)
    A = rng.standard_normal((300, 8)) @ rng.standard_normal((8, 200))
    u, s, vh = rs.rsvd(A, k=8, n_oversamples=8, n_power_iter=1, seed=3)
    assert np.allclose(s, np.linalg.svd(A, compute_uv=False)[:8], rtol=1e-10)
    assert np.abs((u * s) @ vh - A).max() < 1e-10 * np.abs(A).max()


def test_atrg_ising_untruncated_equals_brute_force(gpu):
    from trg_backup_amd.tensor_class.tensor_class import ATRG_Tensor
    from trg_backup_amd.trg import ATRG_3d_new as at
    beta = 0.22
    T = models.to_atrg(models.ising3d_tensor(beta), 64)
    A = ATRG_Tensor(T.U, T.s, T.VH, 64, 3, False, {})
    at.set_rsvd_params(k_factor=2, p=0, q=2)
    try:
        v, _ = at.ln_Z_over_V(A, 64, {"X": 1, "Y": 1, "T": 1})
    finally:
        at.set_rsvd_params()
    assert abs(v - exact.brute_lnZ_ising3d(beta)) < 1e-12


def test_atrg_beta0_is_ln3(gpu):
    from trg_backup_amd.tensor_class.tensor_class import ATRG_Tensor
    from trg_backup_amd.trg import ATRG_3d_new as at
    T = models.to_atrg(models.potts3d_tensor(0.0), 12)
    v, _ = at.ln_Z_over_V(ATRG_Tensor(T.U, T.s, T.VH, 12, 3, False, {}), 12, {"X": 2, "Y": 2, "T": 2})
    assert abs(v - np.log(3)) < 1e-12


@pytest.mark.parametrize("model,beta,chi", [("ising", 0.22, 8), ("potts", 0.55, 10)])
def test_atrg_matches_oracle_identity_basis(gpu, model, beta, chi):
    """Same algorithm and the same Philox sketch on both sides (leg swap in
    the identity basis, sketch of 2 chi columns all handed to truncate, q = 4
    power iterations).  DGENERACY_EPS = 1e-8 on both sides: with 0 the rank
    rule may cut inside a multiplet split only by rounding (Z2 / S3
    symmetry), and which partner survives is then arbitrary in any two
    implementations (LAPACK vs Jacobi)."""
    from trg_backup_amd import atrg
    from trg_backup_amd.tensor_class.tensor_class import ATRG_Tensor
    from trg_backup_amd.trg import ATRG_3d_new as at
    T6 = models.ising3d_tensor(beta) if model == "ising" else models.potts3d_tensor(beta)
    steps = {"X": 2, "Y": 2, "T": 2}
    oat.set_leg_swap_basis("identity")
    # q = 20: the range finder converges to rounding, so the result no longer
    # depends on the sketch (the two sides' tensors differ by a sign gauge of
    # the internal bond, which a fixed Gaussian sketch sees)
    oat.set_rsvd_params(k_factor=2, p=0, q=20, seed=1)
    at.set_rsvd_params(k_factor=2, p=0, q=20, seed=1)
    old = (oat.degeneracy_eps, atrg.degeneracy_eps)
    oat.degeneracy_eps = atrg.degeneracy_eps = 1e-8
    try:
        ref, To = oat.ln_Z_over_V(models.to_atrg(T6, chi), chi, dict(steps))
        Tg = models.to_atrg(T6, chi)
        v, Tn = at.ln_Z_over_V(ATRG_Tensor(Tg.U, Tg.s, Tg.VH, chi, 3, False, {}), chi, dict(steps))
        ref_svd_basis = None
        oat.set_leg_swap_basis("svd")
        ref_svd_basis, _ = oat.ln_Z_over_V(models.to_atrg(T6, chi), chi, dict(steps))
    finally:
        oat.set_leg_swap_basis("svd")
        oat.set_rsvd_params()
        at.set_rsvd_params()
        oat.degeneracy_eps, atrg.degeneracy_eps = old
    print(f"\n[atrg {model} chi={chi}] gpu={v!r} oracle(identity)={ref!r} diff={v - ref:.3e} "
          f"oracle(svd basis)={ref_svd_basis!r} basis diff={ref_svd_basis - ref:.3e}", flush=True)
    assert Tn.U.shape == To.U.shape
    assert abs(v - ref) < 1e-11 * abs(ref)
    # the reference's SVD basis draws the sketch in rotated legs: with a converged
    # range finder it agrees to rounding as well
    assert abs(ref_svd_basis - ref) < 1e-10 * abs(ref)


@pytest.mark.parametrize("q", [2, 4])
def test_atrg_few_power_iterations_within_rsvd_accuracy(gpu, q):
    """The workload's setting (sketch 2 chi, q = 2): device and oracle agree
    to the randomized SVD's own accuracy (they see the sketch in different
    sign gauges of the internal bond)."""
    from trg_backup_amd import atrg
    from trg_backup_amd.tensor_class.tensor_class import ATRG_Tensor
    from trg_backup_amd.trg import ATRG_3d_new as at
    chi = 10
    T6 = models.potts3d_tensor(0.55)
    steps = {"X": 2, "Y": 2, "T": 2}
    oat.set_leg_swap_basis("identity")
    oat.set_rsvd_params(k_factor=2, p=0, q=q, seed=1)
    at.set_rsvd_params(k_factor=2, p=0, q=q, seed=1)
    old = (oat.degeneracy_eps, atrg.degeneracy_eps)
    oat.degeneracy_eps = atrg.degeneracy_eps = 1e-8
    try:
        ref, _ = oat.ln_Z_over_V(models.to_atrg(T6, chi), chi, dict(steps))
        oat.set_rsvd_params(k_factor=2, p=0, q=20, seed=1)
        conv, _ = oat.ln_Z_over_V(models.to_atrg(T6, chi), chi, dict(steps))
        Tg = models.to_atrg(T6, chi)
        v, _ = at.ln_Z_over_V(ATRG_Tensor(Tg.U, Tg.s, Tg.VH, chi, 3, False, {}), chi, dict(steps))
    finally:
        oat.set_leg_swap_basis("svd")
        oat.set_rsvd_params()
        at.set_rsvd_params()
        oat.degeneracy_eps, atrg.degeneracy_eps = old
    print(f"\n[atrg potts q={q}] gpu-oracle={v - ref:.3e} oracle(q)-oracle(converged)={ref - conv:.3e}", flush=True)
    assert abs(v - ref) <= 10 * abs(ref - conv) + 1e-12


@pytest.mark.parametrize("m,l,passes", [(32768, 64, 3), (32768, 64, 1), (4096, 128, 3), (300, 40, 3)])
def test_orthonormalize_tall_blocks(gpu, m, l, passes):
    """Tall blocks take the split (row-slab) Gram; three passes give an orthonormal basis,
    one pass a basis of the same span.  Both reproduce X = Q (Q^+ X)."""
    from trg_backup_amd import linalg
    rng = np.random.default_rng(m + l)
    X = rng.standard_normal((1, m, l)) @ np.diag(np.logspace(0, -6, l))          # graded columns
    Q = linalg.orthonormalize(torch.as_tensor(X).to(gpu), passes=passes)[0].cpu().numpy()
    G = Q.T @ Q
    if passes == 3:
        assert np.abs(G - np.eye(l)).max() < 1e-12
    coef = np.linalg.lstsq(Q, X[0], rcond=None)[0]
    assert np.abs(Q @ coef - X[0]).max() < 1e-9 * np.abs(X).max()


@pytest.mark.parametrize("case", ["decaying", "rank_deficient"])
def test_truncate_known_matches_full_spectrum(gpu, case):
    """The rank rule over the leading values plus the discarded mass equals
    the rule over the whole spectrum (oracle truncate), including an exactly
    rank-deficient spectrum whose computed tail is rounding noise."""
    from trg_backup_amd import linalg
    if case == "decaying":
        full = 10.0 ** (-0.8 * np.arange(40))            # tail rule cuts at 13 < k
        k, deg, known = 16, 0.0, 18
        tail_sq = np.sum(full[known:] ** 2)
    else:
        full = np.concatenate([np.array([1.0, 0.7, 0.7, 0.3, 0.1, 0.02]), np.zeros(30)])
        k, deg, known = 10, 1e-8, 12
        tail_sq = 3e-17                                   # what ||P||^2 - sum s^2 leaves behind
    total = np.sum(full ** 2)
    ref = otsvd.truncate(full, k, deg, 1e-10)
    chi, _ = linalg.truncate_known(_dev(full[:known], gpu), _dev(np.array([tail_sq]), gpu),
                                   _dev(np.array([total]), gpu), k, deg, 1e-10)
    assert chi == ref[0], (chi, ref)


def _atrg_pair(T6, chi, loc=None):
    from trg_backup_amd.tensor_class.tensor_class import ATRG_Tensor
    t = models.to_atrg(T6, chi)
    return ATRG_Tensor(t.U, t.s, t.VH, chi, 3, loc is not None, dict(loc or {}))


def test_atrg_one_point_equals_network_contraction(gpu):
    """Device one-point flow, untruncated (Dcut above every rank), on the
    2x2x2 box: P = Z[impurity at one site] / Z of the eight-tensor network
    for two random impurities carried together, at sites of both parities;
    ln Z/V equals the enumeration."""
    from oracle import HOTRG_3d_QR as hq
    from trg_backup_amd.trg import ATRG_3d_new as at
    rng = np.random.default_rng(21)
    T6 = models.ising3d_tensor(0.22)
    imps = [rng.standard_normal((2,) * 6) + 1.5 * T6, rng.standard_normal((2,) * 6) + 2.0 * T6]
    sites = [(i, j, k) for i in range(2) for j in range(2) for k in range(2)]
    Z = hq.network_value({s: T6 for s in sites}, (2, 2, 2))
    at.set_rsvd_params(k_factor=3, p=0, q=2)
    try:
        for loc in [(0, 0, 0), (1, 0, 1), (1, 1, 1)]:
            ref = [hq.network_value({s: (Ti if s == loc else T6) for s in sites}, (2, 2, 2)) / Z for Ti in imps]
            T = _atrg_pair(T6, 64)
            T0 = [_atrg_pair(Ti, 64, {"X": loc[0], "Y": loc[1], "T": loc[2]}) for Ti in imps]
            T, T0 = at.one_point_function(T, T0, 64, {"X": 1, "Y": 1, "T": 1})
            lnz, P = at.cal_field_expected_value(T, T0, {"X": 1, "Y": 1, "T": 1})
            print(f"\n[atrg 1pt] loc={loc} P={P} ref={ref}", flush=True)
            assert abs(lnz - exact.brute_lnZ_ising3d(0.22)) < 1e-12
            for p, r in zip(P, ref):
                assert abs(p - r) < 1e-10 * abs(r)
            assert all(t.loc == {"X": 0, "Y": 0, "T": 0} for t in T0)
    finally:
        at.set_rsvd_params()


def test_atrg_one_point_matches_oracle_truncated(gpu):
    """Truncated flow (chi 8, two steps per direction) against the oracle with
    the same algorithm and sketch (identity leg-swap basis, converged range
    finder, DGENERACY_EPS 1e-8 on both sides, as in the pure-flow parity
    test): ln Z/V and both impurity values agree to rounding."""
    from trg_backup_amd import atrg
    from trg_backup_amd.trg import ATRG_3d_new as at
    rng = np.random.default_rng(4)
    T6 = models.ising3d_tensor(0.22)
    imps6 = [T6 + 0.3 * rng.standard_normal((2,) * 6), 2.0 * T6 + 0.5 * rng.standard_normal((2,) * 6)]
    loc = {"X": 3, "Y": 2, "T": 1}
    steps = {"X": 2, "Y": 2, "T": 2}
    chi = 8
    oat.set_leg_swap_basis("identity")
    oat.set_rsvd_params(k_factor=2, p=0, q=20, seed=1)
    at.set_rsvd_params(k_factor=2, p=0, q=20, seed=1)
    old = (oat.degeneracy_eps, atrg.degeneracy_eps)
    oat.degeneracy_eps = atrg.degeneracy_eps = 1e-8
    try:
        To = models.to_atrg(T6, chi)
        T0o = []
        for Ti in imps6:
            t = models.to_atrg(Ti, chi)
            t.loc = dict(loc)
            T0o.append(t)
        _, _, lnz_ref, P_ref = oat.one_point_function(To, T0o, chi, dict(steps))
        T, T0 = at.one_point_function(_atrg_pair(T6, chi), [_atrg_pair(Ti, chi, loc) for Ti in imps6], chi,
                                      dict(steps))
        lnz, P = at.cal_field_expected_value(T, T0, steps)
    finally:
        oat.set_leg_swap_basis("svd")
        oat.set_rsvd_params()
        at.set_rsvd_params()
        oat.degeneracy_eps, atrg.degeneracy_eps = old
    print(f"\n[atrg 1pt chi={chi}] lnz {lnz!r} vs {lnz_ref!r}; P {P} vs {P_ref}", flush=True)
    assert abs(lnz - lnz_ref) < 1e-11 * abs(lnz_ref)
    for p, r in zip(P, P_ref):
        assert abs(p - r) < 1e-9 * abs(r)
    assert [t.loc for t in T0] == [t.loc for t in T0o]


@pytest.mark.parametrize("shape,pairs", [
    ((4, 1, 1), [((1, 0, 0), (2, 0, 0)), ((0, 0, 0), (3, 0, 0))]),
    ((1, 4, 1), [((0, 1, 0), (0, 2, 0))]),
    ((1, 1, 4), [((0, 0, 2), (0, 0, 1))]),
    ((2, 2, 2), [((0, 0, 0), (1, 1, 1))]),
])
def test_atrg_two_point_equals_network_contraction(gpu, shape, pairs):
    """Device two-point flow untruncated: the correlator equals the direct
    network contraction of the box with two different non-symmetric
    impurities (rings of 4 fix the pairing by coordinate parity)."""
    import itertools
    from oracle import HOTRG_3d_QR as hq
    from trg_backup_amd.trg import ATRG_3d_new as at
    rng = np.random.default_rng(9)
    T6 = models.ising3d_tensor(0.3)
    I0 = T6 + 0.4 * rng.standard_normal(T6.shape)
    In = 1.5 * T6 + 0.4 * rng.standard_normal(T6.shape)
    sites = list(itertools.product(*[range(L) for L in shape]))
    Z = hq.network_value({s: T6 for s in sites}, shape)
    steps = {"X": shape[0].bit_length() - 1, "Y": shape[1].bit_length() - 1, "T": shape[2].bit_length() - 1}
    at.set_rsvd_params(k_factor=3, p=0, q=2)
    try:
        for r0, rn in pairs:
            Zi = hq.network_value({s: (I0 if s == r0 else In if s == rn else T6) for s in sites}, shape)
            T, T0 = at.two_point_function(_atrg_pair(T6, 64), _atrg_pair(I0, 64, dict(zip("XYT", r0))),
                                          _atrg_pair(In, 64, dict(zip("XYT", rn))), 64, dict(steps))
            lnz, G = at.cal_field_expected_value(T, [T0], steps)
            print(f"\n[atrg 2pt {shape} {r0}-{rn}] gpu={G[0]!r} network={Zi / Z!r}", flush=True)
            assert abs(G[0] - Zi / Z) < 1e-10 * abs(Zi / Z)
            assert abs(lnz - np.log(Z) / len(sites)) < 1e-12
    finally:
        at.set_rsvd_params()


def test_atrg_two_point_matches_oracle_truncated(gpu):
    """Truncated two-point flow (chi 8, two steps per direction, impurities at
    (1,0,1) and (2,3,1)) against the oracle with the same sketch."""
    from trg_backup_amd import atrg
    from trg_backup_amd.trg import ATRG_3d_new as at
    rng = np.random.default_rng(6)
    T6 = models.ising3d_tensor(0.22)
    I0 = T6 + 0.3 * rng.standard_normal(T6.shape)
    In = T6 + 0.3 * rng.standard_normal(T6.shape)
    r0, rn = {"X": 1, "Y": 0, "T": 1}, {"X": 2, "Y": 3, "T": 1}
    steps = {"X": 2, "Y": 2, "T": 2}
    chi = 8
    oat.set_leg_swap_basis("identity")
    oat.set_rsvd_params(k_factor=2, p=0, q=20, seed=1)
    at.set_rsvd_params(k_factor=2, p=0, q=20, seed=1)
    old = (oat.degeneracy_eps, atrg.degeneracy_eps)
    oat.degeneracy_eps = atrg.degeneracy_eps = 1e-8
    try:
        T0o = models.to_atrg(I0, chi)
        T0o.loc = dict(r0)
        Tno = models.to_atrg(In, chi)
        Tno.loc = dict(rn)
        _, _, lnz_ref, G_ref = oat.two_point_function(models.to_atrg(T6, chi), T0o, Tno, chi, dict(steps))
        T, T0 = at.two_point_function(_atrg_pair(T6, chi), _atrg_pair(I0, chi, r0), _atrg_pair(In, chi, rn), chi,
                                      dict(steps))
        lnz, G = at.cal_field_expected_value(T, [T0], steps)
    finally:
        oat.set_leg_swap_basis("svd")
        oat.set_rsvd_params()
        at.set_rsvd_params()
        oat.degeneracy_eps, atrg.degeneracy_eps = old
    print(f"\n[atrg 2pt chi={chi}] G {G[0]!r} vs {G_ref!r}; lnz {lnz!r} vs {lnz_ref!r}", flush=True)
    assert abs(lnz - lnz_ref) < 1e-11 * abs(lnz_ref)
    assert abs(G[0] - G_ref) < 1e-9 * abs(G_ref)


def test_atrg_init_and_inte_rsvds_exact_when_sketch_covers_rank(gpu):
    """utility/randomized_svd.py:100-226 (rsvd_for_3dATRG_tensor_init / _inte):
    with a sketch at least the rank of the product, the singular values equal
    the dense SVD of the product and U s VH rebuilds it.  The sketches are
    wider than the ranks (12 columns for rank 5), which the range finder
    repairs with a 1e-10 perturbation, hence the 1e-9 tolerance."""
    from trg_backup_amd.utility import randomized_svd as rs
    rng = np.random.default_rng(3)
    A = rng.standard_normal((3, 3, 3, 5))
    B = rng.standard_normal((5, 3, 3, 3))
    M = np.einsum("abci,idef->abcdef", A, B).reshape(27, 27)
    s_ref = np.linalg.svd(M, compute_uv=False)
    u, s, vh = rs.rsvd_for_3dATRG_tensor_init(A, B, 8, n_oversamples=4, n_power_iter=2)
    assert np.allclose(s[:5], s_ref[:5], rtol=1e-9, atol=1e-9 * s_ref[0])
    Mr = np.einsum("abcj,j,jdef->abcdef", u[..., :5], s[:5], vh[:5]).reshape(27, 27)
    assert np.abs(Mr - M).max() < 1e-9 * np.abs(M).max()
    Bt = rng.standard_normal((3, 2, 3, 3))      # B_{idbc}
    Ct = rng.standard_normal((2, 3, 3, 3))      # C_{djka}
    M2 = np.einsum("idbc,djka->ijkabc", Bt, Ct).reshape(27, 27)
    s2_ref = np.linalg.svd(M2, compute_uv=False)
    r2 = int((s2_ref > 1e-12 * s2_ref[0]).sum())
    u2, s2, vh2 = rs.rsvd_for_3dATRG_tensor_inte(Bt, Ct, 20, n_oversamples=4, n_power_iter=2)
    assert np.allclose(s2[:r2], s2_ref[:r2], rtol=1e-9, atol=1e-9 * s2_ref[0])
    M2r = np.einsum("ijkl,l,labc->ijkabc", u2[..., :r2], s2[:r2], vh2[:r2]).reshape(27, 27)
    assert np.abs(M2r - M2).max() < 1e-9 * np.abs(M2).max()


@pytest.mark.gpu
@pytest.mark.parametrize("betas", [(0.53, 0.55, 0.57), (0.01, 0.55, 0.30)])
def test_atrg_batched_steps_equal_alone(gpu, betas):
    """atrg.atrg_pure_tensor_batch (trg/ATRG_3d_new.py:441-482 for several
    instances per launch) gives every instance, bit for bit, the tensors it
    gets stepped alone: every kernel is batch-invariant, and where the
    instances' rank decisions part (the second set runs with a 1e-4 tail cut,
    under which the weakly coupled instance keeps fewer values) the step is re-run per group of equal
    decisions.  normalize_batch equals normalize per instance."""
    from trg_backup_amd import atrg
    from trg_backup_amd import models as dm
    chi, steps = 10, 7
    saved = (dict(atrg.RSVD), atrg.degeneracy_eps, atrg.truncate_eps)
    atrg.set_rsvd_params(k_factor=2, p=0, q=2, seed=1)
    atrg.degeneracy_eps = 1e-8
    if 0.01 in betas:
        # a coarse tail cut makes the kept ranks depend on the coupling
        atrg.truncate_eps = 1e-4
    splits0 = atrg.SPLITS
    try:
        alone, batched = [], []
        for b in betas:
            U, s, VH = dm.atrg_initial("potts3d", b)
            s, _ = atrg.normalize(U, s, VH)
            alone.append((U, s, VH))
            batched.append((U, s, VH))
        for n in range(1, steps + 1):
            d = "TXY"[(n - 1) % 3]
            alone = [atrg.atrg_pure_tensor(U, s, VH, chi, d, "f64") for U, s, VH in alone]
            groups = {}
            for i, (U, s, VH) in enumerate(batched):
                groups.setdefault((tuple(U.shape), tuple(VH.shape)), []).append(i)
            nxt = [None] * len(batched)
            for idx in groups.values():
                res = atrg.atrg_pure_tensor_batch(atrg._stack([batched[i][0] for i in idx]),
                                                  atrg._stack([batched[i][1] for i in idx]),
                                                  atrg._stack([batched[i][2] for i in idx]), chi, d, "f64")
                for j, i in enumerate(idx):
                    nxt[i] = res[j]
            batched = nxt
            for i in range(len(betas)):
                for x, y in zip(alone[i], batched[i]):
                    assert x.shape == y.shape and torch.equal(x, y), (betas[i], n)
            # normalisation: batched over the instances of equal shape, each against its own
            idx = [i for i in range(len(betas)) if batched[i][0].shape == batched[0][0].shape
                   and batched[i][2].shape == batched[0][2].shape and batched[i][1].shape == batched[0][1].shape]
            sb, cb = atrg.normalize_batch(atrg._stack([batched[i][0] for i in idx]),
                                          atrg._stack([batched[i][1] for i in idx]),
                                          atrg._stack([batched[i][2] for i in idx]))
            for j, i in enumerate(idx):
                s1, c1 = atrg.normalize(*alone[i])
                assert torch.equal(sb[j], s1) and torch.equal(cb[j:j + 1], c1)
            alone = [(U, atrg.normalize(U, s, VH)[0], VH) for U, s, VH in alone]
            batched = [(U, atrg.normalize(U, s, VH)[0], VH) for U, s, VH in batched]
        print(f"\n[atrg batched {betas}] {steps} steps bitwise equal alone; rank splits {atrg.SPLITS - splits0}",
              flush=True)
        if 0.01 in betas:
            assert atrg.SPLITS > splits0
    finally:
        atrg.RSVD.update(saved[0])
        atrg.degeneracy_eps, atrg.truncate_eps = saved[1], saved[2]
