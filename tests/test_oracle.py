"""CPU tests that pin the NumPy oracle against closed forms, brute force and LAPACK."""
import warnings

import numpy as np
import pytest

from oracle import ATRG_3d_new as at
from oracle import HOTRG_2d as h2
from oracle import HOTRG_3d as h3
from oracle import exact, models, philox, randomized_svd, truncated_svd

warnings.filterwarnings("ignore")

ONSAGER = {0.40: 0.879363820774948, 0.44: 0.928728540456889, 0.48: 0.991524464466794}


@pytest.mark.parametrize("beta", sorted(ONSAGER))
def test_onsager_two_quadratures(beta):
    a = exact.onsager_lnZ(beta)
    b = exact.onsager_lnZ_double(beta)
    assert abs(a - ONSAGER[beta]) < 1e-13
    assert abs(b - ONSAGER[beta]) < 1e-12


def test_brute_force_against_transfer_matrix():
    # 2D torus Ly x Lx via a row transfer matrix: independent of enumeration
    beta, L = 0.37, 3
    confs = ((np.arange(2 ** L)[:, None] >> np.arange(L)[None, :]) & 1) * 2 - 1
    intra = np.sum(confs * np.roll(confs, 1, axis=1), axis=1)
    inter = confs @ confs.T
    Tm = np.exp(beta * (inter + 0.5 * (intra[:, None] + intra[None, :])))
    lnZ = np.log(np.trace(np.linalg.matrix_power(Tm, L))) / (L * L)
    assert abs(lnZ - exact.brute_lnZ_ising2d(beta, L, L)) < 1e-13


def test_hotrg2d_untruncated_equals_brute_force():
    beta = 0.44
    T = models.ising2d_tensor(beta)
    assert abs(h2.ln_Z_over_V(T, 16, 2, 2) - exact.brute_lnZ_ising2d(beta, 4, 4)) < 1e-13


def test_hotrg2d_zero_padding_equals_growth():
    T = models.ising2d_tensor(0.41)
    a = h2.ln_Z_over_V(T, 8, 4, 4)
    b = h2.ln_Z_over_V(models.zero_pad(T, 8), 8, 4, 4)
    assert abs(a - b) < 1e-12


def test_hotrg2d_beta0_is_ln2():
    assert abs(h2.ln_Z_over_V(models.ising2d_tensor(0.0), 8, 6, 6) - np.log(2)) < 1e-13


def test_hotrg2d_chi16_onsager():
    v = h2.ln_Z_over_V(models.ising2d_tensor(0.44), 16, 15, 15)
    assert abs(v - ONSAGER[0.44]) < 1e-6


def test_hotrg2d_projector_properties():
    T = models.ising2d_tensor(0.44)
    for _ in range(3):
        T = h2.new_pure_tensor(T, 8, "Y")
        T, _ = h2.__tensor_normalization__(T)
        T = h2.new_pure_tensor(T, 8, "X")
        T, _ = h2.__tensor_normalization__(T)
    MM1, MM2 = h2.environments(T)
    # environment = M M^T with M the left half-network, built independently by einsum
    ref = np.einsum("aijk,blkm,cijn,dlnm->abcd", T, T, T, T).reshape(MM1.shape)
    assert np.allclose(MM1, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    assert np.allclose(MM1, MM1.T) and np.linalg.eigvalsh(MM1).min() > -1e-12 * np.abs(MM1).max()
    U, _ = h2.__contraction_projector__(T, 8)
    Um = U.reshape(-1, U.shape[2])
    assert np.allclose(Um.T @ Um, np.eye(Um.shape[1]), atol=1e-12)
    # merge equals the one-shot einsum
    ref = np.einsum("acke,bdel,abi,cdj->ijkl", T, T, U, U)
    assert np.allclose(h2.merge(T, T, U, U, nslice=3), ref, atol=1e-12 * np.abs(ref).max())


def test_hotrg3d_untruncated_equals_brute_force():
    beta = 0.22
    assert abs(h3.ln_Z_over_V(models.ising3d_tensor(beta), 16, 1, 1, 1) - exact.brute_lnZ_ising3d(beta)) < 1e-13


def test_hotrg3d_merge_and_env_match_einsum():
    rng = np.random.default_rng(0)
    T = rng.standard_normal((3, 3, 2, 2, 2, 2))
    U1 = np.linalg.qr(rng.standard_normal((9, 4)))[0].reshape(3, 3, 4)
    U2 = np.linalg.qr(rng.standard_normal((4, 3)))[0].reshape(2, 2, 3)
    ref = np.einsum("ijklez,mnopzf,ima,jnb,koc,lpd->abcdef", T, T, U1, U1, U2, U2)
    assert np.allclose(h3.merge(T, T, U1, U1, U2, U2, nslice=2), ref, atol=1e-12)
    ref = np.einsum("efaghm,ijbkml,efcghn,ijdknl->abcd", T, T, T, T).reshape(4, 4)
    assert np.allclose(h3.environment(T, 2), ref, atol=1e-12)


def test_hotrg3d_beta0():
    assert abs(h3.ln_Z_over_V(models.ising3d_tensor(0.0), 4, 2, 2, 2) - np.log(2)) < 1e-13


def test_atrg_ising_untruncated_equals_brute_force():
    at.set_rsvd_params()
    beta = 0.22
    v, _ = at.ln_Z_over_V(models.to_atrg(models.ising3d_tensor(beta), 64), 64, {"X": 1, "Y": 1, "T": 1})
    assert abs(v - exact.brute_lnZ_ising3d(beta)) < 1e-12


@pytest.mark.parametrize("steps,shape", [({"X": 0, "Y": 0, "T": 1}, (1, 1, 2)),
                                         ({"X": 1, "Y": 0, "T": 1}, (2, 1, 2))])
def test_atrg_potts_untruncated_equals_brute_force(steps, shape):
    at.set_rsvd_params(k_factor=1, p=0, q=2)
    try:
        v, _ = at.ln_Z_over_V(models.to_atrg(models.potts3d_tensor(0.55), 243), 243, steps)
    finally:
        at.set_rsvd_params()
    assert abs(v - exact.brute_lnZ_potts3d(0.55, shape=shape)) < 1e-12


def test_atrg_beta0_potts_is_ln3():
    v, _ = at.ln_Z_over_V(models.to_atrg(models.potts3d_tensor(0.0), 16), 16, {"X": 2, "Y": 2, "T": 2})
    assert abs(v - np.log(3)) < 1e-12


def test_atrg_G_H_match_einsum():
    rng = np.random.default_rng(1)
    A = rng.standard_normal((2, 3, 2, 4)); UM = rng.standard_normal((4, 2, 3, 5))
    PX = rng.standard_normal((3, 2, 3)); PY = rng.standard_normal((2, 3, 2))
    ref = np.einsum("Tabi,acX,icdk,bdY->TXYk", A, PX, UM, PY)
    assert np.allclose(at.contract_G(A, PX, UM, PY), ref)
    VMH = rng.standard_normal((5, 4, 3, 2)); Px = rng.standard_normal((3, 3, 2))
    D = rng.standard_normal((4, 2, 2, 3)); Py = rng.standard_normal((2, 2, 3))
    ref = np.einsum("kjab,xac,jtcd,ybd->ktxy", VMH, Px, D, Py)
    assert np.allclose(at.contract_H(VMH, Px, D, Py), ref)


def test_truncate_hand_cases():
    s = np.array([1.0, 0.5, 0.5, 0.1, 1e-12])
    assert truncated_svd.truncate(s, 3, 0, 0) == (3, 0)
    assert truncated_svd.truncate(s, 9, 1e-3, 1e-10) == (5, 0)
    # k = 2 would split the 0.5 pair: utility/truncated_svd.py:52-54 steps down to chi = 1
    chi, err = truncated_svd.truncate(s, 2, 1e-3, 1e-10)
    assert chi == 1 and abs(err - np.sqrt(np.sum(s[1:] ** 2) / np.sum(s ** 2))) < 1e-15
    # k = 3 closes the pair and is kept
    assert truncated_svd.truncate(s, 3, 1e-3, 1e-10)[0] == 3
    chi, _ = truncated_svd.truncate(s, 5, 0.0, 1e-10)      # tail norm 1e-12 < eps -> drop it
    assert chi == 4


def test_truncated_svd_wrappers_reconstruct():
    rng = np.random.default_rng(2)
    A = rng.standard_normal((3, 4, 5))
    u, s, vh = truncated_svd.svd(A, shape=[[0, 2], [1]], k=4, truncate_eps=1e-14, degeneracy_eps=0)
    R = np.einsum("aci,i,ib->abc", u, s, vh)
    assert np.allclose(R, A)
    H = rng.standard_normal((6, 6)); H = H @ H.T
    e, v = truncated_svd.eigh(H.reshape(2, 3, 2, 3), shape=[[0, 1], [2, 3]], k=6, degeneracy_eps=0, truncate_eps=1e-15)
    assert np.allclose(e, np.sort(np.linalg.eigvalsh(H))[::-1])


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32_10
    out = philox.philox4x32_10(0, 0, 0, 0, 0, 0)
    assert [int(x) for x in out] == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    m = 0xFFFFFFFF
    out = philox.philox4x32_10(m, m, m, m, m, m)
    assert [int(x) for x in out] == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    out = philox.philox4x32_10(0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344, 0xa4093822, 0x299f31d0)
    assert [int(x) for x in out] == [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_gaussian_moments():
    z = philox.gaussian(200000, seed=1)
    assert abs(z.mean()) < 0.01 and abs(z.var() - 1) < 0.01


def test_rsvd_exact_when_sketch_covers_rank():
    rng = np.random.default_rng(3)
    A = rng.standard_normal((40, 6)) @ rng.standard_normal((6, 30))
    u, s, vh = randomized_svd.rsvd(A, k=6, n_oversamples=4, n_power_iter=1)
    assert np.allclose(s, np.linalg.svd(A, compute_uv=False)[:6])
    assert np.allclose((u * s) @ vh, A)
    u, s, vh = randomized_svd.rsvd2(A, k=6, n_oversamples=2)
    assert np.allclose((u * s) @ vh, A)


def _check_truncated_factorisation(u, s, vh, M, k):
    """u s vh must be the rank-k truncated SVD of M (gauge-free checks)."""
    U0, s0, V0h = np.linalg.svd(M)
    assert np.allclose(s, s0[:k], rtol=1e-12, atol=1e-12 * s0[0])
    um = u.reshape(-1, u.shape[-1])
    vm = vh.reshape(vh.shape[0], -1)
    assert np.allclose(um.T @ um, np.eye(k), atol=1e-12)
    assert np.allclose(um @ um.T, U0[:, :k] @ U0[:, :k].T, atol=1e-10)
    assert np.allclose((um * s) @ vm, (U0[:, :k] * s0[:k]) @ V0h[:k], atol=1e-11 * s0[0])


@pytest.mark.parametrize("iterator,q", [("QR", 0), ("QR", 2), ("power", 2)])
def test_rsvd_for_3datrg_tensor_init_matches_dense_svd(iterator, q):
    """utility/randomized_svd.py:100-162: the operator is "abci,idef->abcdef" (the sketch
    product at :125 applied to the identity); its adjoint is the :119 string.  The inner
    bond r=5 bounds the rank, and l = Dcut + p = 7 covers it, so the range finder is exact
    and the result must be the dense truncated SVD."""
    rng = np.random.default_rng(11)
    D, r = 3, 5
    A = rng.standard_normal((D, D, D, r))
    B = rng.standard_normal((r, D, D, D))
    M = np.einsum("abci,idef->abcdef", A, B).reshape(D ** 3, D ** 3)
    Q = rng.standard_normal((D ** 3, 4))
    adj = np.einsum("iabc,defi,defl->abcl", B, A, Q.reshape(D, D, D, 4)).reshape(D ** 3, 4)
    assert np.allclose(adj, M.T @ Q)
    u, s, vh = randomized_svd.rsvd_for_3dATRG_tensor_init(A, B, D, n_oversamples=4, n_power_iter=q,
                                                           iterator=iterator, seed=5)
    assert u.shape == (D, D, D, D) and s.shape == (D,) and vh.shape == (D, D, D, D)
    _check_truncated_factorisation(u, s, vh, M, D)


@pytest.mark.parametrize("iterator,q", [("QR", 0), ("QR", 2), ("power", 2)])
def test_rsvd_for_3datrg_tensor_inte_matches_dense_svd(iterator, q):
    """utility/randomized_svd.py:165-226: M_{(i,j,k),(a,b,c)} from the MQ string
    "idbc,djka,abcl->ijkl" (:202) applied to the identity; the first sketch's string
    "iabc,ajkd,dbcl->ijkl" (:209) and the Q^dag M string "abcl,adjk,dbci->lijk" (:222,
    with Q the identity) must give the same operator, and "adjk,dbci,abcl->ijkl" (:196)
    its adjoint.  With an
    inner bond of 1 the rank is at most Dcut^2 = 9 and l = Dcut + p = 10 covers it."""
    rng = np.random.default_rng(12)
    D, x = 3, 1
    B = rng.standard_normal((D, x, D, D))
    C = rng.standard_normal((x, D, D, D))
    n = D ** 3
    eye = np.eye(n).reshape(D, D, D, n)
    M = np.einsum("idbc,djka,abcl->ijkl", B, C, eye).reshape(n, n)
    assert np.allclose(np.einsum("iabc,ajkd,dbcl->ijkl", B, C, eye).reshape(n, n), M)
    assert np.allclose(np.einsum("abcl,adjk,dbci->lijk", eye, B, C).reshape(n, n), M)
    assert np.allclose(np.einsum("adjk,dbci,abcl->ijkl", B, C, eye).reshape(n, n), M.T)
    assert np.linalg.matrix_rank(M) <= D * D
    u, s, vh = randomized_svd.rsvd_for_3dATRG_tensor_inte(B, C, D, n_oversamples=7, n_power_iter=q,
                                                           iterator=iterator, seed=6)
    assert u.shape == (D, D, D, D) and s.shape == (D,) and vh.shape == (D, D, D, D)
    _check_truncated_factorisation(u, s, vh, M, D)


def test_rsvd_for_3datrg_wide_sketch_ignores_the_draw():
    """Two different seeds give the same truncated factorisation once the sketch covers
    the rank (only the gauge of u and vh may differ), so a swapped or transposed operand
    cannot hide behind the random draw."""
    rng = np.random.default_rng(13)
    A = rng.standard_normal((2, 2, 2, 3))
    B = rng.standard_normal((3, 2, 2, 2))
    a = randomized_svd.rsvd_for_3dATRG_tensor_init(A, B, 2, n_oversamples=2, seed=1)
    b = randomized_svd.rsvd_for_3dATRG_tensor_init(A, B, 2, n_oversamples=2, seed=2)
    assert np.allclose(a[1], b[1])
    ua, ub = a[0].reshape(8, 2), b[0].reshape(8, 2)
    assert np.allclose(ua @ ua.T, ub @ ub.T, atol=1e-10)


def test_leg_swap_rsvd_matches_dense_svd():
    rng = np.random.default_rng(4)
    B = rng.standard_normal((4, 3, 5)); C = rng.standard_normal((5, 4, 3))
    M = np.einsum("bit,tjg->igjb", B, C).reshape(9, 16)
    UM, sM, VMH = at.rsvd_for_ATRG_leg_swapping(B, C, chi_k=9, p=0, q=2, eps=1e-12, seed=7)
    assert np.allclose(sM, np.linalg.svd(M, compute_uv=False)[:len(sM)])
    assert np.allclose(np.einsum("igk,k,kjb->igjb", UM, sM, VMH).reshape(9, 16), M)


# ---------------------------------------------------------------------------
# HOTRG 3D with squeezers (trg/HOTRG_3d_QR.py)
# ---------------------------------------------------------------------------
from oracle import HOTRG_3d_QR as hq  # noqa: E402


@pytest.mark.parametrize("kind", ["original", "QR", "fix"])
def test_hotrg3d_qr_untruncated_equals_brute_force(kind):
    """Z, X, Y on a 2x2x2 torus with Dcut = 16: no truncation, so every
    squeezer must give the exact partition function (the oblique QR
    projectors multiply to the identity on the kept space)."""
    beta = 0.3
    v, _ = hq.ln_Z_over_V(models.ising3d_tensor(beta), 16, {"X": 1, "Y": 1, "Z": 1}, kind)
    assert abs(v - exact.brute_lnZ_ising3d(beta)) < 1e-12


@pytest.mark.parametrize("steps,shape", [({"X": 0, "Y": 0, "Z": 1}, (1, 1, 2)),
                                         ({"X": 1, "Y": 0, "Z": 1}, (2, 1, 2))])
@pytest.mark.parametrize("kind", ["original", "QR"])
def test_hotrg3d_qr_potts_boxes(kind, steps, shape):
    v, _ = hq.ln_Z_over_V(models.potts3d_tensor(0.55), 81, steps, kind)
    assert abs(v - exact.brute_lnZ_potts3d(0.55, shape=shape)) < 1e-12


@pytest.mark.parametrize("kind", ["original", "QR", "fix"])
def test_hotrg3d_qr_beta0_is_ln2(kind):
    v, _ = hq.ln_Z_over_V(models.ising3d_tensor(0.0), 4, {"X": 2, "Y": 2, "Z": 2}, kind)
    assert abs(v - np.log(2)) < 1e-13


def test_hotrg3d_qr_environments_and_coarse_graining_match_einsum():
    """The environments against the reference's four-operand strings
    (trg/HOTRG_3d_QR.py:212-215) and the pairwise coarse graining against the
    one-shot einsum of :388, on random tensors with unequal bonds."""
    rng = np.random.default_rng(3)
    T0 = rng.standard_normal((2, 3, 2, 3, 2, 3))
    T1 = rng.standard_normal((2, 3, 2, 3, 3, 2))
    strings = {1: "aibcdm,ejfgmh,akbcdn,elfgnh->ijkl", 0: "iabcdm,jefgmh,kabcdn,lefgnh->ijkl",
               2: "abicdm,efjgmh,abkcdn,eflgnh->ijkl", 3: "abcidm,efgjmh,abckdn,efglnh->ijkl"}
    for leg, s in strings.items():
        ref = np.einsum(s, T0, T1, T0, T1)
        n = T0.shape[leg] * T1.shape[leg]
        assert np.allclose(hq.environment(T0, T1, leg), ref.reshape(n, n))
    Px = rng.standard_normal((4, 2, 2)); PX = rng.standard_normal((3, 3, 5))
    Py = rng.standard_normal((2, 2, 3)); PY = rng.standard_normal((2, 3, 3))
    assert np.allclose(hq.coarse_graining(T0, T1, Px, PX, Py, PY),
                       hq.coarse_graining_einsum(T0, T1, Px, PX, Py, PY))


def test_hotrg3d_qr_projector_identities():
    """Px PX = 1 and PY Py = 1: the QR squeezer's projectors are biorthogonal
    (S^-1/2 U^H (RL RR) V S^-1/2 = 1).  The 'original' isometry is
    orthonormal and spans the top eigenvectors of the x-leg environment (the
    reference's side comparison sums empty slices, so that side always wins)."""
    T = hq.leg_transposition(models.ising3d_tensor(0.25), "transpose", "Z")
    T = hq.new_pure_tensor(T, 3, "Z", "QR")               # a tensor with bonds 3, 3, 3, 3, 2, 2
    Px, PX, Py, PY, _ = hq.squeezer(T, T, T, T, 4, "QR")
    k = Px.shape[0]
    assert np.allclose(Px.reshape(k, -1) @ PX.reshape(-1, k), np.eye(k), atol=1e-10)
    ky = PY.shape[0]
    assert np.allclose(PY.reshape(ky, -1) @ Py.reshape(-1, ky), np.eye(ky), atol=1e-10)
    Px, PX, Py, PY, _ = hq.squeezer(T, T, T, T, 4, "original")
    U = PX.reshape(-1, PX.shape[2])
    assert np.allclose(U.T @ U, np.eye(U.shape[1]), atol=1e-12)
    w, V = np.linalg.eigh(hq.environment(T, T, 0))
    top = V[:, ::-1][:, :U.shape[1]]
    assert np.allclose(U @ U.T, top @ top.T, atol=1e-10)


def test_hotrg3d_qr_slice_legs():
    for leg, sl in [(7, 3), (64, 5), (3, 8), (10, 1), (16, 16)]:
        parts, n = hq.slice_legs(leg, sl)
        assert n == len(parts) == min(leg, max(sl, 1))
        covered = [i for p in parts for i in range(p.start, p.stop)]
        assert covered == list(range(leg))


# ---------------------------------------------------------------------------
# impurity flows (two-point functions) against enumeration
# ---------------------------------------------------------------------------
def _spin_impurity(beta, nlegs, shift=0.0):
    """sum_s (shift + s) prod_leg W_{s,leg}: the site tensor with the spin
    (plus a constant, so that a trace normalisation sees a non-zero trace) inserted."""
    W = models.ising_W(beta)
    sig = np.array([1.0, -1.0]) + shift
    T = np.zeros((2,) * nlegs)
    for s in range(2):
        t = W[s]
        for _ in range(nlegs - 1):
            t = np.multiply.outer(t, W[s])
        T += sig[s] * t
    return T


def test_hotrg2d_ynearest_two_point_equals_enumeration():
    """<s s'> of y-neighbours on the 4x4 torus from the untruncated impurity
    flow (trg/HOTRG_2d.py:300-348): exp(sum lnf_imp) Tr T_imp / Tr T."""
    beta = 0.35
    T, Ti = models.ising2d_tensor(beta), _spin_impurity(beta, 4)
    Tn, Tin, lnf, lnf_imp = h2.ynearest_two_point_func_renorm(T, Ti, Ti, 16, 2, 2)
    corr = np.exp(np.sum(lnf_imp)) * np.einsum("aabb", Tin) / np.einsum("aabb", Tn)
    assert abs(corr - exact.brute_corr_ising(beta, (4, 4), (0, 0), (0, 1))) < 1e-12


def test_hotrg3d_nearest_two_point_equals_enumeration():
    """trg/HOTRG_3d.py:398-453 merges the two impurities in its first (Z)
    step: z-neighbours on the 2x2x2 torus, untruncated.  The 3D flow
    normalises by |Tr|, which is zero for a bare spin impurity, so the
    impurity is 2 + s: <(2 + s)(2 + s')> = 4 + <s s'> (<s> = 0 at h = 0)."""
    beta = 0.25
    T, Ti = models.ising3d_tensor(beta), _spin_impurity(beta, 6, shift=2.0)
    Tn, Tin, lnf, lnf_imp = h3.ynearest_two_point_func_renorm(T, Ti, Ti, 16, 1, 1, 1)
    val = np.exp(np.sum(lnf_imp)) * np.einsum("xxyytt", Tin) / np.einsum("xxyytt", Tn)
    assert abs(val - (4.0 + exact.brute_corr_ising(beta, (2, 2, 2), (0, 0, 0), (0, 0, 1)))) < 1e-12


def test_hotrg2d_two_point_separated_equals_enumeration():
    """Two impurities (nx, ny) = (1, 1) apart (trg/HOTRG_2d.py:219-297) on the 4x4 torus."""
    beta = 0.35
    T, Ti = models.ising2d_tensor(beta), _spin_impurity(beta, 4)
    Tn, Tin, lnf, lnf_imp = h2.two_point_func_renorm(T, Ti, Ti, 1, 1, 16, 2, 2)
    corr = np.exp(np.sum(lnf_imp)) * np.einsum("aabb", Tin) / np.einsum("aabb", Tn)
    assert abs(corr - exact.brute_corr_ising(beta, (4, 4), (0, 0), (1, 1))) < 1e-12


def test_hotrg3d_qr_one_point_equals_network_contraction():
    """trg/HOTRG_3d_QR.py:859-908 untruncated on the 2x2x2 torus: the
    one-point value P = Z(T0 at loc) / Z equals the direct contraction of the
    eight-tensor network, for a random impurity at each of two sites (both
    parities of the impurity coordinate in every direction)."""
    import itertools
    beta = 0.3
    T = models.ising3d_tensor(beta)
    T0 = T + 0.3 * np.random.default_rng(4).standard_normal(T.shape)
    sites = list(itertools.product(range(2), range(2), range(2)))
    Z = hq.network_value({s: T for s in sites}, (2, 2, 2))
    for loc in [(0, 0, 0), (1, 0, 1)]:
        Zi = hq.network_value({s: (T0 if s == loc else T) for s in sites}, (2, 2, 2))
        for kind in ("QR", "original"):
            _, _, P = hq.one_point_function({"Tensor": T, "loc": {}, "factor": {}},
                                            {"Tensor": T0, "loc": {"X": loc[0], "Y": loc[1], "Z": loc[2]},
                                             "factor": {}}, 16, {"X": 1, "Y": 1, "Z": 1}, kind)
            assert abs(P - Zi / Z) < 1e-11 * abs(Zi / Z), (loc, kind, P, Zi / Z)


def test_atrg_one_point_equals_network_contraction():
    """The ATRG one-point flow (untruncated: Dcut above every rank) on the
    2x2x2 box gives P = Z[impurity at one site] / Z of the eight-tensor
    network, for a random impurity at every site (both parities of each
    coordinate) and two impurities carried in one flow."""
    import itertools
    from oracle import HOTRG_3d_QR as hq
    at.set_rsvd_params(k_factor=3, p=0, q=2)          # the sketch covers every rank here
    try:
        rng = np.random.default_rng(21)
        T6 = models.ising3d_tensor(0.22)
        imps = [rng.standard_normal((2,) * 6) + 1.5 * T6, rng.standard_normal((2,) * 6) + 2.0 * T6]
        sites = list(itertools.product(range(2), repeat=3))
        Z = hq.network_value({s: T6 for s in sites}, (2, 2, 2))
        for loc in sites:
            ref = [hq.network_value({s: (Ti if s == loc else T6) for s in sites}, (2, 2, 2)) / Z for Ti in imps]
            T = models.to_atrg(T6, 64)
            T0 = []
            for Ti in imps:
                t = models.to_atrg(Ti, 64)
                t.loc = {"X": loc[0], "Y": loc[1], "T": loc[2]}
                T0.append(t)
            _, _, lnz, P = at.one_point_function(T, T0, 64, {"X": 1, "Y": 1, "T": 1})
            assert abs(lnz - exact.brute_lnZ_ising3d(0.22)) < 1e-12
            for p, r in zip(P, ref):
                assert abs(p - r) < 1e-11 * abs(r), (loc, p, r)
    finally:
        at.set_rsvd_params()


@pytest.mark.parametrize("shape,pairs", [
    ((4, 1, 1), [((0, 0, 0), (1, 0, 0)), ((0, 0, 0), (3, 0, 0)), ((1, 0, 0), (2, 0, 0)), ((2, 0, 0), (0, 0, 0))]),
    ((1, 4, 1), [((0, 1, 0), (0, 2, 0)), ((0, 3, 0), (0, 0, 0))]),
    ((1, 1, 4), [((0, 0, 1), (0, 0, 2)), ((0, 0, 0), (0, 0, 3))]),
    ((2, 2, 2), [((0, 0, 0), (1, 1, 1)), ((1, 0, 1), (0, 1, 1))]),
])
def test_hotrg3d_qr_two_point_equals_network_contraction(shape, pairs):
    """trg/HOTRG_3d_QR.py:910-1033 untruncated: G = Z(T0 at r0, Tn at rn) / Z
    equals the direct contraction of the box's network, with two different
    random (non-symmetric) impurities, so the value depends on the
    displacement's sign as well as its length: rings of 4 along each axis pin
    which tensor of a pair sits at the even coordinate, and the 2x2x2 box the
    absorb step after moves along several axes."""
    import itertools
    rng = np.random.default_rng(9)
    T = models.ising3d_tensor(0.3)
    T0 = T + 0.4 * rng.standard_normal(T.shape)
    Tn = 1.5 * T + 0.4 * rng.standard_normal(T.shape)
    sites = list(itertools.product(*[range(L) for L in shape]))
    Z = hq.network_value({s: T for s in sites}, shape)
    steps = {"X": shape[0].bit_length() - 1, "Y": shape[1].bit_length() - 1, "Z": shape[2].bit_length() - 1}
    for r0, rn in pairs:
        Zi = hq.network_value({s: (T0 if s == r0 else Tn if s == rn else T) for s in sites}, shape)
        _, _, G = hq.two_point_function({"Tensor": T, "loc": {}, "factor": {}},
                                        {"Tensor": T0, "loc": dict(zip("XYZ", r0)), "factor": {}},
                                        {"Tensor": Tn, "loc": dict(zip("XYZ", rn)), "factor": {}},
                                        16, dict(steps))
        assert abs(G - Zi / Z) < 1e-11 * abs(Zi / Z), (shape, r0, rn, G, Zi / Z)


@pytest.mark.parametrize("shape,pairs", [
    ((4, 1, 1), [((1, 0, 0), (2, 0, 0)), ((0, 0, 0), (3, 0, 0))]),
    ((1, 4, 1), [((0, 1, 0), (0, 2, 0))]),
    ((1, 1, 4), [((0, 0, 2), (0, 0, 1))]),
    ((2, 2, 2), [((0, 0, 0), (1, 1, 1))]),
])
def test_atrg_two_point_equals_network_contraction(shape, pairs):
    """ATRG two-point flow (trg/ATRG_3d_new.py:1333-1489 with the one-point
    pairing in the two-point step) untruncated: G = Z(T0 at r0, Tn at rn) / Z
    of the box's network, with two different non-symmetric impurities (the
    rings of 4 fix which tensor of a pair sits at the even coordinate)."""
    import itertools
    from oracle import HOTRG_3d_QR as hq
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
            T0 = models.to_atrg(I0, 64)
            T0.loc = dict(zip("XYT", r0))
            Tn = models.to_atrg(In, 64)
            Tn.loc = dict(zip("XYT", rn))
            _, _, lnz, G = at.two_point_function(models.to_atrg(T6, 64), T0, Tn, 64, dict(steps))
            assert abs(G - Zi / Z) < 1e-11 * abs(Zi / Z), (shape, r0, rn, G, Zi / Z)
            assert abs(lnz - np.log(Z) / len(sites)) < 1e-12
    finally:
        at.set_rsvd_params()


@pytest.mark.parametrize("h", [0.0, 0.15])
def test_ising3d_field_tensors_against_enumeration(h):
    """The field form of the 3D Ising initial tensors (measurement/Ising_model_3d.py:15-79):
    untruncated HOTRG_3d_QR on the 2x2x2 box gives ln Z/V and the one-point value
    <s> = P of the impurity flow equal to the enumeration with field h."""
    beta = 0.2
    T, Timp = models.ising3d_field_tensors(beta, h)
    lnz_ref, m_ref = exact.brute_ising_field(beta, h)
    lnz, _ = hq.ln_Z_over_V(T, 16, {"X": 1, "Y": 1, "Z": 1}, "QR")
    assert abs(lnz - lnz_ref) < 1e-12
    if h == 0.0:
        assert abs(m_ref) < 1e-14
        assert abs(lnz - exact.brute_lnZ_ising3d(beta)) < 1e-12
        return
    _, _, P = hq.one_point_function({"Tensor": T, "loc": {}, "factor": {}},
                                    {"Tensor": Timp, "loc": {"X": 1, "Y": 0, "Z": 1}, "factor": {}}, 16,
                                    {"X": 1, "Y": 1, "Z": 1}, "QR")
    assert abs(P - m_ref) < 1e-11 * abs(m_ref)
