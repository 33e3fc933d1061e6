"""initial_tensor_for_ATRG (tensor_init/ATRG_init.py:7-254).

The oracle is pinned against the dense truncated SVD of M[(T X..), (t x..)]
built by one einsum of the same factors: with a sketch at least as wide as
the number of quadrature points (the rank of M) the randomized SVD is exact.
The device version is compared with the oracle and with the dense SVD.
"""
import numpy as np
import pytest

from oracle import ATRG_init as oinit


def _factors(dim, chi, na, seed):
    rng = np.random.default_rng(seed)
    As, At = rng.standard_normal((na, chi)), rng.standard_normal((na, chi))
    Bs, Bt = rng.standard_normal((chi, na)), rng.standard_normal((chi, na))
    w = rng.uniform(0.1, 1.0, na) * np.exp(-np.arange(na) / 3.0)       # decaying weights
    J = np.full(na, 0.5)
    return dict(dim=dim, J=J, w=w, As=As, Bs=Bs, At=At, Bt=Bt)


def _dense(f, k):
    M = oinit.dense_matrix(**f)
    u, s, vh = np.linalg.svd(M)
    return M, u[:, :k], s[:k], vh[:k]


def _rebuild(u, s, vh):
    n = s.shape[0]
    return (u.reshape(-1, n) * s) @ vh.reshape(n, -1)


@pytest.mark.parametrize("dim,chi,na", [(2, 5, 12), (3, 3, 10), (4, 2, 7)])
def test_oracle_exact_when_the_sketch_covers_the_rank(dim, chi, na):
    f = _factors(dim, chi, na, 1)
    k = min(6, na)
    u, s, vh = oinit.initial_tensor_for_ATRG(**f, k=k, p=na, q=1, seed=3)
    M, u0, s0, vh0 = _dense(f, k)
    assert u.shape == (chi,) * dim + (k,) and vh.shape == (k,) + (chi,) * dim
    assert np.allclose(s, s0, rtol=1e-12)
    assert np.allclose(_rebuild(u, s, vh), (u0 * s0) @ vh0, atol=1e-12 * s0[0])
    uf = u.reshape(-1, k)
    assert np.allclose(uf.T @ uf, np.eye(k), atol=1e-13)


def test_oracle_index_strings_are_the_dense_operator():
    """M applied through the reference's three index strings equals the dense M."""
    f = _factors(3, 3, 8, 2)
    M = oinit.dense_matrix(**f)
    ty, tdq, qdt = oinit._specs(3)
    A, B = oinit._ops(3, f["At"], f["As"], f["Bt"], f["Bs"])
    Y = np.random.default_rng(5).standard_normal((3, 3, 3, 4))
    assert np.allclose(np.einsum(ty, f["w"], f["J"], *A, *B, Y).reshape(27, 4), M @ Y.reshape(27, 4))
    assert np.allclose(np.einsum(tdq, f["w"], f["J"], *B, *A, Y).reshape(27, 4), M.T @ Y.reshape(27, 4))
    assert np.allclose(np.einsum(qdt, f["w"], f["J"], *B, *A, Y).reshape(4, 27), Y.reshape(27, 4).T @ M)


def test_oracle_power_iterations_reach_the_dense_spectrum():
    """A sketch narrower than the rank: with q rounds the leading values converge."""
    f = _factors(3, 4, 40, 4)
    u, s, vh = oinit.initial_tensor_for_ATRG(**f, k=4, p=8, q=6, seed=1)
    _, _, s0, _ = _dense(f, 4)
    assert np.allclose(s, s0, rtol=1e-8)


@pytest.mark.gpu
@pytest.mark.parametrize("dim,chi,na,p,q", [(3, 4, 12, 12, 1), (3, 6, 40, 24, 4), (2, 8, 20, 20, 2)])
def test_device_matches_oracle_and_dense(gpu, dim, chi, na, p, q):
    from trg_backup_amd.tensor_init import ATRG_init as dinit
    f = _factors(dim, chi, na, 7)
    k = 6
    u, s, vh = dinit.initial_tensor_for_ATRG(**f, k=k, p=p, q=q, seed=11)
    uo, so, vho = oinit.initial_tensor_for_ATRG(**f, k=k, p=p, q=q, seed=11)
    assert isinstance(u, np.ndarray) and u.shape == uo.shape and vh.shape == vho.shape
    assert np.allclose(s, so, rtol=1e-11)
    assert np.allclose(_rebuild(u, s, vh), _rebuild(uo, so, vho), atol=1e-11 * so[0])
    if k + p >= na:
        M, u0, s0, vh0 = _dense(f, k)
        assert np.allclose(s, s0, rtol=1e-11)


@pytest.mark.gpu
def test_device_rejects_complex_and_wide_sketches(gpu):
    from trg_backup_amd._lib import TrgError
    from trg_backup_amd.tensor_init import ATRG_init as dinit
    f = _factors(3, 3, 6, 1)
    with pytest.raises(TrgError):
        dinit.initial_tensor_for_ATRG(**f, k=64, p=65, q=0)
    g = dict(f, As=f["As"] * (1 + 0.5j))
    with pytest.raises(TrgError):
        dinit.initial_tensor_for_ATRG(**g, k=4, p=4, q=0)
