"""GPU parity of HOTRG 3D with squeezers (trg/HOTRG_3d_QR.py) against the oracle."""
import numpy as np
import pytest
import torch

from oracle import HOTRG_3d_QR as oq
from oracle import exact, models

pytestmark = pytest.mark.gpu

KINDS = ["original", "QR", "fix"]


def _dev(x, gpu):
    return torch.as_tensor(np.ascontiguousarray(x)).to(gpu)


@pytest.mark.parametrize("kind", KINDS)
def test_hotrg3d_qr_untruncated_equals_brute_force(gpu, kind):
    from trg_backup_amd import hotrg_qr
    beta = 0.3
    v = hotrg_qr.ln_z_over_v(_dev(models.ising3d_tensor(beta), gpu)[None], 16, {"X": 1, "Y": 1, "Z": 1}, kind)
    assert abs(float(v[0].item()) - exact.brute_lnZ_ising3d(beta)) < 1e-12


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("model,beta,chi", [("ising", 0.22, 4), ("ising", 0.23, 6), ("potts", 0.55, 5)])
def test_hotrg3d_qr_matches_oracle(gpu, kind, model, beta, chi):
    """Six steps (Z, X, Y twice) with truncation.  DGENERACY_EPS = 1e-8 on
    both sides: with 0 the rank rule can cut inside an exactly degenerate
    multiplet (Z2 / S3 symmetry), and which member survives is an arbitrary
    choice of the eigensolver (Jacobi here, LAPACK in the oracle)."""
    from trg_backup_amd import hotrg_qr
    T6 = models.ising3d_tensor(beta) if model == "ising" else models.potts3d_tensor(beta)
    steps = {"X": 2, "Y": 2, "Z": 2}
    ref, To = oq.ln_Z_over_V(T6, chi, steps, kind, degeneracy_eps=1e-8)
    groups = hotrg_qr.renorm(_dev(T6, gpu)[None], chi, steps, kind, 1e-8)
    (_, Tn, lnf), = groups
    v = float(hotrg_qr.ln_z_over_v(_dev(T6, gpu)[None], chi, steps, kind, 1e-8)[0].item())
    print(f"\n[hotrg3d_qr {kind} {model} beta={beta} chi={chi}] gpu={v!r} oracle={ref!r} diff={v - ref:.3e}",
          flush=True)
    assert tuple(Tn.shape[1:]) == To["Tensor"].shape
    assert np.allclose(lnf[0].cpu().numpy(), list(To["factor"].values()), rtol=1e-11, atol=1e-14)
    assert abs(v - ref) < 1e-11 * abs(ref)


def test_hotrg3d_qr_batched_equals_alone_and_splits_by_rank(gpu):
    """A batch whose kept ranks differ (beta = 0 is a rank-one tensor, so the
    rank rule keeps one state) splits into groups; each instance's result is
    bitwise the one it gets alone, and beta = 0 gives ln 2."""
    from trg_backup_amd import hotrg_qr
    betas = [0.22, 0.0, 0.23, 0.21]
    steps = {"X": 1, "Y": 1, "Z": 2}
    T = torch.stack([_dev(models.ising3d_tensor(b), gpu) for b in betas])
    zb = hotrg_qr.ln_z_over_v(T, 5, steps, "QR", 1e-8).cpu().numpy()
    groups = hotrg_qr.renorm(T, 5, steps, "QR", 1e-8)
    assert len(groups) == 2 and sorted(i for g in groups for i in g[0]) == [0, 1, 2, 3]
    for i, b in enumerate(betas):
        zi = hotrg_qr.ln_z_over_v(T[i:i + 1].contiguous(), 5, steps, "QR", 1e-8).cpu().numpy()
        assert zi[0] == zb[i], (i, zi[0], zb[i])
    assert abs(zb[1] - np.log(2)) < 1e-13


def test_hotrg3d_qr_dropin_pure_tensor(gpu):
    """The drop-in's dict flow: the reference's "(X,Y,Z)" factor keys and values."""
    from trg_backup_amd.trg import HOTRG_3d_QR as dq
    old = (dq.select_squeezer, dq.degeneracy_eps)
    dq.select_squeezer, dq.degeneracy_eps = "QR", 1e-8
    try:
        T0 = models.ising3d_tensor(0.22)
        steps = {"X": 1, "Y": 1, "Z": 2}
        out = dq.pure_tensor({"Tensor": T0, "loc": {}, "factor": {}}, 4, dict(steps))
        ref = oq.pure_tensor({"Tensor": T0, "loc": {}, "factor": {}}, 4, dict(steps), "QR", 1e-8)
        assert list(out["factor"]) == list(ref["factor"])
        assert np.allclose(list(out["factor"].values()), list(ref["factor"].values()), rtol=1e-11, atol=1e-14)
        assert isinstance(out["Tensor"], np.ndarray) and out["Tensor"].shape == ref["Tensor"].shape
        Px, PX, Py, PY, TdagT = dq.squeezer_QR(T0, T0, T0, T0, 3)
        k = Px.shape[0]
        assert np.allclose(Px.reshape(k, -1) @ PX.reshape(-1, k), np.eye(k), atol=1e-10)
        cg = dq.coarse_graining(T0, T0, Px, PX, Py, PY)
        assert np.allclose(cg, oq.coarse_graining_einsum(T0, T0, Px, PX, Py, PY), rtol=1e-12, atol=1e-12)
    finally:
        dq.select_squeezer, dq.degeneracy_eps = old


@pytest.mark.parametrize("kind", KINDS)
def test_hotrg3d_qr_random_tensor_matches_oracle(gpu, kind):
    """A random positive tensor with no x/X or y/Y symmetry, where the three
    squeezers give different results (for the symmetric model tensors they
    coincide), so each squeezer's choice of environments is pinned."""
    from trg_backup_amd import hotrg_qr
    rng = np.random.default_rng(7)
    T6 = rng.uniform(0.1, 1.0, (3,) * 6)
    steps = {"X": 1, "Y": 1, "Z": 2}
    ref, _ = oq.ln_Z_over_V(T6, 5, steps, kind, 0.0)
    v = float(hotrg_qr.ln_z_over_v(_dev(T6, gpu)[None], 5, steps, kind, 0.0)[0].item())
    print(f"\n[hotrg3d_qr random {kind}] gpu={v!r} oracle={ref!r} diff={v - ref:.3e}", flush=True)
    assert abs(v - ref) < 1e-11 * abs(ref)


@pytest.mark.parametrize("chi,steps", [(16, {"X": 1, "Y": 1, "Z": 1}), (4, {"X": 2, "Y": 1, "Z": 2})])
def test_hotrg3d_qr_one_point_matches_oracle(gpu, chi, steps):
    """one_point_function (trg/HOTRG_3d_QR.py:859-908) with a random impurity at
    (1, 0, 1): the one-point value and ln Z/V against the oracle; untruncated
    (chi = 16, 2x2x2) also against the direct network contraction."""
    import itertools
    from trg_backup_amd.trg import HOTRG_3d_QR as dq
    old = (dq.select_squeezer, dq.degeneracy_eps)
    dq.select_squeezer, dq.degeneracy_eps = "QR", 1e-8
    try:
        T = models.ising3d_tensor(0.3)
        T0 = T + 0.3 * np.random.default_rng(4).standard_normal(T.shape)
        loc = {"X": 1, "Y": 0, "Z": 1}
        Td, T0d = dq.one_point_function({"Tensor": T, "loc": {}, "factor": {}},
                                        {"Tensor": T0, "loc": dict(loc), "factor": {}}, chi, dict(steps))
        n = sum(steps.values())
        lnz, P = dq.cal_field_expected_value(Td, [T0d], steps, False)
        _, _, Pref = oq.one_point_function({"Tensor": T, "loc": {}, "factor": {}},
                                           {"Tensor": T0, "loc": dict(loc), "factor": {}}, chi, dict(steps), "QR",
                                           1e-8)
        print(f"\n[one point chi={chi}] gpu={P[0]!r} oracle={Pref!r}", flush=True)
        assert abs(P[0] - Pref) < 1e-10 * abs(Pref)
        assert T0d["loc"] == {d: loc[d] >> steps[d] for d in "XYZ"}
        if chi == 16:
            sites = list(itertools.product(range(2), range(2), range(2)))
            Z = oq.network_value({s: T for s in sites}, (2, 2, 2))
            Zi = oq.network_value({s: (T0 if s == (1, 0, 1) else T) for s in sites}, (2, 2, 2))
            assert abs(P[0] - Zi / Z) < 1e-10 * abs(Zi / Z)
            assert abs(lnz - np.log(Z) / 8) < 1e-12
    finally:
        dq.select_squeezer, dq.degeneracy_eps = old


@pytest.mark.parametrize("shape,pairs", [
    ((4, 1, 1), [((1, 0, 0), (2, 0, 0)), ((0, 0, 0), (3, 0, 0))]),
    ((1, 4, 1), [((0, 1, 0), (0, 2, 0))]),
    ((1, 1, 4), [((0, 0, 3), (0, 0, 0))]),
    ((2, 2, 2), [((0, 0, 0), (1, 1, 1))]),
])
def test_hotrg3d_qr_two_point_equals_network_contraction(gpu, shape, pairs):
    """two_point_function (trg/HOTRG_3d_QR.py:910-1033) untruncated on small
    boxes with two different non-symmetric impurities: the correlator equals
    the direct network contraction (the rings of 4 fix which tensor of a pair
    sits at the even coordinate), and matches the oracle."""
    import itertools
    from trg_backup_amd.trg import HOTRG_3d_QR as dq
    rng = np.random.default_rng(9)
    T = models.ising3d_tensor(0.3)
    T0 = T + 0.4 * rng.standard_normal(T.shape)
    Tn = 1.5 * T + 0.4 * rng.standard_normal(T.shape)
    sites = list(itertools.product(*[range(L) for L in shape]))
    Z = oq.network_value({s: T for s in sites}, shape)
    steps = {"X": shape[0].bit_length() - 1, "Y": shape[1].bit_length() - 1, "Z": shape[2].bit_length() - 1}
    old = (dq.select_squeezer, dq.degeneracy_eps)
    dq.select_squeezer, dq.degeneracy_eps = "QR", 0.0
    try:
        for r0, rn in pairs:
            Zi = oq.network_value({s: (T0 if s == r0 else Tn if s == rn else T) for s in sites}, shape)
            Td, T0d = dq.two_point_function({"Tensor": T, "loc": {}, "factor": {}},
                                            {"Tensor": T0, "loc": dict(zip("XYZ", r0)), "factor": {}},
                                            {"Tensor": Tn, "loc": dict(zip("XYZ", rn)), "factor": {}},
                                            16, dict(steps))
            _, G = dq.cal_field_expected_value(Td, [T0d], steps, False)
            print(f"\n[two point {shape} {r0}-{rn}] gpu={G[0]!r} network={Zi / Z!r}", flush=True)
            assert abs(G[0] - Zi / Z) < 1e-10 * abs(Zi / Z)
    finally:
        dq.select_squeezer, dq.degeneracy_eps = old


def test_hotrg3d_qr_two_point_matches_oracle_truncated(gpu):
    """Truncated two-point flow (chi = 4, two steps per direction, impurities
    at (1,0,1) and (2,3,1)): the correlator and ln Z/V against the oracle."""
    from trg_backup_amd.trg import HOTRG_3d_QR as dq
    rng = np.random.default_rng(5)
    T = models.ising3d_tensor(0.25)
    T0 = T + 0.3 * rng.standard_normal(T.shape)
    Tn = T + 0.3 * rng.standard_normal(T.shape)
    r0, rn = {"X": 1, "Y": 0, "Z": 1}, {"X": 2, "Y": 3, "Z": 1}
    steps = {"X": 2, "Y": 2, "Z": 2}
    old = (dq.select_squeezer, dq.degeneracy_eps)
    dq.select_squeezer, dq.degeneracy_eps = "QR", 1e-8
    try:
        Td, T0d = dq.two_point_function({"Tensor": T, "loc": {}, "factor": {}},
                                        {"Tensor": T0, "loc": dict(r0), "factor": {}},
                                        {"Tensor": Tn, "loc": dict(rn), "factor": {}}, 4, dict(steps))
        lnz, G = dq.cal_field_expected_value(Td, [T0d], steps, False)
        To, _, Gref = oq.two_point_function({"Tensor": T, "loc": {}, "factor": {}},
                                            {"Tensor": T0, "loc": dict(r0), "factor": {}},
                                            {"Tensor": Tn, "loc": dict(rn), "factor": {}}, 4, dict(steps), "QR",
                                            1e-8)
    finally:
        dq.select_squeezer, dq.degeneracy_eps = old
    lnz_ref = sum(To["factor"].values()) + np.log(np.einsum("xxyyzz", To["Tensor"])) / 2 ** 6
    print(f"\n[two point chi=4] gpu={G[0]!r} oracle={Gref!r} lnz {lnz!r} vs {lnz_ref!r}", flush=True)
    assert abs(G[0] - Gref) < 1e-9 * abs(Gref)
    assert abs(lnz - lnz_ref) < 1e-12 * abs(lnz_ref)
