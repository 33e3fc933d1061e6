"""ATRG diagnostics (trg/ATRG_3d_new.py:768-951), the single-impurity one-point
flow (:1191-1331) and the per-step output files (:197-212, :736-766,
:991-1025, Ising_model_3d.py:40-42).

The oracle's diagnostics are pinned on the CPU: find_max against a brute-force
sort of the full tensor, the diagonal and the density matrices against
einsums of the full tensor, and the scaling dimensions / correlation lengths
against a product tensor whose density matrix is known in closed form.  The
device versions are then compared with the oracle on a coarse-grained tensor.
"""
import os

import numpy as np
import pytest

from oracle import ATRG_3d_new as oat
from oracle import exact, models
from oracle.tensor_class import ATRG_Tensor as OTensor


def _full(T):
    return np.einsum("TXYi,i,itxy->TXYtxy", T.U, T.s, T.VH)


def _coarse_potts(chi=6, beta=0.55):
    """A Potts tensor after one T and one X step of the oracle (identity leg-swap
    basis, converged range finder)."""
    oat.set_rsvd_params(k_factor=2, p=0, q=20, seed=1)
    oat.set_leg_swap_basis("identity")
    try:
        T = models.to_atrg(models.potts3d_tensor(beta), chi)
        for d in "TX":
            T = oat.atrg_pure_tensor(T, chi, d)
            T, _ = oat.normalization(T)
    finally:
        oat.set_leg_swap_basis("svd")
        oat.set_rsvd_params()
    return T


def _product_tensor(A, B, C, Dcut=6):
    """T_{TXYtxy} = A_{Tt} B_{Xx} C_{Yy}: its density matrices are tr(B) tr(C) A,
    tr(A) tr(C) B and tr(A) tr(B) C."""
    full = np.einsum("Tt,Xx,Yy->TXYtxy", A, B, C)
    dims = A.shape[0], B.shape[0], C.shape[0]
    n = int(np.prod(dims))
    u, s, vh = np.linalg.svd(full.reshape(n, n))
    return OTensor(u.reshape(*dims, -1), s, vh.reshape(-1, *dims), Dcut, 3, False, {})


def _sym(e, seed):
    Q = np.linalg.qr(np.random.default_rng(seed).standard_normal((len(e), len(e))))[0]
    return (Q * np.asarray(e)) @ Q.T


# ---------------------------------------------------------------------------
# oracle pins (CPU)
# ---------------------------------------------------------------------------
def test_oracle_find_max_is_the_top_of_the_full_tensor():
    T = _coarse_potts(4)
    full = _full(T)
    v, ind = oat.find_max(T, 7)
    ref = np.sort(np.abs(full).ravel())[::-1][:7]
    assert np.allclose(np.abs(v), ref, rtol=1e-13)
    assert all(abs(full[tuple(i)] - x) < 1e-13 * abs(x) for i, x in zip(ind, v))


def test_oracle_diagonal_and_density_matrices_match_einsum():
    T = _coarse_potts(4)
    full = _full(T)
    assert np.allclose(oat.tensor_diagonal(T), np.einsum("ijkijk->ijk", full).ravel(), atol=1e-13)
    assert np.allclose(oat.find_1stto3rdmax(T), sorted(np.einsum("ijkijk->ijk", full).ravel(), key=abs)[::-1][:3],
                       rtol=1e-13)
    assert np.allclose(oat.density_matrix(T, "t"), np.einsum("TXYtXY->Tt", full), atol=1e-12)
    assert np.allclose(oat.density_matrix(T, "x"), np.einsum("TXYTxY->Xx", full), atol=1e-12)
    assert np.allclose(oat.density_matrix(T, "y"), np.einsum("TXYTXy->Yy", full), atol=1e-12)
    assert abs(np.trace(oat.density_matrix(T, "t")) - np.einsum("TXYTXY", full)) < 1e-12


def test_oracle_scaling_dimensions_and_correlation_closed_form():
    a, b, c = [2.0, 0.5, 0.125, -0.01], [1.0, 0.3], [1.0, 0.6, 0.2]
    T = _product_tensor(_sym(a, 3), _sym(b, 4), _sym(c, 5), Dcut=6)
    lam = np.sum(b) * np.sum(c) * np.array(a)
    lam = np.concatenate([np.maximum(lam, 1e-16), [1e-16, 1e-16]])      # floored, padded to Dcut
    assert np.allclose(oat.cal_scaling_dimensions(T), np.log(lam[0] / lam) / (2 * np.pi), rtol=1e-10, atol=1e-12)
    ev, lnxi = oat.cal_coorelation(T, {"T": 3, "X": 1, "Y": 2})
    assert np.allclose(ev["t"], np.abs(a + [0, 0]) / 2.0, atol=1e-13)
    assert np.allclose(ev["x"], b + [0] * 4, atol=1e-13)
    assert np.allclose(ev["y"], c + [0] * 3, atol=1e-13)
    for leg, n, e1 in (("t", 3, 0.25), ("x", 1, 0.3), ("y", 2, 0.6)):
        assert abs(lnxi[leg] - (n * np.log(2) - np.log(-1 / np.log(e1)))) < 1e-12


# ---------------------------------------------------------------------------
# device versus oracle
# ---------------------------------------------------------------------------
def _device_tensor(t, chi, loc=None):
    from trg_backup_amd.tensor_class.tensor_class import ATRG_Tensor
    return ATRG_Tensor(t.U, t.s, t.VH, chi, 3, loc is not None, dict(loc or {}))


@pytest.mark.gpu
def test_atrg_diagnostics_match_oracle(gpu):
    from trg_backup_amd.trg import ATRG_3d_new as at
    T = _coarse_potts(6)
    Td = _device_tensor(T, 6)
    v, ind = at.find_max(Td, 5)
    v0, ind0 = oat.find_max(T, 5)
    assert np.allclose(v, v0, rtol=1e-12) and np.array_equal(ind, ind0)
    assert np.allclose(at.find_1stto3rdmax(Td), oat.find_1stto3rdmax(T), rtol=1e-12)
    assert np.allclose(at.cal_scaling_dimensions(Td), oat.cal_scaling_dimensions(T), rtol=1e-9, atol=1e-12)
    at.rgstep.update({"X": 1, "Y": 0, "T": 1})
    try:
        ev, lnxi = at.cal_coorelation(Td)
    finally:
        at.rgstep.update({"X": 0, "Y": 0, "T": 0})
    ev0, lnxi0 = oat.cal_coorelation(T, {"X": 1, "Y": 0, "T": 1})
    for leg in "txy":
        assert np.allclose(ev[leg], ev0[leg], rtol=1e-9, atol=1e-13)
        assert abs(lnxi[leg] - lnxi0[leg]) < 1e-9


@pytest.mark.gpu
def test_atrg_diagnostics_closed_form(gpu):
    """The product tensor A (x) B (x) C on the device: scaling dimensions and
    correlation spectra from the closed-form density matrices."""
    from trg_backup_amd.trg import ATRG_3d_new as at
    a, b, c = [2.0, 0.5, 0.125, -0.01], [1.0, 0.3], [1.0, 0.6, 0.2]
    T = _product_tensor(_sym(a, 3), _sym(b, 4), _sym(c, 5), Dcut=6)
    Td = _device_tensor(T, 6)
    lam = np.sum(b) * np.sum(c) * np.array(a)
    lam = np.concatenate([np.maximum(lam, 1e-16), [1e-16, 1e-16]])
    assert np.allclose(at.cal_scaling_dimensions(Td), np.log(lam[0] / lam) / (2 * np.pi), rtol=1e-10, atol=1e-12)
    ev, _ = at.cal_coorelation(Td)
    assert np.allclose(ev["t"], np.abs(a + [0, 0]) / 2.0, atol=1e-13)
    assert np.allclose(ev["y"], c + [0] * 3, atol=1e-13)


@pytest.mark.gpu
def test_one_point_function1_equals_network_contraction(gpu):
    """The single-impurity flow (:1246-1331), untruncated on the 2x2x2 box:
    P = Z[impurity at one site] / Z of the eight-tensor network and ln Z/V is
    the enumeration; the per-step field file has one line per step."""
    from oracle import HOTRG_3d_QR as hq
    from trg_backup_amd.trg import ATRG_3d_new as at
    rng = np.random.default_rng(8)
    T6 = models.ising3d_tensor(0.22)
    Ti = rng.standard_normal((2,) * 6) + 1.5 * T6
    sites = [(i, j, k) for i in range(2) for j in range(2) for k in range(2)]
    Z = hq.network_value({s: T6 for s in sites}, (2, 2, 2))
    loc = (1, 0, 1)
    ref = hq.network_value({s: (Ti if s == loc else T6) for s in sites}, (2, 2, 2)) / Z
    at.set_rsvd_params(k_factor=3, p=0, q=2)
    try:
        T = _device_tensor(models.to_atrg(T6, 64), 64)
        T0 = _device_tensor(models.to_atrg(Ti, 64), 64, {"X": loc[0], "Y": loc[1], "T": loc[2]})
        T, T0 = at.one_point_function1(T, T0, 64, {"X": 1, "Y": 1, "T": 1})
        lnz, P = at.cal_field_expected_value1(T, T0, {"X": 1, "Y": 1, "T": 1})
    finally:
        at.set_rsvd_params()
    print(f"\n[atrg 1pt single] P={P!r} ref={ref!r}", flush=True)
    assert abs(lnz - exact.brute_lnZ_ising3d(0.22)) < 1e-12
    assert abs(P - ref) < 1e-10 * abs(ref)
    assert T0.loc == {"X": 0, "Y": 0, "T": 0}


@pytest.mark.gpu
def test_flows_write_the_reference_step_files(gpu, tmp_path, monkeypatch):
    """OUTPUT_DIR files of the pure flow (X.dat, tensor_sv/, leg_exchange_pure/)
    and of the single-impurity flow (field_expected_value_rgsteps_dependence.dat)."""
    from trg_backup_amd.trg import ATRG_3d_new as at
    monkeypatch.setenv("OUTPUT_DIR", str(tmp_path))
    Td = at.pure_tensor(_device_tensor(models.to_atrg(models.potts3d_tensor(0.5), 6), 6), 6,
                        {"X": 1, "Y": 1, "T": 1})
    xs = (tmp_path / "X.dat").read_text().splitlines()
    assert len(xs) == 4 and all(len(l.split()) == 3 for l in xs)
    assert np.allclose([float(v) for v in xs[-1].split()], [float(v) for v in Td.cal_X()], rtol=1e-12)
    steps = [(0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1)]
    assert sorted(os.listdir(tmp_path / "tensor_sv")) == sorted(
        f"tensor_sv_lx{x}_ly{y}_lt{t}.dat" for x, y, t in steps)
    last = (tmp_path / "tensor_sv" / "tensor_sv_lx1_ly1_lt1.dat").read_text().splitlines()
    s = np.asarray(Td.s.double().cpu() if hasattr(Td.s, "cpu") else Td.s, dtype=float)
    assert last[0] == f"#s1={s.max():.12e}" and len(last) == 1 + max(6, len(s))
    assert abs(float(last[1]) - 1.0) < 1e-15
    assert sorted(os.listdir(tmp_path / "leg_exchange_pure")) == sorted(
        f"leg_exchange_sv_lx{x}_ly{y}_lt{t}.dat" for x, y, t in steps[1:])
    leg = (tmp_path / "leg_exchange_pure" / "leg_exchange_sv_lx1_ly1_lt1.dat").read_text().splitlines()
    assert leg[0] == f"#s1={1.0:.12e}" and len(leg) == 1 + 3 * 6

    T6 = models.potts3d_tensor(0.5)
    T = _device_tensor(models.to_atrg(T6, 6), 6)
    T0 = _device_tensor(models.to_atrg(T6, 6), 6, {"X": 1, "Y": 0, "T": 1})
    at.one_point_function1(T, T0, 6, {"X": 1, "Y": 1, "T": 1})
    lines = (tmp_path / "field_expected_value_rgsteps_dependence.dat").read_text().splitlines()
    assert [int(l.split()[0]) for l in lines] == [0, 1, 2, 3]
    assert all(abs(float(l.split()[2]) - 1.0) < 1e-6 for l in lines)     # the impurity is the pure tensor


def test_result_line_format(tmp_path):
    from trg_backup_amd.measurement import Ising_model_3d as meas
    f = tmp_path / "result.dat"
    line = meas.write_result(str(f), 0.22, 0.0, 0.7812345678901234, 0.123)
    assert f.read_text() == line + "\n"
    assert line == "0.22 0.0  7.812345678901e-01  1.230000000000e-01"


# ---------------------------------------------------------------------------
# HOTRG 3D diagnostics (trg/HOTRG_3d_QR.py:559-713, trg/HOTRG_3d.py:240-245)
# ---------------------------------------------------------------------------
from oracle import HOTRG_3d_QR as oq  # noqa: E402


def _product6(A, B, C):
    """T[x,X,y,Y,t,T] = A[x,X] B[y,Y] C[t,T]."""
    return np.einsum("xX,yY,tT->xXyYtT", A, B, C)


def _qr_tensor(chi=4, beta=0.21):
    """A 3D Ising tensor after a Z and an X step of the QR oracle."""
    T = models.ising3d_tensor(beta)
    for d in "ZX":
        T = oq.new_pure_tensor(T, chi, d)
        T, _ = oq.normalization(T)
    return T


def test_oracle_qr_diagnostics_closed_form():
    a, b, c = [2.0, 0.5, -0.25], [1.0, 0.3], [1.0, 0.6, 0.2, 0.05]
    A, B, C = _sym(a, 6), _sym(b, 7), _sym(c, 8)
    T = _product6(A, B, C)
    ta, tb, tc = sum(a), sum(b), sum(c)
    Xt, Xx, Xy = oq.cal_X(T)
    assert np.allclose([Xt, Xx, Xy], [tc ** 2 / np.sum(np.square(c)), ta ** 2 / np.sum(np.square(a)),
                                      tb ** 2 / np.sum(np.square(b))], rtol=1e-12)
    d = np.einsum("x,y,t->xyt", np.diag(A), np.diag(B), np.diag(C)).ravel()
    assert np.allclose(oq.find_1stto3rdmax(T), sorted(d, key=abs)[::-1][:3], rtol=1e-12)
    lam = np.concatenate([ta * tb * np.array(c), [1e-16] * 2])                # padded to Dcut = 6
    assert np.allclose(oq.cal_scaling_dimensions(T, 6), np.log(lam[0] / lam) / (2 * np.pi), rtol=1e-10)
    ev, lnxi = oq.cal_coorelation(T, 3, {"X": 2, "Y": 1, "Z": 3})
    assert np.allclose(ev["t"], np.array(c + [0] * 5), atol=1e-13)            # padded to Dcut^2 = 9
    assert np.allclose(ev["x"], np.abs(a + [0] * 6) / 2.0, atol=1e-13)
    for leg, n, e1 in (("t", 3, 0.6), ("x", 2, 0.5 / 2.0), ("y", 1, 0.3)):
        assert abs(lnxi[leg] - (n * np.log(2) - np.log(-1 / np.log(e1)))) < 1e-12
    s, smax = oq.tensor_spectrum(T, 2)
    sv = np.sort(np.abs(np.einsum("i,j,k->ijk", a, b, c)).ravel())[::-1]
    assert abs(smax - sv[0]) < 1e-12 and np.allclose(s, sv[:4] / sv[0], atol=1e-13)


def test_oracle_qr_diagnostics_match_einsum():
    T = _qr_tensor(4)
    tr = np.einsum("xxyytt", T)
    assert abs(oq.cal_X(T)[0] - tr ** 2 / np.einsum("xxyytT,aabbTt", T, T)) < 1e-12
    D = np.einsum("xxyytT->tT", T)
    assert np.allclose(oq.density_matrix(T, "t"), D.T, atol=1e-14)
    assert abs(np.trace(oq.density_matrix(T, "x")) - tr) < 1e-12
    assert abs(np.trace(oq.density_matrix(T, "y")) - tr) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["product", "coarse"])
def test_qr_diagnostics_match_oracle(gpu, case):
    """Device cal_X / find_1stto3rdmax / cal_scaling_dimensions / cal_coorelation
    and the tensor spectrum against the oracle; the product tensor uses a
    non-symmetric t factor, so the density matrix's transpose and the lower
    triangle the eigensolver reads are pinned."""
    import torch

    from trg_backup_amd.trg import HOTRG_3d as h3
    from trg_backup_amd.trg import HOTRG_3d_QR as hq
    if case == "product":
        rng = np.random.default_rng(9)
        T = _product6(_sym([2.0, 0.5, -0.25], 6), _sym([1.0, 0.3], 7), rng.standard_normal((4, 4)) + 2 * np.eye(4))
        Dcut = 4
    else:
        T = _qr_tensor(6)
        Dcut = 6
    Td = {"Tensor": torch.as_tensor(T).to(gpu)}
    assert np.allclose(hq.cal_X(Td), oq.cal_X(T), rtol=1e-12)
    assert np.allclose(hq.find_1stto3rdmax(Td), oq.find_1stto3rdmax(T), rtol=1e-12)
    assert np.allclose(hq.cal_scaling_dimensions(Td, Dcut), oq.cal_scaling_dimensions(T, Dcut), rtol=1e-9,
                       atol=1e-12)
    hq.rgstep.update({"X": 1, "Y": 2, "Z": 1})
    try:
        ev, lnxi = hq.cal_coorelation(Td, Dcut)
    finally:
        hq.rgstep.update({"X": 0, "Y": 0, "Z": 0})
    ev0, lnxi0 = oq.cal_coorelation(T, Dcut, {"X": 1, "Y": 2, "Z": 1})
    for leg in "txy":
        assert np.allclose(ev[leg], ev0[leg], rtol=1e-9, atol=1e-13)
        assert abs(lnxi[leg] - lnxi0[leg]) < 1e-9
    s, smax = h3.__get_tensor_spectrum(Td["Tensor"], Dcut)
    s0, smax0 = oq.tensor_spectrum(T, Dcut)
    assert abs(smax - smax0) < 1e-12 * smax0 and np.allclose(s, s0, rtol=1e-8, atol=1e-12)


@pytest.mark.gpu
def test_qr_pure_tensor_writes_the_reference_step_files(gpu, tmp_path, monkeypatch):
    """With OUTPUT_DIR the QR flow writes X.dat, lnξ*.dat, scaling_dimensions.dat and
    density_matrix_t/ for the initial and every coarse-grained tensor, and its
    factors equal those of the flow without files."""
    from trg_backup_amd.trg import HOTRG_3d_QR as hq
    T6 = models.ising3d_tensor(0.21)
    steps = {"X": 1, "Y": 1, "Z": 1}
    ref = hq.pure_tensor({"Tensor": T6.copy(), "factor": {}}, 4, steps)
    monkeypatch.setenv("OUTPUT_DIR", str(tmp_path))
    out = hq.pure_tensor({"Tensor": T6.copy(), "factor": {}}, 4, steps)
    assert out["factor"] == ref["factor"]
    for name in ("X.dat", "lnξt.dat", "lnξx.dat", "lnξy.dat", "scaling_dimensions.dat"):
        assert len((tmp_path / name).read_text().splitlines()) == 4, name
    sd = (tmp_path / "scaling_dimensions.dat").read_text().splitlines()
    assert all(len(l.split()) == 4 for l in sd)
    assert sorted(os.listdir(tmp_path / "density_matrix_t")) == sorted(
        f"DentMatT_lx{x}_ly{y}_lt{z}.dat" for x, y, z in [(0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1)])
    last = np.array([float(v) for v in (tmp_path / "X.dat").read_text().splitlines()[-1].split()])
    assert np.allclose(last, hq.cal_X(out), rtol=1e-12)
