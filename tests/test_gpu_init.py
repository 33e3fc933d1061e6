"""GPU tests: initial-tensor drop-ins (measurement/Ising_model_3d.py, tensor_init/Potts_model.py)."""
import numpy as np
import pytest

from oracle import exact

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rgscheme", ["hotrg", "atrg"])
def test_ising3d_field_init_lnz_and_magnetisation(gpu, rgscheme, monkeypatch):
    """init_tensor / init_tensor_imp with a field, through the device flows
    untruncated on the 2x2x2 box: ln Z/V (ln_Z_over_V with RGSCHEME) and the
    magnetisation <s> from the one-point flow equal the enumeration."""
    from trg_backup_amd.measurement import Ising_model_3d as im
    beta, h = 0.2, 0.15
    lnz_ref, m_ref = exact.brute_ising_field(beta, h)
    monkeypatch.setenv("RGSCHEME", rgscheme)
    if rgscheme == "atrg":
        from trg_backup_amd.trg import ATRG_3d_new as at
        at.set_rsvd_params(k_factor=3, p=0, q=2)
    try:
        lnz = im.ln_Z_over_V(beta, h, 64, 1, 1, 1)
        assert abs(lnz - lnz_ref) < 1e-12, (lnz, lnz_ref)
        T, Timp = im.init_tensor_imp(beta, h, 64, rgscheme)
        if rgscheme == "hotrg":
            from trg_backup_amd.trg import HOTRG_3d_QR as q
            Td, T0d = q.one_point_function({"Tensor": T, "loc": {}, "factor": {}},
                                           {"Tensor": Timp, "loc": {"X": 1, "Y": 1, "Z": 0}, "factor": {}}, 16,
                                           {"X": 1, "Y": 1, "Z": 1})
            _, P = q.cal_field_expected_value(Td, [T0d], {"X": 1, "Y": 1, "Z": 1}, False)
        else:
            Timp.loc = {"X": 1, "Y": 1, "T": 0}
            T, (T0,) = at.one_point_function(T, [Timp], 64, {"X": 1, "Y": 1, "T": 1})
            _, P = at.cal_field_expected_value(T, [T0], {"X": 1, "Y": 1, "T": 1})
    finally:
        if rgscheme == "atrg":
            at.set_rsvd_params()
    print(f"\n[init {rgscheme}] lnz={lnz!r} ({lnz_ref!r}) <s>={P[0]!r} ({m_ref!r})", flush=True)
    assert abs(P[0] - m_ref) < 1e-10 * abs(m_ref)


def test_potts_infinite_density_init(gpu):
    """tensor_init/Potts_model.py infinite_density at h = mu = 0: the HOTRG and
    ATRG initial tensors through the device flows (untruncated boxes of 2 and
    4 sites) give the enumerated ln Z/V; a field raises."""
    from trg_backup_amd.tensor_init.Potts_model import infinite_density
    from trg_backup_amd.trg import ATRG_3d_new as at
    from trg_backup_amd.trg import HOTRG_3d_QR as q
    from trg_backup_amd import ops
    k = 0.55
    model = infinite_density(3, 243, 3, k, 0.0, 0.0)
    # HOTRG: one Z step joins the two sites of the (1,1,2) box (bond 9, no truncation)
    T = q.pure_tensor({"Tensor": model.cal_HOTRG_init_tensor(), "loc": {}, "factor": {}}, 9,
                      {"X": 0, "Y": 0, "Z": 1})
    v = sum(T["factor"].values()) + np.log(float(ops.trace_pairs(ops.contiguous(T["Tensor"][None]))[0].item())) / 2
    ref = exact.brute_lnZ_potts3d(k, shape=(1, 1, 2))
    assert abs(v - ref) < 1e-12, (v, ref)
    # ATRG on the (2,1,2) box (T then X step), sketch covering every rank
    at.set_rsvd_params(k_factor=1, p=0, q=2)
    try:
        va, _ = at.ln_Z_over_V(model.cal_ATRG_init_tensor(), 243, {"X": 1, "Y": 0, "T": 1})
    finally:
        at.set_rsvd_params()
    ref2 = exact.brute_lnZ_potts3d(k, shape=(2, 1, 2))
    assert abs(va - ref2) < 1e-12, (va, ref2)
    with pytest.raises(NotImplementedError):
        infinite_density(3, 64, 3, k, 0.1, 0.0).cal_HOTRG_init_tensor()


@pytest.mark.parametrize("rgscheme", ["hotrg", "atrg"])
def test_ising3d_magnetization_and_internal_energy(gpu, rgscheme, monkeypatch):
    """measurement/Ising_model_3d.py magnetization and internal_energy on the
    device, untruncated on the 2x2x2 box with a field: |<s>| and the
    nearest-neighbour <s_0 s_1> along t equal the enumeration."""
    from trg_backup_amd.measurement import Ising_model_3d as im
    beta, h = 0.2, 0.15
    lnz_ref, m_ref = exact.brute_ising_field(beta, h)
    c_ref = exact.brute_corr_ising_field(beta, h, (2, 2, 2), (0, 0, 0), (0, 0, 1))
    monkeypatch.setenv("RGSCHEME", rgscheme)
    Dcut = 16 if rgscheme == "hotrg" else 64
    if rgscheme == "atrg":
        from trg_backup_amd.trg import ATRG_3d_new as at
        at.set_rsvd_params(k_factor=3, p=0, q=2)
    try:
        lnz, m = im.magnetization(beta, h, Dcut, 1, 1, 1)
        lnz2, c = im.internal_energy(beta, h, Dcut, 1, 1, 1)
    finally:
        if rgscheme == "atrg":
            at.set_rsvd_params()
    print(f"\n[{rgscheme}] <s>={m!r} ({m_ref!r}) <s0 s1>={c!r} ({c_ref!r})", flush=True)
    assert abs(lnz - lnz_ref) < 1e-12 and abs(lnz2 - lnz_ref) < 1e-12
    assert abs(m - abs(m_ref)) < 1e-10 * abs(m_ref)
    assert abs(c - c_ref) < 1e-10 * abs(c_ref)
