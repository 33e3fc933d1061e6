"""CPU checks of the drop-in surface: module aliases and the reference's public names.

The name lists are the public functions of the reference modules on the path
(trg/HOTRG_2d.py:8-300, trg/HOTRG_3d.py:18-398, trg/HOTRG_3d_QR.py:29-1033, trg/ATRG_3d_new.py:18-974,
utility/truncated_svd.py:17-151, utility/randomized_svd.py:5-165,
tensor_class/tensor_class.py:5-41).
"""
import sys

import pytest

SURFACE = {
    "trg.HOTRG_2d": ["__tensor_legs_rearrange__", "__tensor_legs_restore__", "__contraction_projector__",
                     "new_pure_tensor", "new_impuer_tensor_2imp", "new_impuer_tensor_2to1imp",
                     "new_impuer_tensor_1imp", "__tensor_normalization__", "pure_tensor_renorm",
                     "two_point_func_renorm", "ynearest_two_point_func_renorm"],
    "trg.HOTRG_3d": ["__tensor_legs_rearrange__", "__tensor_legs_restore__", "__contraction_projector__",
                     "new_pure_tensor", "new_impuer_tensor_2imp", "new_impuer_tensor_2to1imp",
                     "new_impuer_tensor_1imp", "__tensor_normalization__", "pure_tensor_renorm",
                     "two_point_func_renorm", "ynearest_two_point_func_renorm", "__get_tensor_spectrum"],
    "trg.HOTRG_3d_QR": ["leg_transposition", "squeezer_fix", "squeezer_original", "squeezer_QR", "slice_legs",
                        "coarse_graining", "squeezer", "new_pure_tensor", "new_impuer_tensor_2points",
                        "new_impuer_tensor_absorb", "new_impuer_tensor_1point", "cal_X", "find_1stto3rdmax",
                        "cal_scaling_dimensions", "cal_coorelation", "normalization", "exec_rgstep_pure_tensor",
                        "pure_tensor", "cal_field_expected_value", "exec_rgstep_1point_func", "one_point_function",
                        "exec_rgstep_2point_func", "two_point_function", "__get_tensor_spectrum"],
    "trg.ATRG_3d_new": ["leg_transposition", "rsvd_for_ATRG_leg_swapping", "leg_swapping", "squeezer",
                        "leg_slicing", "coarse_graining", "atrg_pure_tensor", "normalization",
                        "exec_pure_tensor_renorm", "pure_tensor", "new_impuer_tensor_2points",
                        "new_impuer_tensor_absorb", "new_impuer_tensor_1point",
                        "new_impuer_tensor_1point_MultipleImpureTensors", "save_singularvalues", "cal_X",
                        "find_max", "find_1stto3rdmax", "cal_scaling_dimensions", "cal_coorelation",
                        "sqr_distance", "cal_field_expected_value", "exec_rgstep_1point_func",
                        "one_point_function", "cal_field_expected_value1", "exec_rgstep_1point_func1",
                        "one_point_function1", "exec_rgstep_2point_func", "two_point_function"],
    "utility.truncated_svd": ["truncate", "svd", "eigh"],
    "utility.randomized_svd": ["rsvd", "rsvd2", "__power_iteration__", "__QR_iteration__",
                               "rsvd_for_3dATRG_tensor_init", "rsvd_for_3dATRG_tensor_inte"],
    "tensor_class.tensor_class": ["HOTRG_Tensor", "Tensor", "ATRG_Tensor"],
    "measurement.Ising_model_3d": ["init_tensor", "init_tensor_imp", "ln_Z_over_V", "magnetization",
                                   "internal_energy"],
    "tensor_init.Potts_model": ["infinite_density"],
    "tensor_init.ATRG_init": ["initial_tensor_for_ATRG"],
}


@pytest.fixture
def aliases():
    from trg_backup_amd import dropin
    saved = {k: sys.modules.pop(k) for k in list(dropin.ALIASES) if k in sys.modules}
    names = dropin.install()
    yield names
    dropin.uninstall()
    sys.modules.update(saved)


def test_aliases_resolve_to_the_gpu_modules(aliases):
    import importlib
    assert set(aliases) == set(SURFACE) | {"trg", "utility", "tensor_class", "measurement", "tensor_init"}
    for alias in SURFACE:
        mod = importlib.import_module(alias)
        assert mod.__name__ == "trg_backup_amd." + alias


@pytest.mark.parametrize("alias", sorted(SURFACE))
def test_reference_names_exist(aliases, alias):
    import importlib
    mod = importlib.import_module(alias)
    missing = [n for n in SURFACE[alias] if not hasattr(mod, n)]
    assert not missing, missing


def test_numpy_input_without_library_or_gpu_fails_loudly(monkeypatch):
    """No CPU fallback: with the library missing, a call raises TrgError."""
    import numpy as np

    from trg_backup_amd import _lib
    from trg_backup_amd.utility import truncated_svd
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libtrg.so")
    with pytest.raises(Exception) as e:
        truncated_svd.svd(np.eye(4), shape=[[0], [1]], k=2)
    assert isinstance(e.value, (_lib.TrgError, RuntimeError, AssertionError))


@pytest.mark.parametrize("leg,sl", [(7, 3), (64, 5), (3, 8), (10, 1), (16, 16), (100, 7)])
def test_hotrg3d_qr_slice_legs_matches_oracle(leg, sl):
    """The drop-in's host slicing equals the oracle restatement of trg/HOTRG_3d_QR.py:284-316."""
    from oracle import HOTRG_3d_QR as oq
    from trg_backup_amd.trg import HOTRG_3d_QR as dq
    assert dq.slice_legs(leg, sl) == oq.slice_legs(leg, sl)


def test_hotrg3d_qr_schedule_and_keys():
    """"ZXY" cycle with per-direction counts (trg/HOTRG_3d_QR.py:780-786) and the "(X,Y,Z)" factor keys."""
    from trg_backup_amd import hotrg_qr
    assert hotrg_qr.schedule({"X": 2, "Y": 1, "Z": 3}) == ["Z", "X", "Y", "Z", "X", "Z"]
    assert hotrg_qr.factor_keys({"X": 1, "Y": 1, "Z": 1}) == ["(0,0,0)", "(0,0,1)", "(1,0,1)", "(1,1,1)"]
