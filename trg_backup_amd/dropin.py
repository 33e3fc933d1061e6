"""Module aliases that let the reference's own driver scripts run unchanged.

The reference's scripts import the path's modules by their top-level names
(``import trg.HOTRG_2d as hotrg``, ``import trg.ATRG_3d_new as atrg``,
``from utility.randomized_svd import rsvd``, ``from tensor_class.tensor_class
import ATRG_Tensor as Tensor``; e.g. measurement/Ising_model_3d.py).  After

    import trg_backup_amd.dropin
    trg_backup_amd.dropin.install()

those names resolve to the GPU modules of this package.  Modules of the
reference outside the path (``trg.TRG_2d``, ``trg.gauge_fixing``, ...) are
not provided and still fail to import.
"""
import importlib
import sys

ALIASES = {
    "trg": "trg_backup_amd.trg",
    "trg.HOTRG_2d": "trg_backup_amd.trg.HOTRG_2d",
    "trg.HOTRG_3d": "trg_backup_amd.trg.HOTRG_3d",
    "trg.ATRG_3d_new": "trg_backup_amd.trg.ATRG_3d_new",
    "trg.HOTRG_3d_QR": "trg_backup_amd.trg.HOTRG_3d_QR",
    "utility": "trg_backup_amd.utility",
    "utility.truncated_svd": "trg_backup_amd.utility.truncated_svd",
    "utility.randomized_svd": "trg_backup_amd.utility.randomized_svd",
    "tensor_class": "trg_backup_amd.tensor_class",
    "tensor_class.tensor_class": "trg_backup_amd.tensor_class.tensor_class",
    "measurement": "trg_backup_amd.measurement",
    "measurement.Ising_model_3d": "trg_backup_amd.measurement.Ising_model_3d",
    "tensor_init": "trg_backup_amd.tensor_init",
    "tensor_init.Potts_model": "trg_backup_amd.tensor_init.Potts_model",
    "tensor_init.ATRG_init": "trg_backup_amd.tensor_init.ATRG_init",
}


def install(overwrite=False):
    """Register the aliases in sys.modules; returns the names registered.
    Existing entries (e.g. the reference itself already imported) are kept
    unless overwrite=True."""
    done = []
    for alias, target in ALIASES.items():
        if alias in sys.modules and not overwrite:
            continue
        mod = importlib.import_module(target)
        sys.modules[alias] = mod
        parent, _, leaf = alias.rpartition(".")
        if parent and parent in sys.modules:
            setattr(sys.modules[parent], leaf, mod)
        done.append(alias)
    return done


def uninstall():
    for alias, target in ALIASES.items():
        mod = sys.modules.get(alias)
        if mod is not None and getattr(mod, "__name__", "") == target:
            del sys.modules[alias]
