"""MI355X-native coarse-graining hot path of LUOXIAO92/TRG_backup.

Drop-in modules (same names and call surface as the reference):
  trg_backup_amd.trg.HOTRG_2d, trg_backup_amd.trg.HOTRG_3d,
  trg_backup_amd.trg.ATRG_3d_new, trg_backup_amd.utility.truncated_svd,
  trg_backup_amd.utility.randomized_svd, trg_backup_amd.tensor_class.tensor_class.
Every computation runs on the GPU through libtrg.so (hand-written gfx950
kernels); there is no CPU fallback.
"""
__version__ = "0.1.0"
