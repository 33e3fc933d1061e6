"""Drop-in coarse-graining modules (trg/HOTRG_2d.py, trg/HOTRG_3d.py, trg/HOTRG_3d_QR.py,
trg/ATRG_3d_new.py)."""
