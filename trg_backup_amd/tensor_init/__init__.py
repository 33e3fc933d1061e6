"""Drop-ins for the reference's tensor_init modules (the step before the coarse-graining path)."""
