"""Drop-in for the reference's utility package (truncated and randomized SVD)."""
