"""The benchmark workloads' couplings and precision schedules, shared by
bench.py and the parity tests so that both run exactly the same solves.

* 2D Ising HOTRG, chi = 64, 40 RG steps (Y first, alternating), 32 couplings
  beta_i = 0.40 + 0.08 i / 31 (trg/HOTRG_2d.py:187-216).
* Batched 3D Ising HOTRG, chi = 16, 24 RG steps (Z, Y, X), 128 couplings
  drawn by numpy.random.default_rng(0).uniform(0.20, 0.24, 128)
  (trg/HOTRG_3d.py:247-322).
* 3D q=3 Potts ATRG, beta = 0.55, chi = 64, 30 RG steps (T, X, Y), leg-swap
  rSVD with k = chi, p = chi, 2 power iterations, sketch seed 1
  (trg/ATRG_3d_new.py:974-1039).
"""
import numpy as np

from . import hotrg


def hotrg2d_betas(nb=32):
    return [0.40 + 0.08 * i / (nb - 1) for i in range(nb)] if nb > 1 else [0.44]


def hotrg2d_policy(fp64_steps=6, split6_env_from=6, split6_merge_from=6, split_merge_from=10,
                   bf16_env_from=12, bf16_merge_from=22, eig_q=((12, 8), (22, 4))):
    """float64 through step 6, then float32 storage with: three-piece split-bf16
    (fp32 accuracy) environments for steps 7-12 and merges for 7-10, split-bf16
    merges for 11-22, bf16 merges after 22 and bf16 environments after 12 (a
    value below 0 switches a tier off); the isometry eigensolver runs its 12
    subspace iterations through step 12, 8 through step 22 and 4 after (an
    isometry error enters ln Z/V at second order and with the step's weight
    2^-n; eig_q=() keeps 12 everywhere)."""
    opt = lambda v: None if v is None or v < 0 else v
    return hotrg.Policy(fp64_steps=fp64_steps, low=hotrg.FP32, bf16_env_from=opt(bf16_env_from),
                        split_merge_from=opt(split_merge_from), bf16_merge_from=opt(bf16_merge_from),
                        split6_merge_from=opt(split6_merge_from), split6_env_from=opt(split6_env_from),
                        eig_q=eig_q)


def hotrg3d_betas(nb=128):
    return np.random.default_rng(0).uniform(0.20, 0.24, 128)[:nb].tolist()


def hotrg3d_policy(fp64_steps=6, merge="bf16x6", split_merge_from=16, bf16_merge_from=None,
                   bf16_env_from=None, fp64_merge="f64i8"):
    """float64 through step 6, then float32 storage with three-piece split-bf16
    GEMMs (fp32 accuracy; float32 MFMA with merge="f32"), split-bf16 merges
    after `split_merge_from`, bf16 merges after `bf16_merge_from` and bf16
    environments after `bf16_env_from` (None or a value below 0: never).
    Step 4 is the first full chi^11 step; in the ordered phase of the
    128-coupling draw a float32-accurate step 6 still moves ln Z/V by 5.8e-8
    (beta = 0.2387), so the first six steps stay float64: worst error then
    7.9e-9 over the eight checked couplings, the split-bf16 merges after
    step 16 included.  The merges of those float64 steps run as 6-slice int8
    MFMA products (fp64_merge="f64i8": 42 bits per row in fixed point, so
    entries far below their row's maximum keep fewer; 1.4x faster steps,
    about 1e-9 in ln Z/V, worst validated error 8.3e-9) or on the fp64 MFMA
    (fp64_merge="f64")."""
    opt = lambda v: None if v is None or v < 0 else v
    split6 = fp64_steps if merge == "bf16x6" else None
    return hotrg.Policy(fp64_steps=fp64_steps, low=hotrg.FP32, split6_merge_from=split6, split6_env_from=split6,
                        split_merge_from=opt(split_merge_from), bf16_merge_from=opt(bf16_merge_from),
                        bf16_env_from=opt(bf16_env_from), fp64_merge=fp64_merge)


def tiers(policy, nsteps, first=1):
    """[(first step, Precision)] of each run of equal precision in steps first..nsteps."""
    out = []
    for n in range(first, nsteps + 1):
        p = policy(n)
        if not out or (p.store, p.env, p.merge) != (out[-1][1].store, out[-1][1].env, out[-1][1].merge):
            out.append((n, p))
    return out


def schedule2d(nsteps):
    """(n, direction) as trg/HOTRG_2d.py:195-214 runs them: Y first, then alternating."""
    return [(n, "Y" if n % 2 == 1 else "X") for n in range(1, nsteps + 1)]


def schedule3d(nsteps):
    """(n, direction) in the order of trg/HOTRG_3d.py:259-316: Z, Y, X, ..."""
    return [(n, "ZYX"[(n - 1) % 3]) for n in range(1, nsteps + 1)]
