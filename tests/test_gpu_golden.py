"""Parity at the benchmark configurations against golden vectors written by
the float64 oracle (tests/golden/make_*.py): the device float64 path to 1e-12
relative in ln Z/V and the benchmark's mixed-precision schedule to 1e-8
(2D Ising chi = 64, 3D Ising chi = 16, 3D Potts ATRG).

2D Ising HOTRG chi = 64, 40 RG steps (trg/HOTRG_2d.py:187-216): the oracle
uses LAPACK eigh of the full 4096 x 4096 environments (:33-50); the device
uses the top-k subspace eigensolver.
"""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD2D = os.path.join(HERE, "golden", "hotrg2d_chi64.json")


def _load(path):
    if not os.path.exists(path):
        pytest.fail(f"golden file {path} is missing (tests/golden/make_*.py writes it)")
    with open(path) as f:
        return json.load(f)


def _hotrg2d(gpu, policy):
    import torch

    from trg_backup_amd import hotrg, models
    g = _load(GOLD2D)
    betas = sorted(g["betas"], key=float)
    T0 = torch.stack([models.initial_tensor("ising2d", float(b), gpu) for b in betas])
    print(f"hotrg2d chi={g['chi']} {g['XLOOPS'] + g['YLOOPS']} steps, betas {betas}", flush=True)
    T, lnf = hotrg.renorm2d(T0, g["chi"], g["XLOOPS"], g["YLOOPS"], policy)
    z = hotrg.lnz2d(T, lnf, g["XLOOPS"] + g["YLOOPS"]).cpu().numpy()
    ref = np.array([g["betas"][b]["lnZ"] for b in betas])
    lnf = lnf.cpu().numpy()
    lref = np.array([g["betas"][b]["lnf"] for b in betas])
    # per-step normalisation constants ln c_n = lnf[n] 2^n (diagnostic)
    w = 2.0 ** np.arange(lref.shape[1])
    dc = np.abs(lnf * w - lref * w).max(axis=0)
    print("max |d ln c_n| per step:", " ".join(f"{v:.1e}" for v in dc), flush=True)
    for b, a, r in zip(betas, z, ref):
        print(f"beta={b}: device {a!r} oracle {r!r} diff {a - r:.3e} rel {abs(a - r) / abs(r):.3e}", flush=True)
    return z, ref


@pytest.mark.gpu
def test_hotrg2d_chi64_fp64_matches_oracle(gpu):
    from trg_backup_amd import hotrg
    z, ref = _hotrg2d(gpu, hotrg.ALL_FP64)
    assert np.all(np.abs(z - ref) <= 1e-12 * np.abs(ref))


@pytest.mark.gpu
def test_hotrg2d_chi64_bench_schedule_within_1e8_of_oracle(gpu):
    from trg_backup_amd import workloads
    z, ref = _hotrg2d(gpu, workloads.hotrg2d_policy())
    assert np.all(np.abs(z - ref) <= 1e-8)


def test_hotrg2d_chi64_golden_against_onsager():
    """The oracle's chi = 64 values sit within the truncation error of the
    exact free energy: 1e-6 off criticality and 1e-4 at beta_c (where HOTRG
    converges slowest in chi)."""
    from oracle import exact
    g = _load(GOLD2D)
    for b, v in g["betas"].items():
        beta = float(b)
        tol = 1e-4 if abs(beta - 0.4406867935) < 2e-3 else 1e-6
        assert abs(v["lnZ"] - exact.onsager_lnZ(beta)) < tol, (beta, v["lnZ"], exact.onsager_lnZ(beta))
        assert abs(sum(v["lnf"]) + np.log(v["trace"]) / 2 ** 40 - v["lnZ"]) < 1e-15


# ---------------------------------------------------------------------------
# Batched 3D Ising HOTRG chi = 16, 24 steps (trg/HOTRG_3d.py:247-322): the
# oracle uses LAPACK eigh of the four 256 x 256 environments and a merge
# sliced over the output leg; couplings of the benchmark's draw
# ---------------------------------------------------------------------------
GOLD3D = os.path.join(HERE, "golden", "hotrg3d_chi16.json")


def _hotrg3d(gpu, policy):
    import torch

    from trg_backup_amd import hotrg, models
    g = _load(GOLD3D)
    betas = sorted(g["betas"], key=float)
    nsteps = g["XLOOPS"] + g["YLOOPS"] + g["ZLOOPS"]
    T0 = torch.stack([models.initial_tensor("ising3d", float(b), gpu) for b in betas])
    print(f"hotrg3d chi={g['chi']} {nsteps} steps, betas {betas}", flush=True)
    T, lnf = hotrg.renorm3d(T0, g["chi"], g["XLOOPS"], g["YLOOPS"], g["ZLOOPS"], policy)
    z = hotrg.lnz3d(T, lnf, nsteps).cpu().numpy()
    ref = np.array([g["betas"][b]["lnZ"] for b in betas])
    lnf = lnf.cpu().numpy()
    lref = np.array([g["betas"][b]["lnf"] for b in betas])
    w = 2.0 ** np.arange(lref.shape[1])
    print("max |d ln c_n| per step:", " ".join(f"{v:.0e}" for v in np.abs(lnf * w - lref * w).max(axis=0)), flush=True)
    for b, a, r in zip(betas, z, ref):
        print(f"beta={b}: device {a!r} oracle {r!r} diff {a - r:.3e} rel {abs(a - r) / abs(r):.3e}", flush=True)
    return z, ref


@pytest.mark.gpu
def test_hotrg3d_chi16_fp64_matches_oracle(gpu):
    from trg_backup_amd import hotrg
    z, ref = _hotrg3d(gpu, hotrg.ALL_FP64)
    assert np.all(np.abs(z - ref) <= 1e-12 * np.abs(ref))


@pytest.mark.gpu
def test_hotrg3d_chi16_fp64_int8_merges_match_oracle(gpu):
    """Float64 storage throughout, every merge with M, N, K >= 256 (the chi^11
    GEMM) as 6-slice int8 MFMA products, against the oracle's exact-eigh
    float64 solve, to the 1e-8 of the precision tiers.  Measured: 3.2e-11
    and 9.5e-10.  The slices are fixed point per row, so entries far below
    their row's maximum keep few bits: on these tensors the products carry
    about 1e-10 relative error in the step's trace, not the 2^-42 of a
    row's largest entries."""
    from trg_backup_amd import hotrg
    z, ref = _hotrg3d(gpu, hotrg.Policy(fp64_merge="f64i8"))
    assert np.all(np.abs(z - ref) <= 1e-8)


@pytest.mark.gpu
@pytest.mark.parametrize("fp64_merge", ["f64i8", "f64"])
def test_hotrg3d_chi16_bench_schedule_within_1e8_of_oracle(gpu, fp64_merge):
    """The benchmark's schedule (its float64-step merges on the int8 MFMA by
    default, or on the fp64 MFMA) against the oracle's exact-eigh solve."""
    from trg_backup_amd import workloads
    z, ref = _hotrg3d(gpu, workloads.hotrg3d_policy(fp64_merge=fp64_merge))
    assert np.all(np.abs(z - ref) <= 1e-8)


def test_hotrg3d_golden_steps_consistent():
    """The stored per-step constants reassemble the stored ln Z/V, the
    couplings are the benchmark's draw, and the trace is positive."""
    from trg_backup_amd import workloads
    g = _load(GOLD3D)
    draw = workloads.hotrg3d_betas(128)
    for b, v in g["betas"].items():
        assert float(b) == draw[v["draw_index"]]
        assert v["trace"] > 0
        assert abs(sum(v["lnf"]) + np.log(v["trace"]) / 2 ** 24 - v["lnZ"]) < 1e-15
        assert 0.69 < v["lnZ"] < 0.85          # between ln 2 (beta = 0) and the ordered phase


# ---------------------------------------------------------------------------
# 3D q=3 Potts ATRG, beta = 0.55, 30 steps (trg/ATRG_3d_new.py:974-1039), the
# benchmark's leg-swap rSVD (k = chi, p = chi, seed 1) at q = 2 and q = 20
# ---------------------------------------------------------------------------
GOLD_ATRG = os.path.join(HERE, "golden", "atrg3d_potts.json")


def _atrg_runs():
    if not os.path.exists(GOLD_ATRG):
        return []
    with open(GOLD_ATRG) as f:
        return sorted({int(k.split("_")[0][3:]) for k in json.load(f)["runs"]})


@pytest.mark.gpu
@pytest.mark.parametrize("chi", _atrg_runs() or [pytest.param(0, marks=pytest.mark.skip("no ATRG golden file"))])
def test_atrg_potts_matches_golden(gpu, chi):
    """Device float64 ATRG against the oracle with the same sketch: with the
    range finder at q = 20 to 1e-10 relative in ln Z/V (its residual
    dependence on the sketch, which the two sides see in different sign
    gauges of the internal bond, is ~2e-11 at chi = 12, against 2e-6 at
    q = 2); at the benchmark's q = 2 within ten times the oracle's own q = 2
    vs q = 20 difference."""
    from oracle import models
    from trg_backup_amd import atrg
    from trg_backup_amd.tensor_class.tensor_class import ATRG_Tensor
    from trg_backup_amd.trg import ATRG_3d_new as at
    g = _load(GOLD_ATRG)
    ref20, ref2 = g["runs"][f"chi{chi}_q20"], g["runs"][f"chi{chi}_q2"]
    out = {}
    old = atrg.degeneracy_eps
    atrg.degeneracy_eps = g["DGENERACY_EPS"]
    try:
        for q in (20, 2):
            at.set_rsvd_params(k_factor=1, p=chi, q=q, seed=1)
            t = models.to_atrg(models.potts3d_tensor(g["beta"]), chi)
            v, T = at.ln_Z_over_V(ATRG_Tensor(t.U, t.s, t.VH, chi, 3, False, {}), chi, dict(g["steps"]))
            out[q] = (v, T)
    finally:
        atrg.degeneracy_eps = old
        at.set_rsvd_params()
    v20, T20 = out[20]
    v2, _ = out[2]
    print(f"\n[atrg golden chi={chi}] q20 {v20!r} vs {ref20['lnZ']!r} ({(v20 - ref20['lnZ']) / ref20['lnZ']:.2e} rel); "
          f"q2 {v2!r} vs {ref2['lnZ']!r} (oracle q2-q20 {ref2['lnZ'] - ref20['lnZ']:.2e})", flush=True)
    # the kept ranks may differ where the squeezer spectrum's tail sits below float64
    # resolution of the discarded mass (the device cuts there, the oracle's LAPACK values
    # still exceed truncate_eps): those directions weigh ~1e-10 and do not move ln Z/V
    print(f"final U {list(T20.U.shape)} vs oracle {ref20['final_bond']}", flush=True)
    # chi <= 16: the squeezer's SVD is the exact Jacobi one on both sides, 2.3e-11 relative at chi = 12.
    # Above, the device takes the leading triplets by a randomized block iteration: at chi = 24 the
    # two sides agree to 2.6e-9 relative, and neither 60 iterations nor a block of 4 chi moves the
    # device value by more than 2e-9, so the residual is the q = 20 range finder's dependence on the
    # sketch (seen in different sign gauges of the internal bond) at this bond dimension
    tol = 1e-10 if chi <= 16 else 5e-9
    assert abs(v20 - ref20["lnZ"]) < tol * abs(ref20["lnZ"])
    assert abs(v2 - ref2["lnZ"]) <= 10 * abs(ref2["lnZ"] - ref20["lnZ"]) + 1e-12
