"""Write tests/golden/atrg3d_potts.json: float64 oracle ln Z/V of 3D q=3 Potts ATRG
(trg/ATRG_3d_new.py:974-1039) at beta = 0.55, 30 RG steps (T, X, Y), with the
benchmark's leg-swap randomized SVD (k = chi, p = chi, Philox sketch seed 1)
at q = 2 power iterations and with a converged range finder (q = 20).

Both sides of the comparison use the leg swap in the identity basis and
DGENERACY_EPS = 1e-8: with 0 the rank rule may cut inside an exactly
degenerate S3 multiplet, and which member survives is then a free choice of
each SVD implementation.  The initial tensor is the oracle's
models.to_atrg(potts3d_tensor(0.55)) (U, s, VH), which the GPU test feeds to
the device flow unchanged.

Usage:  python tests/golden/make_atrg3d_potts.py CHI [CHI ...]
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import ATRG_3d_new as oat  # noqa: E402
from oracle import models  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "atrg3d_potts.json")
BETA, STEPS = 0.55, {"X": 10, "Y": 10, "T": 10}


def run(chi, q):
    oat.set_leg_swap_basis("identity")
    oat.set_rsvd_params(k_factor=1, p=chi, q=q, seed=1)
    old = oat.degeneracy_eps
    oat.degeneracy_eps = 1e-8
    t0 = time.time()
    try:
        lnz, T = oat.ln_Z_over_V(models.to_atrg(models.potts3d_tensor(BETA), chi), chi, dict(STEPS))
    finally:
        oat.degeneracy_eps = old
        oat.set_leg_swap_basis("svd")
        oat.set_rsvd_params()
    return {"lnZ": lnz, "normal_const": {k: float(v) for k, v in T.normal_const.items()},
            "final_bond": list(T.U.shape), "seconds": round(time.time() - t0, 1)}


def main():
    data = {"model": "3D q=3 Potts, h = 0, ATRG initial tensor of models.to_atrg", "beta": BETA,
            "steps": STEPS, "order": "T, X, Y", "leg_swap_basis": "identity", "DGENERACY_EPS": 1e-8,
            "rsvd": "k_factor 1, p = chi, Philox seed 1", "runs": {}}
    if os.path.exists(OUT):
        with open(OUT) as f:
            data = json.load(f)
    for a in sys.argv[1:]:
        chi = int(a)
        for q in (2, 20):
            res = run(chi, q)
            data["runs"][f"chi{chi}_q{q}"] = res
            print(f"chi={chi} q={q} lnZ/V={res['lnZ']!r} ({res['seconds']} s)", flush=True)
            with open(OUT + ".tmp", "w") as f:
                json.dump(data, f, indent=1, sort_keys=True)
            os.replace(OUT + ".tmp", OUT)


if __name__ == "__main__":
    main()
