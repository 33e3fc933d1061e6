"""Write tests/golden/hotrg3d_chi16.json: float64 oracle ln Z/V of 3D Ising HOTRG
at chi=16, 24 RG steps (Z, Y, X; trg/HOTRG_3d.py:247-322) for couplings of the
batched benchmark's draw numpy.random.default_rng(0).uniform(0.20, 0.24, 128).

The oracle takes LAPACK eigh of the four 256 x 256 environments
(trg/HOTRG_3d.py:55-58) and evaluates the merge (:107-116) sliced over the
output leg, one slice of 2 chi^10 flop at a time (the unsliced chi^8
intermediate is 34 GB in float64).  A steady step costs about two minutes on
8 host cores.

Usage:  python tests/golden/make_hotrg3d_chi16.py INDEX [INDEX ...]   (indices into the draw)
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import HOTRG_3d as h3  # noqa: E402
from oracle import models  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hotrg3d_chi16.json")
CHI, LOOPS = 16, 8


def run(beta):
    T = models.ising3d_tensor(beta)
    t0 = time.time()
    lnf = np.zeros(3 * LOOPS + 1)
    T, c = h3.__tensor_normalization__(T)
    lnf[0] = np.log(c)
    n = 0
    flags = []
    for _ in range(LOOPS):
        for direction in ("Z", "Y", "X"):
            n += 1
            Tr = h3.__tensor_legs_rearrange__(T, direction)
            U1, U2, PM1, PM2 = h3.__contraction_projector__(Tr, CHI)
            T = h3.__tensor_legs_restore__(h3._merge_pm(Tr, Tr, U1, U2, PM1, PM2), direction)
            T, c = h3.__tensor_normalization__(T)
            lnf[n] = np.log(c) / 2 ** n
            flags.append(PM1[0] + PM2[0])
            print(f"beta={beta} step {n} {direction} {PM1}/{PM2} c={c:.6e} t={time.time() - t0:.0f}s",
                  file=sys.stderr, flush=True)
    tr = float(np.einsum("xxyytt", T))
    lnz = float(np.sum(lnf) + np.log(tr) / 2 ** (3 * LOOPS))
    return {"lnZ": lnz, "lnf": [float(v) for v in lnf], "trace": tr, "PM": flags, "seconds": round(time.time() - t0, 1)}


def main():
    draw = np.random.default_rng(0).uniform(0.20, 0.24, 128)
    for a in sys.argv[1:]:
        i = int(a)
        beta = float(draw[i])
        res = run(beta)
        res["draw_index"] = i
        data = {"model": "3D Ising, h=0, T = sum_s prod_leg W_s,leg, W W^T = [[e^b, e^-b], [e^-b, e^b]]",
                "chi": CHI, "XLOOPS": LOOPS, "YLOOPS": LOOPS, "ZLOOPS": LOOPS, "order": "Z, Y, X",
                "draw": "numpy.random.default_rng(0).uniform(0.20, 0.24, 128)",
                "eigh": "numpy.linalg.eigh (LAPACK) of the four 256 x 256 environments", "betas": {}}
        if os.path.exists(OUT):
            with open(OUT) as f:
                data = json.load(f)
        data["betas"][repr(beta)] = res
        with open(OUT + ".tmp", "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(OUT + ".tmp", OUT)
        print(f"beta={beta} lnZ/V={res['lnZ']!r}", flush=True)


if __name__ == "__main__":
    main()
