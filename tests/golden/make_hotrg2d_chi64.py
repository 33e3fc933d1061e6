"""Write tests/golden/hotrg2d_chi64.json: float64 oracle ln Z/V of 2D Ising HOTRG
at chi=64 (40 RG steps, Y first, alternating; trg/HOTRG_2d.py:187-216).

The oracle follows the reference step for step with LAPACK eigh of the full
4096 x 4096 environments (trg/HOTRG_2d.py:33-50).  One chi=64 step costs about
half a minute on 8 host cores, so a coupling takes about 20 minutes.

Usage:  python tests/golden/make_hotrg2d_chi64.py BETA [BETA ...]
Each coupling is merged into the JSON file as soon as it finishes.
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import HOTRG_2d as h2  # noqa: E402
from oracle import models  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hotrg2d_chi64.json")
CHI, XLOOPS, YLOOPS = 64, 20, 20


def run(beta):
    T = models.ising2d_tensor(beta)
    t0 = time.time()
    nx = ny = n = 0
    lnf = np.zeros(XLOOPS + YLOOPS + 1)
    T, c = h2.__tensor_normalization__(T)
    lnf[0] = np.log(c)
    flags = []
    while nx < XLOOPS or ny < YLOOPS:
        for direction in ("Y", "X"):
            if direction == "Y" and ny < YLOOPS:
                ny += 1
            elif direction == "X" and nx < XLOOPS:
                nx += 1
            else:
                continue
            n += 1
            Tr = h2.__tensor_legs_rearrange__(T, direction)
            U, LDRU = h2.__contraction_projector__(Tr, CHI)
            T = h2.__tensor_legs_restore__(h2._merge_ldru(Tr, Tr, U, LDRU), direction)
            T, c = h2.__tensor_normalization__(T)
            lnf[n] = np.log(c) / 2 ** n
            flags.append(LDRU)
            print(f"beta={beta} step {n} {direction} {LDRU} c={c:.6e} t={time.time() - t0:.0f}s",
                  file=sys.stderr, flush=True)
    tr = float(np.einsum("aabb", T))
    lnz = float(np.sum(lnf) + np.log(tr) / 2 ** (XLOOPS + YLOOPS))
    return {"lnZ": lnz, "lnf": [float(v) for v in lnf], "trace": tr, "LDRU": "".join(f[0] for f in flags),
            "seconds": round(time.time() - t0, 1)}


def main():
    for b in sys.argv[1:]:
        res = run(float(b))
        data = {"model": "2D Ising, h=0, T_ijkl = sum_s W_si W_sj W_sk W_sl, W W^T = [[e^b, e^-b], [e^-b, e^b]]",
                "chi": CHI, "XLOOPS": XLOOPS, "YLOOPS": YLOOPS, "order": "Y first, then alternating",
                "eigh": "numpy.linalg.eigh (LAPACK) of both 4096 x 4096 environments", "betas": {}}
        if os.path.exists(OUT):
            with open(OUT) as f:
                data = json.load(f)
        data["betas"][repr(float(b))] = res
        with open(OUT + ".tmp", "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(OUT + ".tmp", OUT)
        print(f"beta={b} lnZ/V={res['lnZ']!r}", flush=True)


if __name__ == "__main__":
    main()
