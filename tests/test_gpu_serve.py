"""GPU tests of the continuous-batching server: alone == in the mix, admission."""
import numpy as np
import pytest

from oracle import HOTRG_2d as o2
from oracle import models as omodels

pytestmark = pytest.mark.gpu


def _reqs():
    from trg_backup_amd.serve import Request
    return [Request(0, "ising2d", 16, 0.44, 12, 0.0),
            Request(1, "ising2d", 16, 0.43, 12, 0.0),
            Request(2, "ising2d", 16, 0.45, 14, 0.01),
            Request(3, "ising2d", 8, 0.41, 10, 0.02),
            Request(4, "ising3d", 6, 0.22, 6, 0.0),
            Request(5, "potts3d", 6, 0.55, 6, 0.03),
            Request(6, "ising2d", 16, 0.44, 12, 0.05),
            Request(7, "potts3d", 6, 0.56, 6, 0.0),
            Request(8, "potts3d", 6, 0.55, 6, 0.04)]


def test_each_result_is_the_same_alone_and_in_the_mix(gpu):
    from trg_backup_amd.serve import Server
    reqs = _reqs()
    done = Server(streams=2).run(reqs, realtime=False)
    assert all(f.error is None and np.isfinite(f.result) for f in done), [(f.req.rid, f.error) for f in done]
    mixed = {f.req.rid: f.result for f in done}
    alone = {}
    for r in reqs:
        f = Server(streams=1).run([r], realtime=False)[0]
        assert f.error is None, f.error
        alone[r.rid] = f.result
    for rid in alone:
        print(f"\n[serve] rid={rid} alone={alone[rid]!r} mixed={mixed[rid]!r}", flush=True)
        assert mixed[rid] == alone[rid], (rid, mixed[rid], alone[rid])
    assert mixed[0] == mixed[6]      # identical requests, different arrival and batch
    assert mixed[5] == mixed[8]      # the same for ATRG flows stepped in batched groups


def test_served_2d_result_matches_oracle(gpu):
    from trg_backup_amd.serve import Request, Server
    done = Server().run([Request(0, "ising2d", 16, 0.44, 30, 0.0)], realtime=False)
    ref = o2.ln_Z_over_V(omodels.ising2d_tensor(0.44), 16, 15, 15)
    assert abs(done[0].result - ref) < 1e-8


def test_admission_by_footprint(gpu):
    from trg_backup_amd.serve import Request, Server, footprint
    reqs = [Request(i, "ising2d", 16, 0.40 + 0.01 * i, 8, 0.0) for i in range(6)]
    reqs.append(Request(99, "ising2d", 200, 0.44, 8, 0.0))         # far above the budget
    fp = footprint(reqs[0])
    srv = Server(hbm_budget=int(2.5 * fp))
    done = srv.run(reqs, realtime=False)
    assert sorted(f.req.rid for f in done) == list(range(6))
    assert [r.rid for r in srv.rejected] == [99]
    assert srv.max_used <= 2.5 * fp
