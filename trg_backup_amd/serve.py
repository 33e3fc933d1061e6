"""Continuous-batching server for phase-diagram solves (mixed model, chi, beta).

A request asks for ln Z/V of one model at one coupling, bond dimension and
number of RG steps.  The server

  * admits queued requests in arrival order while the sum of the HBM
    footprints of the admitted flows stays within the budget (requests whose
    footprint alone exceeds the budget are rejected);
  * keeps every admitted flow's tensors resident on the GPU between steps;
  * runs RG steps on a small pool of HIP streams (one host thread each):
    whenever a stream is free it takes the ready flow of the earliest
    request (so requests complete in arrival order rather than all
    advancing together and finishing late) together with
    every other ready flow of the same scheme at the same step shape,
    direction and precision, stacked into one batched launch sequence (an
    ATRG group whose rank decisions part re-runs the step per group of equal
    decisions); a flow is ready again as soon as its step has finished, so
    the work of different requests overlaps on the device and short flows
    are not held back by long ones;
  * admits new arrivals whenever they come, as memory frees up;
  * records the latency arrival -> result of every request.

Every kernel is batch-invariant (an instance's arithmetic does not depend on
which other instances share its launch), so a request's result is the same
alone and in the mix; tests/test_gpu_serve.py checks that.

Model -> scheme: 2D Ising runs HOTRG (trg/HOTRG_2d.py); 3D Ising and 3D q=3
Potts run ATRG (trg/ATRG_3d_new.py, the reference's RGSCHEME=atrg path,
measurement/Ising_model_3d.py:114-122, measurement/Potts3_model_3d.py:32-40),
because 3D HOTRG needs chi^6 storage and 2 chi^11 flop per step (275 GB and
1.5e20 flop at chi = 64).
"""
import math
import threading
import time
from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor, wait
from dataclasses import dataclass, field

import numpy as np
import torch

from . import atrg, hotrg, models, ops


@dataclass
class Request:
    rid: int
    model: str            # "ising2d" | "ising3d" | "potts3d"
    chi: int
    beta: float
    steps: int
    arrival: float = 0.0  # seconds after the start of the stream


@dataclass
class Flow:
    req: Request
    scheme: str
    footprint: int
    step: int = 0
    T: object = None           # 2D HOTRG tensor (1, D, D, D, D)
    lnf: object = None         # (1, steps + 1) float64
    U: object = None
    s: object = None
    VH: object = None
    consts: list = field(default_factory=list)
    admitted_at: float = 0.0
    result: float = None
    finished_at: float = None
    error: str = None          # set when a step's health word reports a breakdown or a bad norm


def footprint(req, policy=None):
    """HBM bytes a flow holds at its peak step (tensor, environments,
    eigensolver and merge / coarse-graining intermediates)."""
    chi = req.chi
    if req.model == "ising2d":
        t = chi ** 4 * 8
        env = 2 * chi ** 4 * 8 * 3
        per_c = 2 * chi ** 4 * chi * 4
        merge = 4 * per_c * max(1, min(chi, hotrg.INSTANCE_BUDGET // max(per_c, 1)))
        return int(3 * t + env + merge)
    # ATRG: U, VH, G, H (chi^4 each), squeezer environments and roots (chi^4 float64 each),
    # range-finder blocks (chi^3 x 2 chi)
    return int(16 * chi ** 4 * 8 + 6 * chi ** 3 * 2 * chi * 8 + chi ** 5 * 8)


class Server:
    def __init__(self, hbm_budget=None, max_batch=32, streams=3, fp64_steps=10, atrg_q=2, atrg_deg_eps=1e-8):
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info()
            hbm_budget = int(0.6 * free)
        self.budget = hbm_budget
        self.max_batch = max_batch
        self.policy = hotrg.Policy(fp64_steps=fp64_steps, low=hotrg.FP32)
        self.atrg_q = atrg_q
        self.atrg_deg_eps = atrg_deg_eps
        self.streams = [torch.cuda.Stream() for _ in range(streams)]   # default priority
        self.setup_stream = torch.cuda.Stream()                       # flow set-up (main thread)
        self.pool = ThreadPoolExecutor(max_workers=streams)
        self.lock = threading.Lock()
        self.rejected = []
        self.max_used = 0

    # ---- flow set-up and stepping ----------------------------------------
    def _start(self, fl):
        with torch.cuda.stream(self.setup_stream):
            self._start_on_stream(fl)

    def _start_on_stream(self, fl):
        r = fl.req
        if fl.scheme == "hotrg2d":
            T = models.initial_tensor(r.model, r.beta)[None].contiguous()
            fl.lnf = ops.fill_(torch.empty((1, r.steps + 1), dtype=torch.float64, device=T.device), 0.0)
            hotrg.normalize_maxabs(T, fl.lnf, 0)
            fl.T = T
        else:
            U, s, VH = models.atrg_initial(r.model, r.beta)
            s, c = atrg.normalize(U, s, VH)
            fl.U, fl.s, fl.VH = U, s, VH
            fl.consts = [c]

    def _hotrg_direction(self, step):
        return "Y" if step % 2 == 1 else "X"       # trg/HOTRG_2d.py:195-214 (Y first)

    def _atrg_direction(self, step):
        return "TXY"[(step - 1) % 3]                # trg/ATRG_3d_new.py:999-1000

    def _step_group(self, flows, stream):
        """One RG step of every flow in `flows` on `stream` (a worker thread).
        The tensors a step replaces stay referenced until the stream has
        drained: the caching allocator may hand their blocks to another
        stream as soon as they are released."""
        keep = []
        with torch.cuda.stream(stream):
            health = ops.Health(stream.device)
            health.__enter__()
        try:
            self._step_group_on(flows, stream, keep)
        finally:
            health.__exit__(None, None, None)
        stream.synchronize()
        bad = health.value()
        if bad:
            msg = "; ".join(m for b, m in ops.Health.BITS.items() if bad & b)
            for f in flows:
                f.error = f"step {f.step} failed on the device: {msg}"
        del keep
        return flows

    def _step_group_on(self, flows, stream, keep):
        with torch.cuda.stream(stream):
            if flows[0].scheme == "hotrg2d":
                n0 = flows[0].step + 1
                prec = self.policy(n0)
                if len(flows) > 1:
                    T = torch.empty((len(flows),) + tuple(flows[0].T.shape[1:]), dtype=flows[0].T.dtype,
                                    device=flows[0].T.device)
                    for i, f in enumerate(flows):
                        ops.copy_into(T[i:i + 1], f.T)
                else:
                    T = flows[0].T
                T = hotrg.step2d(T, flows[0].req.chi, self._hotrg_direction(n0), prec)
                c = ops.maxabs(T, batch=T.shape[0])
                for i, f in enumerate(flows):
                    n = f.step + 1
                    Ti = ops.contiguous(T[i:i + 1])
                    ops.normalize_(Ti, c[i:i + 1], f.lnf, n, batch=1, lnf_stride=f.lnf.shape[1])
                    keep.append(f.T)
                    f.T = Ti
                    f.step = n
                keep += [T, c]
            else:
                n0 = flows[0].step + 1
                d = self._atrg_direction(n0)
                if len(flows) > 1:
                    # equal bond shapes: one batched step (re-run per group where ranks part)
                    res = atrg.atrg_pure_tensor_batch(atrg._stack([f.U for f in flows]),
                                                      atrg._stack([f.s for f in flows]),
                                                      atrg._stack([f.VH for f in flows]), flows[0].req.chi, d, "f64")
                else:
                    res = [atrg.atrg_pure_tensor(flows[0].U, flows[0].s, flows[0].VH, flows[0].req.chi, d, "f64")]
                for f, (U, s, VH) in zip(flows, res):
                    s, c = atrg.normalize(U, s, VH)
                    keep.append((f.U, f.s, f.VH))
                    f.U, f.s, f.VH = U, s, VH
                    f.consts.append(c)
                    f.step = n0

    def _finish(self, fl, now):
        r = fl.req
        if fl.error:
            fl.result = None
            fl.T = fl.lnf = fl.U = fl.s = fl.VH = None
        elif fl.scheme == "hotrg2d":
            fl.result = float(hotrg.lnz2d(fl.T, fl.lnf, r.steps)[0].item())
            fl.T = fl.lnf = None
        else:
            tr = atrg.trace(fl.U, fl.s, fl.VH)
            lg = sum(math.log(abs(float(c[0].item()))) / 2 ** k for k, c in enumerate(fl.consts))
            fl.result = lg + math.log(abs(float(tr[0].item()))) / 2 ** r.steps
            fl.U = fl.s = fl.VH = None
        fl.finished_at = now

    # ---- the loop -----------------------------------------------------------
    def run(self, requests, realtime=True, log=None):
        """Serve `requests` (arrival times in seconds).  Returns the finished flows."""
        atrg_params = dict(atrg.RSVD)
        atrg.set_rsvd_params(k_factor=2, p=0, q=self.atrg_q, seed=1)
        old_deg = atrg.degeneracy_eps
        atrg.degeneracy_eps = self.atrg_deg_eps
        try:
            return self._run(sorted(requests, key=lambda r: (r.arrival, r.rid)), realtime, log)
        finally:
            atrg.RSVD.update(atrg_params)
            atrg.degeneracy_eps = old_deg

    def _group_key(self, f):
        if f.scheme == "hotrg2d":
            n = f.step + 1
            return ("h", f.req.chi, tuple(f.T.shape[1:]), f.T.dtype, self._hotrg_direction(n), self.policy(n))
        n = f.step + 1
        return ("a", f.req.chi, tuple(f.U.shape), tuple(f.s.shape), tuple(f.VH.shape), self._atrg_direction(n))

    def _take_group(self, ready):
        """The oldest ready flow, with every other ready flow that can share
        its launches (same scheme, step shape, direction and precision)."""
        key = self._group_key(ready[0])
        group, rest = [], []
        for f in ready:
            if len(group) < self.max_batch and self._group_key(f) == key:
                group.append(f)
            else:
                rest.append(f)
        ready[:] = rest
        return group

    def _run(self, queue, realtime, log):
        """Event-driven loop: a flow goes back to the ready list as soon as its
        step has finished, and a free stream takes the next ready group at once,
        so fast flows are never held back by slow ones."""
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        ready, running, done, used, qi = [], {}, [], 0, 0
        free = list(range(len(self.streams)))
        last_note = 0.0
        while qi < len(queue) or ready or running:
            admitted = False
            while qi < len(queue) and (not realtime or queue[qi].arrival <= now()):
                r = queue[qi]
                fp = footprint(r)
                if fp > self.budget:
                    self.rejected.append(r)
                    qi += 1
                    continue
                if used + fp > self.budget:
                    break
                fl = Flow(req=r, scheme="hotrg2d" if r.model == "ising2d" else "atrg3d", footprint=fp,
                          admitted_at=now())
                self._start(fl)
                ready.append(fl)
                used += fp
                self.max_used = max(self.max_used, used)
                admitted = True
                qi += 1
            if admitted:
                # set-up kernels ran on the set-up stream; the workers' streams must see them done
                self.setup_stream.synchronize()
            while free and ready:
                group = self._take_group(ready)
                si = free.pop(0)
                running[self.pool.submit(self._step_group, group, self.streams[si])] = (group, si)
            if not running:
                if qi < len(queue) and realtime:
                    time.sleep(max(0.0, min(0.01, queue[qi].arrival - now())))
                continue
            timeout = max(0.0, queue[qi].arrival - now()) if (realtime and qi < len(queue)) else None
            finished, _ = wait(list(running), timeout=timeout, return_when=FIRST_COMPLETED)
            for fu in finished:
                group, si = running.pop(fu)
                fu.result()
                free.append(si)
                for f in group:
                    if f.step >= f.req.steps or f.error:
                        self._finish(f, now())
                        used -= f.footprint
                        done.append(f)
                        if log:
                            res = f"lnZ/V={f.result:.12f}" if f.error is None else f"error: {f.error}"
                            # without real time the arrival offsets are not waited for, so only
                            # the time since the start of the run means anything
                            t = (f"latency={f.finished_at - f.req.arrival:.3f}s" if realtime
                                 else f"done at {f.finished_at:.3f}s")
                            log(f"done rid={f.req.rid} {f.req.model} chi={f.req.chi} steps={f.req.steps} {res} {t}")
                    else:
                        ready.append(f)
            # oldest request first: a stream always takes the earliest admitted ready flow (with
            # every flow that can share its launches), so requests complete in arrival order
            # instead of all advancing together
            ready.sort(key=lambda f: (f.req.arrival, f.req.rid))
            if log and now() - last_note >= 10.0:
                last_note = now()
                log(f"t={last_note:.1f}s: {len(done)} of {len(queue)} done, "
                    f"{sum(len(g) for g, _ in running.values())} flows stepping, {len(ready)} ready, "
                    f"{used / 2 ** 30:.1f} GiB admitted")
        return done


def poisson_stream(n, rate, seed, chis=(16, 24, 32, 48, 64), steps=(20, 40),
                   models_=("ising2d", "ising3d", "potts3d"), window=0.1):
    """The serving workload: Poisson arrivals (exponential gaps, `rate` per
    second), model uniform over `models_`, chi uniform over `chis`, beta
    uniform in [1 - window, 1 + window] x beta_c, steps uniform in [lo, hi]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        m = models_[rng.integers(len(models_))]
        chi = int(chis[rng.integers(len(chis))])
        b = models.BETA_C[m] * rng.uniform(1 - window, 1 + window)
        st = int(rng.integers(steps[0], steps[1] + 1))
        out.append(Request(rid=i, model=m, chi=chi, beta=float(b), steps=st, arrival=float(t)))
    return out


def summarize(done, t_total):
    lat = np.array([f.finished_at - f.req.arrival for f in done])
    return {"requests": len(done), "solves_per_s": len(done) / t_total if t_total > 0 else None,
            "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None,
            "rg_steps": int(sum(f.req.steps for f in done))}
