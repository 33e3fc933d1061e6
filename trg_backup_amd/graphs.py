"""HIP-graph capture of steady-state batched HOTRG steps.

Once the bonds stop growing, every Y (and every X) step of a batch has the
same shapes, so the whole step -- environments, batched top-k eigensolver,
on-device LD/RU choice, merge, max-|T| normalisation and the log-norm write
-- is recorded once into a HIP graph and replayed.  The step contains no host
synchronisation and no memset: flags and counters are cleared by fill
kernels, and the log-norm slot comes from a device-side step counter that the
graph itself advances (trg_normalize with step_ptr).

    step = device int32 counter shared by the Y and X graphs
    gY = GraphedStep2D(batch, chi, "Y", prec, lnf, step)
    gX = GraphedStep2D(batch, chi, "X", prec, lnf, step)
    set_step(step, n); T = gY(T); T = gX(T); ...

GraphedStep3D does the same for steady-state 3D HOTRG steps (trace
normalisation, trg/HOTRG_3d.py:232-238).
"""
import torch

from . import hotrg, ops


def step_counter(device, value=0):
    s = torch.empty(1, dtype=torch.int32, device=device)
    ops.fill_(s, value)
    return s


def set_step(counter, n):
    ops.fill_(counter, int(n))


class GraphedStep2D:
    """`health` (an ops.Health, optional) receives the breakdown flags and bad
    norms of every replay; check() reads it.  Graphs that are replayed one
    after another on one stream may share a memory `pool`
    (torch.cuda.graph_pool_handle()): each step's intermediates are dead once
    it ends, so the steps of a cycle reuse one another's workspace and only
    their outputs stay resident."""

    def __init__(self, batch, chi, direction, prec, lnf, step, warmup=1, health=None, pool=None):
        dev = lnf.device
        self.batch, self.chi, self.direction, self.prec = batch, chi, direction, prec
        self.inp = torch.empty((batch,) + (chi,) * getattr(self, "ndim", 4), dtype=prec.store, device=dev)
        ops.fill_(self.inp, 1.0)
        self.health = health if health is not None else ops.Health(dev)
        # warm-up on a side stream writes into scratch log-norm / counter buffers (its dummy
        # input reports to no health word)
        scratch_lnf = ops.fill_(torch.empty(lnf.shape, dtype=torch.float64, device=dev), 0.0)
        scratch_step = step_counter(dev)
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s), ops.isolated_health(None):
            for _ in range(warmup):
                self._body(scratch_lnf, scratch_step)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, pool=pool), ops.isolated_health(self.health):
            self.out = self._body(lnf, step)
        torch.cuda.synchronize()

    def check(self):
        self.health.check(f"graphed {self.direction} step")

    def _body(self, lnf, step):
        T = hotrg.step2d(self.inp, self.chi, self.direction, self.prec)
        c = ops.maxabs(T, batch=self.batch)
        ops.normalize_(T, c, lnf, 0, batch=self.batch, lnf_stride=lnf.shape[1], step_ptr=step)
        return T

    def __call__(self, T):
        if T.data_ptr() != self.inp.data_ptr():
            ops.copy_into(self.inp, T)
        self.graph.replay()
        return self.out


class GraphedStep3D(GraphedStep2D):
    """A captured steady-state batched 3D HOTRG step (legs chi^6)."""

    def __init__(self, batch, chi, direction, prec, lnf, step, warmup=1, health=None, pool=None):
        self.ndim = 6
        super().__init__(batch, chi, direction, prec, lnf, step, warmup, health, pool)

    def _body(self, lnf, step):
        T = hotrg.step3d(self.inp, self.chi, self.direction, self.prec)
        c = ops.trace_pairs(T, batch=self.batch)
        ops.normalize_(T, c, lnf, 0, batch=self.batch, lnf_stride=lnf.shape[1], step_ptr=step)
        return T
