"""Benchmark: batched mixed-precision HOTRG 2D on one MI355X.

Workload (the JSON line's config): 2D Ising HOTRG, chi = 64, 40 RG steps (Y first, then
alternating; 20 per axis), 32 couplings beta_i = 0.40 + 0.08 i / 31 batched into every launch,
starting from the seeded initial tensor T = sum_s W^4 (D0 = 2).  A "step" of this bench is one
RG step of all 32 instances along one axis.  The timed region is the whole solve: steps 1..K of
the schedule (K = 40 by default: bond growth, float64 steps, float32 steps, bf16-merge steps)
with the inputs resident in HBM, then ln Z/V, closed by a device synchronisation.
value = 32 * K / (timed seconds)  [RG steps/s, whole GPU].

Precision schedule (hotrg.Policy, workloads.hotrg2d_policy): steps 1..6 in float64, then float32
storage with float64 isometry eigensolvers, traces and log-norms; the merge GEMMs run on the
three-piece split-bf16 ("bf16x6", fp32 accuracy on six bf16 MFMAs) up to step 10, on split-bf16
("bf16x3", about 16 significant bits per operand) up to step 22 and on bf16 after that; the
environment GEMMs on bf16x6 up to step 12 and on bf16 after.  An error made at step n enters
ln Z/V with weight 2^-n, which is what lets later steps run at lower precision.

Warm-up (untimed): the same schedule with every precision tier after step 6 cut to its first
W steps, which loads every kernel and lets the GEMM autotuner time each tile configuration on
every problem shape the timed solve meets.

Also on the JSON line:
  * roofline: the kernel with the most time in the timed solve (the chi^7 merge GEMM), its
    launches timed live with HIP events on the launching stream; algorithmic flops 2 M N K per
    launch; the merge GEMMs of the other precisions are listed beside it;
  * lnZ_err_vs_fp64: max |ln Z/V (timed solve) - ln Z/V (float64 solve)| over all 32 couplings;
  * plain_fp64: one instance, float64 throughout, a host sync per step, the same solve (the
    per-instance float64 GPU path); speedup_vs_plain_fp64 = value / its steps/s;
  * serving: a Poisson stream of mixed solves through serve.Server (p50/p99 latency);
  * cpu_baseline: the NumPy float64 oracle timed on the host for one steady-state chi = 64 RG
    step of one instance.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X published dense MFMA peaks by the MFMA type a kernel issues (split-bf16 issues bf16
# MFMAs, three per algorithmic product, so its algorithmic flop/s can reach a third of this)
PEAK_TFLOPS = {"f32": 157.3, "bf16": 2500.0, "bf16x3": 2500.0, "bf16x6": 2500.0, "f64": 78.6}
MFMA_PER_PRODUCT = {"f32": 1, "bf16": 1, "bf16x3": 3, "bf16x6": 6, "f64": 1}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="hotrg2d",
                    choices=["hotrg2d", "hotrg3d", "hotrg3d_qr", "atrg3d", "serve"],
                    help="hotrg2d: the headline line; hotrg3d / hotrg3d_qr / atrg3d: the batched 3D Ising HOTRG "
                         "(trg/HOTRG_3d.py, trg/HOTRG_3d_QR.py) and the 3D q=3 Potts ATRG configurations; serve: "
                         "a Poisson request stream (secondary lines)")
    ap.add_argument("--steps", type=int, default=None,
                    help="hotrg2d: minimum number of timed RG steps, rounded up to whole 40-step solves (the "
                         "solve is never shortened); hotrg3d / hotrg3d_qr / atrg3d: timed RG steps of the schedule "
                         "(24 / 6 / 30)")
    ap.add_argument("--solve-steps", type=int, default=40, help="hotrg2d: RG steps of one solve (20 per axis)")
    ap.add_argument("--warmup", type=int, default=2, help="warm-up steps per precision tier")
    ap.add_argument("--chi", type=int, default=None, help="hotrg2d 64, hotrg3d 16, atrg3d 64")
    ap.add_argument("--nbeta", type=int, default=None,
                    help="hotrg2d 32, hotrg3d 128 (of the 128-value draw), atrg3d 1 (more: a batched sweep around 0.55)")
    ap.add_argument("--fp64-steps", type=int, default=None,
                    help="hotrg2d 6, hotrg3d 6, atrg3d 14, hotrg3d_qr all (its oblique projectors amplify "
                         "float32 rounding: f32 after step 4 gives 5e-8)")
    ap.add_argument("--atrg-q", type=int, default=2)
    ap.add_argument("--squeezer", default="QR", choices=["original", "QR", "fix"], help="hotrg3d_qr squeezer")
    ap.add_argument("--atrg-prec", default="f32", choices=["f32", "f64"])
    ap.add_argument("--atrg-deg-eps", type=float, default=0.0,
                    help="atrg3d: DGENERACY_EPS of the rank rule (trg/ATRG_3d_new.py:15; the reference reads it "
                         "from the environment with no default)")
    ap.add_argument("--merge3d", default="bf16x6", choices=["f32", "bf16x6"],
                    help="hotrg3d: compute type of the steady steps' GEMMs (bf16x6 = three-piece split-bf16, "
                         "fp32 accuracy)")
    ap.add_argument("--split6-env-from", type=int, default=6,
                    help="three-piece split-bf16 environments (fp32 accuracy) after this step (-1: never)")
    ap.add_argument("--split6-merge-from", type=int, default=6,
                    help="bf16x6 merge (three-piece split-bf16, fp32 accuracy) after this step (-1: never)")
    ap.add_argument("--split-merge-from", type=int, default=10, help="bf16x3 merge after this step (-1: never)")
    ap.add_argument("--bf16-env-from", type=int, default=12, help="bf16 environments after this step (-1: never)")
    ap.add_argument("--bf16-merge-from", type=int, default=22, help="bf16 merge after this step (-1: never)")
    ap.add_argument("--eig-q-tiers", default="12:8,22:4",
                    help="hotrg2d: isometry eigensolver iterations after a step, as step:q pairs ('none': 12 throughout)")
    ap.add_argument("--h3-split-from", type=int, default=16, help="hotrg3d: bf16x3 merges after this step (-1: never)")
    ap.add_argument("--h3-fp64-merge", choices=("f64", "f64i8"), default="f64i8",
                    help="hotrg3d: merges of the float64 steps on the fp64 MFMA or as int8-MFMA slice products")
    ap.add_argument("--h3-validate", type=int, default=8,
                    help="hotrg3d: couplings re-solved in float64 for the accuracy check (spread over the batch)")
    ap.add_argument("--h3-bf16-from", type=int, default=-1, help="hotrg3d: bf16 merges after this step (-1: never)")
    ap.add_argument("--h3-bf16-env-from", type=int, default=-1,
                    help="hotrg3d: bf16 environments after this step (-1: never)")
    ap.add_argument("--validate-betas", type=int, default=0,
                    help="hotrg2d: validate this many couplings against float64 (0 = all of them)")
    ap.add_argument("--validate-all", action="store_true", help="validate every coupling (the hotrg2d default)")
    ap.add_argument("--no-validate", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-n", type=int, default=24)
    ap.add_argument("--serve-chis", default="16,24,32", help="bond dimensions of the request stream")
    ap.add_argument("--serve-steps", default="20,30", help="RG-step range of the request stream")
    ap.add_argument("--serve-check", type=int, default=3,
                    help="ATRG requests of the stream re-solved alone for the accuracy check (smallest chi first)")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--no-graphs", action="store_true", help="hotrg3d: eager steady steps instead of HIP graphs")
    ap.add_argument("--progress", action="store_true",
                    help="hotrg3d: synchronise and log after every timed step (long solves)")
    ap.add_argument("--load-tune", default=None, help="GEMM tile table (JSON) to start from")
    ap.add_argument("--save-tune", default=None, help="write the GEMM tile table after the warm-up")
    ap.add_argument("--gemm-sweep", action="store_true",
                    help="time every tile configuration of the 2D chi=64 merge GEMM (bf16, split-bf16, three-piece) "
                         "and check each against the first, bit for bit")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="run one steady step with each merge precision plus a 1 GiB copy of known size "
                         "(for hardware-counter runs), then exit")
    return ap.parse_args()


def _line(metric_cfg, value, K, warmup, dt, dtype, roof, extra):
    out = {"metric": "RG steps/sec and p50/p99 solve latency at |dlnZ/V|<=1e-8 vs FP64 plain path",
           "value": value, "unit": "RG steps/s", "n_gpus": 1, "steps": K, "warmup": warmup,
           "ms_per_step": dt / K * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": dtype, "data": "synthetic", "config": metric_cfg}
    if roof is not None:
        out["roofline"] = roof
    out.update(extra)
    return out


def _roof(rec, names, peak):
    import torch  # noqa: F401  (events need the runtime)
    evs = [(e0.elapsed_time(e1), fl) for e0, e1, fl in rec.events]
    if not evs:
        return None
    ms = sum(e[0] for e in evs)
    fl = sum(e[1] for e in evs)
    ach = fl / (ms * 1e-3) / 1e12
    return {"bound": "mfma", "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
            "traffic": None, "kernel": names, "launches": len(evs), "avg_ms": ms / len(evs),
            "flops_per_launch": fl / len(evs)}


def bench_hotrg3d(args):
    """Batched 3D Ising HOTRG: beta drawn from numpy.random.default_rng(0).uniform(0.20, 0.24, 128)
    (the first --nbeta of them), chi = 16, steps Z, Y, X (trg/HOTRG_3d.py:259-316) from the initial
    tensor; float64 for the first --fp64-steps steps, then three-piece split-bf16 GEMMs (fp32 accuracy;
    --merge3d f32 for the float32 MFMA), split-bf16 merges after --h3-split-from, with float64
    eigensolvers; --h3-validate couplings are re-solved in float64 for the accuracy check."""
    import numpy as np
    import torch

    from trg_backup_amd import _lib, autotune, hotrg, models, ops, workloads
    _lib.load()
    dev = torch.device("cuda:0")
    chi, nb, K = args.chi, args.nbeta, args.steps
    betas = workloads.hotrg3d_betas(nb)
    policy = workloads.hotrg3d_policy(args.fp64_steps, args.merge3d, args.h3_split_from, args.h3_bf16_from,
                                      args.h3_bf16_env_from, fp64_merge=args.h3_fp64_merge)
    log = lambda *a: print("[bench]", *a, file=sys.stderr, flush=True)
    T0 = torch.stack([models.initial_tensor("ising3d", b, dev) for b in betas])
    sched = workloads.schedule3d(K)
    steady = workloads.tiers(policy, K, args.fp64_steps + 1)

    def run(steps, pol, T0_, what=None):
        lnf = ops.fill_(torch.empty((T0_.shape[0], max(K, steps[-1][0]) + 1), dtype=torch.float64, device=dev),
                        0.0)
        T = ops.contiguous(T0_, pol(0).store)
        hotrg.normalize_trace(T, lnf, 0)
        for n, d in steps:
            T = hotrg.step3d(T, chi, d, pol(n))
            hotrg.normalize_trace(T, lnf, n)
            if what:
                torch.cuda.current_stream().synchronize()
                log(f"{what}: step {n} {d} done at {time.perf_counter() - t0:.1f} s")
        return T, lnf

    autotune.enable(not args.no_tune)
    t0 = time.perf_counter()
    # the merge GEMMs run in sub-batches of at most 16 instances (hotrg.MERGE_BUDGET), so a
    # 16-instance warm-up meets (and tunes) the same chi^11 problems as the timed solve: the
    # float64 steps, then three steps (Z, Y, X) in each steady precision tier
    nfirst = max(args.fp64_steps, 3)
    rep = list(range(nfirst + 1))
    for n_t, _ in steady:
        rep += [max(n_t, nfirst + 1)] * min(3, args.warmup + 1)
    wsched = workloads.schedule3d(len(rep) - 1)
    run(wsched, lambda n: policy(rep[n]), T0[:min(nb, 16)], "warm-up")
    torch.cuda.synchronize()
    autotune.enable(False)
    log(f"warm-up in {time.perf_counter() - t0:.1f} s")
    torch.cuda.empty_cache()
    graphed = {}
    health = ops.Health(dev)
    if not args.no_graphs and K > args.fp64_steps and len(steady) == 1:
        # one captured batched launch sequence per steady direction (float32 steps; a ladder
        # of several steady tiers runs eagerly: a batched chi = 16 step is seconds of GPU work)
        from trg_backup_amd import graphs
        counter = graphs.step_counter(dev)
        lnf_g = ops.fill_(torch.empty((nb, K + 1), dtype=torch.float64, device=dev), 0.0)
        # the Z, Y, X graphs replay in turn on one stream and share one workspace pool
        pool = torch.cuda.graph_pool_handle()
        for d in sorted({dd for n, dd in sched if n > args.fp64_steps}):
            graphed[d] = graphs.GraphedStep3D(nb, chi, d, policy(args.fp64_steps + 1), lnf_g, counter,
                                              health=health, pool=pool)
            torch.cuda.empty_cache()
        log(f"captured {len(graphed)} steady-step graphs; "
            f"{torch.cuda.memory_reserved(dev) / 2**30:.1f} GiB reserved")

    def run_timed():
        if not graphed:
            # a batched chi = 16 step is seconds of GPU work: the per-step sync of --progress costs
            # nothing measurable
            return run(sched, policy, T0, "timed" if args.progress else None)
        ops.fill_(lnf_g, 0.0)
        T = ops.contiguous(T0, policy(0).store)
        hotrg.normalize_trace(T, lnf_g, 0)
        for n, d in sched:
            if n <= args.fp64_steps:
                T = hotrg.step3d(T, chi, d, policy(n))
                hotrg.normalize_trace(T, lnf_g, n)
            else:
                if n == args.fp64_steps + 1:
                    T = ops.cast(T, policy(n).store)
                    graphs.set_step(counter, n)
                T = graphed[d](T)
            if args.progress:
                # a batched chi = 16 step is seconds of GPU work: the sync costs nothing measurable
                torch.cuda.current_stream().synchronize()
                log(f"step {n} {d} done at {time.perf_counter() - t0:.1f} s")
        return T, lnf_g

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with health:
        T, lnf = run_timed()
        z = hotrg.lnz3d(T, lnf, K)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    health.check("timed solve")
    zt = z.cpu().numpy()
    log(f"timed {K} steps x {nb} instances in {dt:.2f} s ({'graph replays' if graphed else 'eager'})")
    # the chi^11 GEMM's launch time, from one more eager steady step (graph replays carry no
    # per-launch events)
    rec = ops.GemmRecorder("merge3d_x3")
    with rec:
        n_last, d_last = sched[-1]
        hotrg.step3d(ops.cast(T, policy(n_last).store), chi, d_last, policy(n_last))
        torch.cuda.synchronize()
    del T
    torch.cuda.empty_cache()
    extra = {"lnZ_per_site": {"betas": betas[:3], "values": zt[:3].tolist()}}
    if not args.no_validate:
        nv = max(1, min(nb, args.h3_validate))
        sel = sorted({round(i * (nb - 1) / max(1, nv - 1)) for i in range(nv)} | {0})
        Td, ld = run(sched, hotrg.ALL_FP64, ops.contiguous(T0[sel]), "float64 validation")
        zd = hotrg.lnz3d(Td, ld, K).cpu().numpy()
        errs = [abs(float(zt[i]) - float(zd[j])) for j, i in enumerate(sel)]
        extra["lnZ_err_vs_fp64"] = max(errs)
        extra["lnZ_check"] = {"betas": [betas[i] for i in sel], "abs_err": errs}
        gold = os.path.join(ROOT, "tests", "golden", "hotrg3d_chi16.json")
        if os.path.exists(gold) and chi == 16 and K == 24:
            with open(gold) as f:
                g = json.load(f)["betas"]
            oerr = {}
            for j, i in enumerate(sel):
                key = repr(float(betas[i]))
                if key in g:
                    oerr[key] = {"fp64_rel_err": abs(float(zd[j]) - g[key]["lnZ"]) / abs(g[key]["lnZ"]),
                                 "mixed_abs_err": abs(float(zt[i]) - g[key]["lnZ"])}
            extra["vs_oracle_golden"] = oerr
        log(f"validated {sel}: {errs}")
        del Td
        torch.cuda.empty_cache()
    if not args.no_plain:
        # plain per-instance float64 path: one instance, float64 throughout, host sync per step
        T1 = ops.contiguous(T0[:1])
        l1 = ops.fill_(torch.empty((1, K + 1), dtype=torch.float64, device=dev), 0.0)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        hotrg.normalize_trace(T1, l1, 0)
        for n, d in sched:
            T1 = hotrg.step3d(T1, chi, d, hotrg.FP64)
            hotrg.normalize_trace(T1, l1, n)
            torch.cuda.synchronize()
        hotrg.lnz3d(T1, l1, K).cpu()
        tp1 = time.perf_counter() - t1
        extra["plain_fp64"] = {"steps_per_s": K / tp1, "solve_s": tp1, "beta": betas[0]}
        extra["speedup_vs_plain_fp64"] = (nb * K / dt) / (K / tp1)
        log(f"plain float64 solve {tp1:.1f} s")
        del T1
        torch.cuda.empty_cache()
    cfg = {"workload": f"3D Ising HOTRG chi={chi}, {K} RG steps (Z,Y,X) from the initial tensor, {nb} beta "
                       "from default_rng(0).uniform(0.20,0.24) batched",
           "chi": chi, "nbeta": nb, "rg_steps": K, "policy": policy.describe(), "parallelism": "single GPU, batched"}
    last = policy(K).merge
    if last in ("bf16", "bf16x3", "bf16x6"):
        roof = _roof(rec, f"merge chi^11 GEMM (T.U1.U2 x T, {last}: {MFMA_PER_PRODUCT[last]} bf16 MFMAs per product, "
                          f"last tier of the solve)", PEAK_TFLOPS[last])
        if roof:
            roof["mfma_issue_frac"] = roof["frac"] * MFMA_PER_PRODUCT[last]
    elif last == "f64i8":
        # 21 int8 products per float64 product: the int8 peak over 21 is this GEMM's ceiling
        roof = _roof(rec, "merge chi^11 GEMM (T.U1.U2 x T, float64 as 6-slice int8 MFMA products)", 5000.0 / 21)
        if roof:
            roof["mfma_issue_frac"] = roof["frac"]
    elif last == "f64":
        roof = _roof(rec, "merge chi^11 GEMM (T.U1.U2 x T, f64 MFMA)", 78.6)
    else:
        roof = _roof(rec, "merge chi^11 GEMM (T.U1.U2 x T, f32 MFMA)", PEAK_TFLOPS["f32"])
    tp = os.path.join(ROOT, "profiles", "gemm_merge3d_x3_traffic.json")
    if roof and os.path.exists(tp):
        # HBM bytes per launch (16 couplings) from the rocprof FETCH_SIZE / WRITE_SIZE runs of the probe
        with open(tp) as f:
            roof["traffic"] = json.load(f).get("bytes_per_launch", {}).get(last)
    print(json.dumps(_line(cfg, nb * K / dt, K, args.warmup, dt, dtype_label(policy, K), roof, extra)), flush=True)
    return 0


def bench_hotrg3d_qr(args):
    """Batched 3D Ising HOTRG with squeezers (trg/HOTRG_3d_QR.py, the 3D HOTRG the reference's 3D
    drivers run): beta from numpy.random.default_rng(0).uniform(0.20, 0.24, 128) (the first --nbeta),
    chi = 16, steps in the "ZXY" cycle from the initial tensor, squeezer --squeezer; float64 for the
    first --fp64-steps steps (default: every step), then float32 MFMA merges with float64
    environments and projectors."""
    import numpy as np
    import torch

    from trg_backup_amd import _lib, autotune, hotrg, hotrg_qr, models, ops
    _lib.load()
    dev = torch.device("cuda:0")
    chi, nb, K = args.chi, args.nbeta, args.steps
    betas = np.random.default_rng(0).uniform(0.20, 0.24, 128)[:nb].tolist()
    # float32 MFMA merges after the float64 steps: the split-bf16 tiers of the hotrg3d workload are
    # not used here (the oblique projectors amplify rounding, and no split-compute parity test
    # covers this flow)
    policy = hotrg.Policy(fp64_steps=args.fp64_steps, low=hotrg.FP32)
    log = lambda *a: print("[bench]", *a, file=sys.stderr, flush=True)
    T0 = torch.stack([models.initial_tensor("ising3d", b, dev) for b in betas])
    order = hotrg_qr.schedule({"X": K, "Y": K, "Z": K})[:K]
    tot = {d: order.count(d) for d in "XYZ"}

    autotune.enable(not args.no_tune)
    t0 = time.perf_counter()
    warm = hotrg_qr.schedule({"X": K, "Y": K, "Z": K})[:min(K, max(args.fp64_steps, 3) + 1)]
    hotrg_qr.renorm(T0, chi, {d: warm.count(d) for d in "XYZ"}, args.squeezer, 0.0, policy)
    torch.cuda.synchronize()
    autotune.enable(False)
    log(f"warm-up in {time.perf_counter() - t0:.1f} s")
    torch.cuda.empty_cache()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    groups = hotrg_qr.renorm(T0, chi, tot, args.squeezer, 0.0, policy)
    z = torch.empty(nb, dtype=torch.float64, device=dev)
    for idx, Tg, lg in groups:
        zg = hotrg.lnz3d(Tg, lg, K)
        for j, b in enumerate(idx):
            ops.copy_into(z[b:b + 1], zg[j:j + 1])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    zt = z.cpu().numpy()
    log(f"timed {K} steps x {nb} instances in {dt:.2f} s ({len(groups)} rank group(s))")
    # the chi^11 GEMM's launch time at the last step's precision, from one more steady step
    rec = ops.GemmRecorder("merge3d_x3")
    with rec:
        idx, Tg, _ = groups[0]
        hotrg_qr.step(Tg, chi, order[-1], args.squeezer, 0.0, policy(K))
        torch.cuda.synchronize()
    del groups, Tg
    extra = {"lnZ_per_site": {"betas": betas[:3], "values": zt[:3].tolist()}, "squeezer": args.squeezer}
    if not args.no_validate:
        sel = [0, nb - 1] if nb > 1 else [0]
        zd = hotrg_qr.ln_z_over_v(ops.contiguous(T0[sel]), chi, tot, args.squeezer, 0.0).cpu().numpy()
        extra["lnZ_err_vs_fp64"] = float(max(abs(zt[i] - zd[j]) for j, i in enumerate(sel)))
    cfg = {"workload": f"3D Ising HOTRG_3d_QR ({args.squeezer} squeezer) chi={chi}, {K} RG steps (ZXY cycle) "
                       f"from the initial tensor, {nb} beta from default_rng(0).uniform(0.20,0.24) batched",
           "chi": chi, "nbeta": nb, "rg_steps": K, "policy": policy.describe(), "parallelism": "single GPU, batched"}
    last = "f64" if policy(K).store == torch.float64 else policy(K).merge
    roof = _roof(rec, f"merge chi^11 GEMM (T.Px.Py x T, {last} MFMA)", PEAK_TFLOPS[last])
    print(json.dumps(_line(cfg, nb * K / dt, K, args.warmup, dt, last, roof, extra)), flush=True)
    return 0


def bench_atrg3d(args):
    """3D q=3 Potts ATRG, beta = 0.55, chi = 64, steps T, X, Y (trg/ATRG_3d_new.py:974-1039) from the
    eigen-factorised initial tensor; leg-swap rSVD with k = chi, p = chi, --atrg-q power iterations,
    Philox sketch from seed 1; contractions in --atrg-prec after the first --fp64-steps steps,
    decompositions float64 and batched where independent.  The plain per-instance float64 path
    runs the same solve with float64 contractions and the decompositions one at a time, in the
    reference's order; its ln Z/V is the accuracy reference (the decompositions give the same
    bits batched or not, so the difference measures the contraction precision alone)."""
    import math

    import torch

    from trg_backup_amd import _lib, atrg, autotune, models, ops
    _lib.load()
    chi, K = args.chi, args.steps
    log = lambda *a: print("[bench]", *a, file=sys.stderr, flush=True)
    saved = (dict(atrg.RSVD), atrg.BATCH_DECOMPOSITIONS, atrg.degeneracy_eps)
    atrg.degeneracy_eps = args.atrg_deg_eps

    nb = args.nbeta
    # one coupling (the configuration's 0.55), or nb couplings around it stepped as one batch
    betas = [0.55] if nb == 1 else [round(0.55 + 0.005 * (i - (nb - 1) // 2), 6) for i in range(nb)]

    def solve(nsteps, prec, q, fp64_steps, what, betas=(0.55,)):
        """ln Z/V of each coupling; several couplings step together through
        atrg_pure_tensor_batch, grouped by bond shape."""
        atrg.set_rsvd_params(k_factor=1, p=chi, q=q, seed=1)
        st = []
        for b in betas:
            U, s, VH = models.atrg_initial("potts3d", b)
            s, c = atrg.normalize(U, s, VH)
            st.append([U, s, VH, [c]])
        t0 = time.perf_counter()
        for n in range(1, nsteps + 1):
            d, pr = "TXY"[(n - 1) % 3], "f64" if n <= fp64_steps else prec
            if len(st) == 1:
                res = [atrg.atrg_pure_tensor(st[0][0], st[0][1], st[0][2], chi, d, pr)]
            else:
                groups, res = {}, [None] * len(st)
                for i, x in enumerate(st):
                    groups.setdefault((tuple(x[0].shape), tuple(x[1].shape), tuple(x[2].shape)), []).append(i)
                for idx in groups.values():
                    out = atrg.atrg_pure_tensor_batch(atrg._stack([st[i][0] for i in idx]),
                                                      atrg._stack([st[i][1] for i in idx]),
                                                      atrg._stack([st[i][2] for i in idx]), chi, d, pr)
                    for j, i in enumerate(idx):
                        res[i] = out[j]
            for x, (U, s, VH) in zip(st, res):
                s, c = atrg.normalize(U, s, VH)
                x[0], x[1], x[2] = U, s, VH
                x[3].append(c)
            if n % 5 == 0 or n == nsteps:
                log(f"{what}: step {n} at {time.perf_counter() - t0:.1f} s")
        zs = []
        for U, s, VH, consts in st:
            tr = atrg.trace(U, s, VH)
            lg = sum(math.log(abs(float(c[0].item()))) / 2 ** k for k, c in enumerate(consts))
            zs.append(lg + math.log(abs(float(tr[0].item()))) / 2 ** nsteps)
        return zs

    try:
        autotune.enable(not args.no_tune)
        t0 = time.perf_counter()
        # the bonds reach chi within the first few steps: the warm-up runs far enough into the
        # steady shapes that the tuner meets every GEMM problem of both precision tiers
        nw = min(K, max(8, 3 + args.warmup))
        solve(nw, args.atrg_prec, args.atrg_q, nw, "warm-up", betas)
        solve(nw, args.atrg_prec, args.atrg_q, 0, "warm-up", betas)
        if not args.no_plain and nb > 1:
            solve(nw, args.atrg_prec, args.atrg_q, nw, "warm-up (one coupling)")
        torch.cuda.synchronize()
        autotune.enable(False)
        log(f"warm-up in {time.perf_counter() - t0:.1f} s")
        torch.cuda.synchronize()
        rec = ops.GemmRecorder("atrg_cg")
        t0 = time.perf_counter()
        with rec:
            zs = solve(K, args.atrg_prec, args.atrg_q, args.fp64_steps, "timed", betas)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        z = zs[betas.index(0.55)]
        log(f"timed {K} steps x {nb} couplings in {dt:.2f} s, ln Z/V(0.55) = {z!r}")
        extra = {"lnZ_per_site": {"beta": 0.55, "value": z}}
        if nb > 1:
            extra["lnZ_per_site"] = {"betas": betas, "values": zs}
            extra["atrg_rank_splits"] = atrg.SPLITS
        if not args.no_plain:
            atrg.BATCH_DECOMPOSITIONS = False
            t1 = time.perf_counter()
            zd = solve(K, "f64", args.atrg_q, K, "plain float64")[0]
            tp = time.perf_counter() - t1
            atrg.BATCH_DECOMPOSITIONS = True
            extra["plain_fp64"] = {"steps_per_s": K / tp, "solve_s": tp, "value": zd, "beta": 0.55,
                                   "path": "one coupling, float64 contractions, decompositions one at a time, "
                                           "host syncs per step"}
            extra["speedup_vs_plain_fp64"] = (nb * K / dt) / (K / tp)
            extra["lnZ_err_vs_fp64"] = abs(z - zd)
    finally:
        atrg.RSVD.update(saved[0])
        atrg.BATCH_DECOMPOSITIONS = saved[1]
        atrg.degeneracy_eps = saved[2]
    # coarse-graining GEMMs (G and H, 4 chi^7 flop per step) of the steps after the float64 ones
    names = {_lib.F32: "f32", _lib.F64: "f64", _lib.BF16: "bf16", _lib.BF16X3: "bf16x3", _lib.BF16X6: "bf16x6"}
    evs = [(e0.elapsed_time(e1), fl, cfg) for (e0, e1, fl), cfg in zip(rec.events, rec.event_configs)]
    groups = {}
    for ms, fl, cfg in evs:
        groups.setdefault(names[cfg[2]], []).append((ms, fl))
    byp = {}
    for nm, v in groups.items():
        ms, fl = sum(x[0] for x in v), sum(x[1] for x in v)
        byp[nm] = {"launches": len(v), "total_ms": ms, "achieved_tflops": fl / (ms * 1e-3) / 1e12,
                   "peak_tflops": PEAK_TFLOPS[nm], "frac": fl / (ms * 1e-3) / 1e12 / PEAK_TFLOPS[nm]}
    roof = None
    if byp:
        dom = max(byp, key=lambda k: byp[k]["total_ms"])
        d = byp[dom]
        roof = {"bound": "mfma", "achieved": d["achieved_tflops"], "peak": d["peak_tflops"], "unit": "TFLOP/s",
                "frac": d["frac"], "traffic": None,
                "kernel": f"ATRG coarse-graining GEMMs G and H ({dom} MFMA), live HIP events",
                "cg_gemm_share_of_solve": sum(v["total_ms"] for v in byp.values()) / (dt * 1e3),
                "by_precision": byp,
                "note": "the rest of the step is float64 decompositions (CholeskyQR range finders, Jacobi "
                        "SVDs, PSD Cholesky roots), latency-bound; see the rocprof summary in profiles/"}
    sweep = "beta=0.55" if nb == 1 else f"{nb} couplings beta in [{betas[0]}, {betas[-1]}] batched"
    cfg = {"workload": f"3D q=3 Potts ATRG chi={chi}, {sweep}, {K} RG steps (T,X,Y), rSVD k=chi p=chi "
                       f"q={args.atrg_q} seed 1", "chi": chi, "rg_steps": K, "nbeta": nb,
           "DGENERACY_EPS": args.atrg_deg_eps,
           "precision": f"f64 for steps <= {args.fp64_steps}, then {args.atrg_prec} contractions "
                        "(decompositions and squeezer environments f64)",
           "parallelism": ("single GPU, one instance, independent decompositions batched" if nb == 1 else
                           "single GPU, couplings of equal bond shapes stepped as one batch")}
    dtype = f"mixed (steps 1-{args.fp64_steps}: f64; after: {args.atrg_prec} contractions; f64 decompositions)"
    print(json.dumps(_line(cfg, nb * K / dt, K, args.warmup, dt, dtype, roof, extra)), flush=True)
    return 0


def serve_leg(args):
    """The request stream through serve.Server: Poisson arrivals (rate 50/s, seed 2), model uniform
    over {2D Ising HOTRG, 3D Ising ATRG, 3D q=3 Potts ATRG}, chi uniform over --serve-chis, beta
    uniform within 10% of the model's critical coupling, steps uniform in --serve-steps.  2D flows
    run float64 for 10 steps then float32; ATRG flows run float64 (rSVD k = 2 chi, q = 2), flows of equal
    bond shapes stepped in batched groups."""
    import torch

    from trg_backup_amd import atrg, serve
    chis = tuple(int(c) for c in args.serve_chis.split(","))
    lo, hi = (int(c) for c in args.serve_steps.split(","))
    reqs = serve.poisson_stream(args.serve_n, 50.0, 2, chis=chis, steps=(lo, hi))
    srv = serve.Server()
    t1 = time.perf_counter()
    done = srv.run(reqs, realtime=True, log=lambda m: print("[serve]", m, file=sys.stderr, flush=True))
    ts = time.perf_counter() - t1
    summ = serve.summarize(done, ts)
    # accuracy of every model: the 2D flows (float32 after step 10) against float64 solves of the
    # same requests; the ATRG flows (float64, stepped in batched groups) against a sample of the same
    # requests solved alone, one flow at a time (the per-instance float64 path)
    twod = [f.req for f in done if f.req.model == "ising2d"]
    slog = lambda m: print("[serve]", m, file=sys.stderr, flush=True)
    slog(f"stream served in {ts:.1f} s; float64 re-solves of the {len(twod)} 2D requests")
    ref = ({f.req.rid: f.result for f in serve.Server(fp64_steps=10 ** 9).run(twod, realtime=False, log=slog)}
           if twod else {})
    sample = [f.req for f in done if f.req.model != "ising2d" and f.error is None]
    sample = sorted(sample, key=lambda r: (r.chi, r.rid))[:args.serve_check]
    for r in sample:
        ref[r.rid] = serve.Server(streams=1).run([r], realtime=False)[0].result
        slog(f"rid={r.rid} {r.model} chi={r.chi} re-solved alone")
    errs = {}
    for f in done:
        if f.req.rid in ref and f.error is None:
            errs[f.req.model] = max(errs.get(f.req.model, 0.0), abs(f.result - ref[f.req.rid]))
    checked = {m: sum(1 for f in done if f.req.rid in ref and f.req.model == m) for m in errs}
    by = {}
    for f in done:
        by.setdefault(f.req.model, []).append(f.finished_at - f.req.arrival)
    torch.cuda.empty_cache()
    return dict(summ, worst_lnZ_err_vs_fp64=max(errs.values() or [0.0]), lnZ_err_by_model=errs,
                lnZ_checked_by_model=checked, errors=sum(1 for f in done if f.error),
                atrg_rank_splits=atrg.SPLITS, rejected=len(srv.rejected), wall_s=ts,
                max_hbm_admitted_gb=srv.max_used / 2 ** 30,
                latency_by_model={m: {"n": len(v), "p50_s": float(sorted(v)[len(v) // 2]), "max_s": max(v)}
                                  for m, v in by.items()},
                stream=f"{args.serve_n} Poisson requests (rate 50/s, seed 2), models {{2D Ising HOTRG, 3D Ising "
                       f"ATRG, 3D Potts ATRG}}, chi in {set(chis)}, {lo}-{hi} RG steps, beta within 10% of beta_c")


def bench_serve(args):
    from trg_backup_amd import _lib
    _lib.load()
    res = serve_leg(args)
    cfg = {"workload": res["stream"], "parallelism": "single GPU, continuous batching"}
    extra = {"serving": res, "p50_latency_s": res["p50_latency_s"], "p99_latency_s": res["p99_latency_s"]}
    out = _line(cfg, res["rg_steps"] / res["wall_s"], res["rg_steps"], 0, res["wall_s"], "f32/f64", None, extra)
    print(json.dumps(out), flush=True)
    return 0


def gemm_sweep(args):
    """Every tile configuration of the chi=64 merge GEMM of the 2D solve (M = 8 chi^2, N = K = chi^2,
    batch 32, the solve's two-level output index), per compute type: average of 3 launches after one
    warm-up launch, and bitwise equality with configuration 4 (the 256x256 K-contiguous tile)."""
    import torch

    from trg_backup_amd import _lib, ops
    log = lambda *a: print("[bench]", *a, file=sys.stderr, flush=True)
    lib = _lib.load()
    bt, M, N, K = 32, 32768, 4096, 4096
    torch.manual_seed(0)
    A = torch.randn(bt, M, K, device="cuda")
    B = torch.randn(bt, N, K, device="cuda")           # B(k, n) at n*K + k
    C, Cref = torch.empty(bt * M * N, device="cuda"), torch.empty(bt * M * N, device="cuda")
    lay = (M * N, 32768, 1, 4096, 64, 64, 512)          # [k,i,l,c,d] destination of the merge
    flops = 2.0 * M * N * K * bt
    rows = []
    for name, dt, cc in (("bf16 (bf16 operands)", _lib.BF16, _lib.BF16), ("split-bf16", _lib.F32, _lib.BF16X3),
                         ("three-piece split", _lib.F32, _lib.BF16X6)):
        a, b = (A.to(torch.bfloat16), B.to(torch.bfloat16)) if dt == _lib.BF16 else (A, B)
        n = lib.trg_gemm_config_count(dt, _lib.F32, cc)
        for cfg in [4] + [c for c in range(n) if c != 4]:
            out = Cref if cfg == 4 else C

            def go():
                _lib.call("trg_gemm_ex", cfg, dt, _lib.F32, cc, bt, M, N, K, 1.0, a.data_ptr(), M * K, K, 1,
                          b.data_ptr(), N * K, 1, K, 0.0, out.data_ptr(), *lay, ops.stream_ptr())
            go()
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            ev[0].record()
            for _ in range(3):
                go()
            ev[1].record()
            ev[1].synchronize()
            ms = ev[0].elapsed_time(ev[1]) / 3
            row = {"compute": name, "config": cfg,
                   "tile": lib.trg_gemm_config_name(dt, _lib.F32, cc, cfg).decode(), "ms": ms,
                   "tflops": flops / ms / 1e9, "bitwise_equal_to_config_4": cfg == 4 or bool(torch.equal(C, Cref))}
            rows.append(row)
            log(f"{name:22s} cfg {cfg:2d} {row['tile']:32s} {ms:8.2f} ms {row['tflops']:7.1f} TF/s "
                f"equal={row['bitwise_equal_to_config_4']}")
        del a, b
        torch.cuda.empty_cache()
    # library reference point: torch.bmm (hipBLASLt) on the same bf16 product, contiguous operands,
    # bf16 output in a plain row-major destination
    del C, Cref
    torch.cuda.empty_cache()
    a, b = A.to(torch.bfloat16), B.transpose(1, 2).contiguous().to(torch.bfloat16)
    del A, B
    c = torch.bmm(a, b)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for _ in range(3):
        torch.bmm(a, b, out=c)
    ev[1].record()
    ev[1].synchronize()
    ms = ev[0].elapsed_time(ev[1]) / 3
    lib_row = {"compute": "torch.bmm bf16 (hipBLASLt), plain layout", "ms": ms, "tflops": flops / ms / 1e9}
    log(f"torch.bmm bf16 {ms:8.2f} ms {lib_row['tflops']:7.1f} TF/s")
    print(json.dumps({"gemm_sweep": rows, "library": lib_row}), flush=True)
    return 0 if all(r["bitwise_equal_to_config_4"] for r in rows) else 1


def _copy_calibration(dev):
    """A plain 1 GiB float32 copy whose FETCH_SIZE / WRITE_SIZE calibrate the counters."""
    import torch

    from trg_backup_amd import ops
    src = ops.fill_(torch.empty(1 << 28, dtype=torch.float32, device=dev), 1.0)
    dst = torch.empty_like(src)
    ops.copy_into(dst, src)
    torch.cuda.synchronize()
    print("[probe] copy of 1 GiB float32 done", flush=True)


def traffic_probe_3d(args):
    """The batched 3D solve of the bench on 16 couplings (one merge sub-batch) up to the first
    steady three-piece split step (two chi^11 GEMM launches tuned, then counted), then the copy."""
    import torch

    from trg_backup_amd import _lib, autotune, hotrg, models, ops, workloads
    _lib.load()
    dev = torch.device("cuda:0")
    betas = workloads.hotrg3d_betas(16)
    policy = workloads.hotrg3d_policy(args.fp64_steps, args.merge3d)
    T = ops.contiguous(torch.stack([models.initial_tensor("ising3d", b, dev) for b in betas]))
    lnf = ops.fill_(torch.empty((len(betas), 16), dtype=torch.float64, device=dev), 0.0)
    hotrg.normalize_trace(T, lnf, 0)
    autotune.enable(not args.no_tune)
    for n, d in workloads.schedule3d(args.fp64_steps + 1):
        if n == args.fp64_steps + 1:
            hotrg.step3d(T, args.chi, d, policy(n))      # tune first, then the step whose dispatches are read
            torch.cuda.synchronize()
            print(f"[probe] counted step {n} {d} merge={policy(n).merge}", flush=True)
        T = hotrg.step3d(T, args.chi, d, policy(n))
        hotrg.normalize_trace(T, lnf, n)
        torch.cuda.synchronize()
    autotune.enable(False)
    _copy_calibration(dev)
    return 0


def traffic_probe(args):
    """One steady step of the bench batch with each merge precision (float32, bf16x6, bf16x3,
    bf16), then a plain 1 GiB float32 copy whose FETCH_SIZE / WRITE_SIZE calibrate the counters."""
    import torch

    from trg_backup_amd import _lib, autotune, hotrg, models, ops
    if args.workload == "hotrg3d":
        return traffic_probe_3d(args)
    _lib.load()
    dev = torch.device("cuda:0")
    chi, nb = args.chi, args.nbeta
    betas = [0.40 + 0.08 * i / (nb - 1) for i in range(nb)] if nb > 1 else [0.44]
    T = ops.contiguous(torch.stack([models.initial_tensor("ising2d", b, dev) for b in betas]))
    lnf = ops.fill_(torch.empty((nb, 16), dtype=torch.float64, device=dev), 0.0)
    n, d = 0, "Y"
    while T.shape[1] < chi or T.shape[3] < chi:
        n += 1
        T = hotrg.step2d(T, chi, d, hotrg.FP64)
        hotrg.normalize_maxabs(T, lnf, n)
        d = "X" if d == "Y" else "Y"
    T = ops.cast(T, torch.float32)
    if args.load_tune:
        autotune.load(args.load_tune)
    autotune.enable(not args.no_tune)
    for prec in (hotrg.FP32, hotrg.Precision(torch.float32, "f32", "bf16x6"), hotrg.Precision(torch.float32, "f32", "bf16x3"),
                 hotrg.Precision(torch.float32, "f32", "bf16")):
        # tune first, then the step whose dispatches are read
        hotrg.step2d(T, chi, d, prec)
        torch.cuda.synchronize()
        print(f"[probe] counted step {d} merge={prec.merge}", flush=True)
        T = hotrg.step2d(T, chi, d, prec)
        torch.cuda.synchronize()
    autotune.enable(False)
    _copy_calibration(dev)
    return 0


def dtype_label(policy, nsteps):
    """The arithmetic of each precision tier of a solve, in step order."""
    tiers, last = [], None
    for n in range(1, nsteps + 1):
        p = policy(n)
        if p.store.is_floating_point and p.store.itemsize == 8:
            key = ("f64 storage, merges as 6-slice int8 MFMA products (per-row fixed point)"
                   if p.merge == "f64i8" else "f64")
        else:
            key = f"env {p.env}, merge {p.merge}"
        if key != last:
            tiers.append([n, n, key])
            last = key
        else:
            tiers[-1][1] = n
    body = "; ".join(f"steps {a}-{b}: {k}" for a, b, k in tiers)
    return f"mixed ({body}; float32 storage after the f64 steps; f64 eigensolvers, traces and log-norms)"


def schedule(nsteps):
    """(n, direction) as trg/HOTRG_2d.py:195-214 runs them: Y first, then alternating."""
    from trg_backup_amd import workloads
    return workloads.schedule2d(nsteps)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU; --gpus above 1 is not supported "
                                   "(independent couplings shard across GPUs as whole flows, with no "
                                   "exchange inside a step)"}))
        return 2
    wl = args.workload
    args.chi = args.chi or {"hotrg2d": 64, "hotrg3d": 16, "hotrg3d_qr": 16, "atrg3d": 64}.get(wl, 64)
    args.nbeta = args.nbeta or {"atrg3d": 1, "hotrg3d": 128}.get(wl, 32)
    args.steps_requested = args.steps
    args.steps = args.steps or {"hotrg2d": 40, "hotrg3d": 24, "hotrg3d_qr": 6, "atrg3d": 30}.get(wl, 40)
    args.fp64_steps = (args.fp64_steps if args.fp64_steps is not None
                       else {"hotrg2d": 6, "hotrg3d": 6, "hotrg3d_qr": 10 ** 6, "atrg3d": 14}.get(wl, 0))
    if args.gemm_sweep:
        return gemm_sweep(args)
    if args.traffic_probe:
        return traffic_probe(args)
    if wl == "hotrg3d":
        return bench_hotrg3d(args)
    if wl == "hotrg3d_qr":
        return bench_hotrg3d_qr(args)
    if wl == "atrg3d":
        return bench_atrg3d(args)
    if wl == "serve":
        return bench_serve(args)
    import numpy as np
    import torch

    from trg_backup_amd import _lib, autotune, hotrg, models, ops, workloads

    _lib.load()
    dev = torch.device("cuda:0")
    # the solve is never shortened: --steps is a minimum, rounded up to whole solves
    K = args.solve_steps
    nsolve = max(1, -(-args.steps // K))
    chi, nb = args.chi, args.nbeta
    betas = workloads.hotrg2d_betas(nb)
    eig_q = () if args.eig_q_tiers == "none" else tuple(tuple(int(x) for x in t.split(":"))
                                                          for t in args.eig_q_tiers.split(","))
    policy = workloads.hotrg2d_policy(args.fp64_steps, args.split6_env_from, args.split6_merge_from,
                                      args.split_merge_from, args.bf16_env_from, args.bf16_merge_from, eig_q)
    log = lambda *a: print("[bench]", *a, file=sys.stderr, flush=True)
    T0 = torch.stack([models.initial_tensor("ising2d", b, dev) for b in betas])

    def run(steps, lnf):
        T = ops.contiguous(T0)
        hotrg.normalize_maxabs(T, lnf, 0)
        for n, d in steps:
            T = hotrg.step2d(T, chi, d, policy(n))
            hotrg.normalize_maxabs(T, lnf, n)
        return T

    # ---- warm-up: every precision tier after step 6 cut to its first W steps (untimed) ----
    warm, seen = [], {}
    for n, d in schedule(K):
        p = policy(n)
        key = (p.store, p.env, p.merge)
        if n <= 6 or seen.get(key, 0) < args.warmup:
            warm.append((n, d))
            if n > 6:
                seen[key] = seen.get(key, 0) + 1
    if args.load_tune:
        autotune.load(args.load_tune)
    autotune.enable(not args.no_tune)
    t0 = time.perf_counter()
    lw = ops.fill_(torch.empty((nb, K + 1), dtype=torch.float64, device=dev), 0.0)
    run(warm, lw)
    torch.cuda.synchronize()
    autotune.enable(False)
    if args.save_tune:
        autotune.save(args.save_tune)
    log(f"warm-up {len(warm)} steps in {time.perf_counter() - t0:.1f} s; "
        f"{len(autotune.table())} GEMM problems tuned")
    del lw
    torch.cuda.empty_cache()

    # ---- timed: nsolve whole solves from the initial tensor ----
    lnf = ops.fill_(torch.empty((nb, K + 1), dtype=torch.float64, device=dev), 0.0)
    rec = ops.GemmRecorder("merge2d_z2")
    health = ops.Health(dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with rec, health:
        for _ in range(nsolve):
            T = run(schedule(K), lnf)
            z = hotrg.lnz2d(T, lnf, K)
        torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    health.check("timed solve")
    Kt = K * nsolve
    value = nb * Kt / dt
    zt = z.cpu().numpy()
    log(f"timed {nsolve} solve(s) of {K} steps x {nb} instances in {dt:.2f} s -> {value:.2f} RG steps/s")
    del T
    torch.cuda.empty_cache()

    # merge GEMM launches grouped by (storage, compute)
    names = {_lib.F32: "f32", _lib.BF16: "bf16", _lib.BF16X3: "bf16x3", _lib.BF16X6: "bf16x6", _lib.F64: "f64"}
    groups = {}
    for (e0, e1, fl), cfg in zip(rec.events, rec.event_configs):
        groups.setdefault(names[cfg[2]], []).append((e0.elapsed_time(e1), fl, cfg))
    kern = {}
    for nm, evs in groups.items():
        ms = sum(e[0] for e in evs)
        fl = sum(e[1] for e in evs)
        ach = fl / (ms * 1e-3) / 1e12
        kern[nm] = {"launches": len(evs), "avg_ms": ms / len(evs), "total_ms": ms,
                    "flops_per_launch": fl / len(evs), "achieved_tflops": ach,
                    "peak_tflops": PEAK_TFLOPS[nm], "frac": ach / PEAK_TFLOPS[nm],
                    "mfma_issue_frac": ach * MFMA_PER_PRODUCT[nm] / PEAK_TFLOPS[nm],
                    "tiles": sorted({autotune.config_name(c[0], c[2], c[3], c[1]) for _, _, c in evs})}
    dom = max(kern, key=lambda k: kern[k]["total_ms"])
    d = kern[dom]

    traffic = None
    tp = os.path.join(ROOT, "profiles", "gemm_merge2d_z2_traffic.json")
    if os.path.exists(tp):
        with open(tp) as f:
            traffic = json.load(f).get("bytes_per_launch", {}).get(dom)

    out = {
        "metric": "RG steps/sec and p50/p99 solve latency at |dlnZ/V|<=1e-8 vs FP64 plain path",
        "value": value, "unit": "RG steps/s", "n_gpus": 1, "steps": Kt, "warmup": args.warmup,
        "ms_per_step": dt / Kt * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": dtype_label(policy, K), "data": "synthetic",
        "config": {"workload": f"2D Ising HOTRG chi={chi}, {K} RG steps from the initial tensor "
                               f"(Y first, alternating), {nb} beta in [0.40,0.48] batched",
                   "chi": chi, "nbeta": nb, "rg_steps": K, "solves_timed": nsolve,
                   "steps_requested": args.steps_requested, "policy": policy.describe(),
                   "parallelism": "single GPU, batched"},
        "steps_note": "a bench step is one RG step of all nbeta couplings along one axis; --steps is a minimum, "
                      "rounded up to whole solves, so steps = rg_steps * solves_timed",
        "roofline": {"bound": "mfma", "achieved": d["achieved_tflops"], "peak": d["peak_tflops"],
                     "unit": "TFLOP/s", "frac": d["frac"], "traffic": traffic,
                     "kernel": f"merge chi^7 GEMM ({dom} MFMA, tiles {', '.join(d['tiles'])}; "
                               "M=nc*chi^2 N=chi^2 K=chi^2, batch=nbeta)",
                     "launches": d["launches"], "avg_ms": d["avg_ms"], "flops_per_launch": d["flops_per_launch"],
                     "mfma_issue_frac": d["mfma_issue_frac"],
                     "note": "achieved = algorithmic 2MNK flops / live launch time; peak = published dense peak "
                             "of the MFMA type issued; mfma_issue_frac counts the MFMA work issued (bf16x3: "
                             "3 bf16 products per algorithmic product, bf16x6: 6)",
                     "merge_gemm_by_precision": kern},
    }
    out["lnZ_per_site"] = {"betas": [betas[0], betas[nb // 2], betas[-1]],
                           "values": [float(zt[0]), float(zt[nb // 2]), float(zt[-1])]}

    # ---- accuracy against the float64 path (untimed) ----
    if not args.no_validate:
        idx = list(range(nb)) if (args.validate_all or args.validate_betas <= 0) else sorted(
            set(np.linspace(0, nb - 1, max(2, args.validate_betas)).round().astype(int).tolist()) | {nb // 2})
        errs = []
        for c0 in range(0, len(idx), 16):
            sel = idx[c0:c0 + 16]
            Td, ld = hotrg.renorm2d(ops.contiguous(T0[sel]), chi, K // 2, K - K // 2, hotrg.ALL_FP64)
            zd = hotrg.lnz2d(Td, ld, K).cpu().numpy()
            errs += [abs(float(zt[i]) - float(zd[j])) for j, i in enumerate(sel)]
            del Td
            torch.cuda.empty_cache()
            log(f"validated betas {[round(betas[i], 4) for i in sel]}: max err so far {max(errs):.3e}")
        out["lnZ_err_vs_fp64"] = max(errs)
        out["lnZ_check"] = {"betas": [betas[i] for i in idx], "abs_err": errs, "n": len(errs),
                            "all_within_1e-8": bool(max(errs) <= 1e-8)}

    # ---- plain per-instance float64 path: the same solve, one instance, host sync per step ----
    if not args.no_plain:
        T1 = ops.contiguous(T0[nb // 2:nb // 2 + 1])
        l1 = ops.fill_(torch.empty((1, K + 1), dtype=torch.float64, device=dev), 0.0)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        hotrg.normalize_maxabs(T1, l1, 0)
        for n, dd in schedule(K):
            T1 = hotrg.step2d(T1, chi, dd, hotrg.FP64)
            hotrg.normalize_maxabs(T1, l1, n)
            torch.cuda.synchronize()
        hotrg.lnz2d(T1, l1, K).cpu()
        tp1 = time.perf_counter() - t1
        out["plain_fp64"] = {"steps_per_s": K / tp1, "solve_s": tp1, "beta": betas[nb // 2]}
        out["speedup_vs_plain_fp64"] = value / out["plain_fp64"]["steps_per_s"]
        log(f"plain float64 solve {tp1:.1f} s")
        del T1
        torch.cuda.empty_cache()

    # ---- serving: a Poisson stream of mixed solves (latency) ----
    if not args.no_serve:
        res = serve_leg(args)
        out["serving"] = res
        out["p50_latency_s"] = res["p50_latency_s"]
        out["p99_latency_s"] = res["p99_latency_s"]
        log(f"serving: {res}")

    # ---- CPU baseline: the NumPy float64 oracle on the host ----
    if not args.no_cpu:
        from oracle import HOTRG_2d as oh2
        from oracle import models as omodels
        Tc = omodels.ising2d_tensor(0.44)
        s, dd = 0, "Y"
        while Tc.shape[0] < chi or Tc.shape[2] < chi:
            s += 1
            Tc, _ = oh2.__tensor_normalization__(oh2.new_pure_tensor(Tc, chi, dd))
            dd = "X" if dd == "Y" else "Y"
        t1 = time.perf_counter()
        Tc, _ = oh2.__tensor_normalization__(oh2.new_pure_tensor(Tc, chi, dd))
        tc = time.perf_counter() - t1
        try:
            from threadpoolctl import threadpool_info
            cores = max(d["num_threads"] for d in threadpool_info() if d.get("user_api") == "blas")
        except (ImportError, ValueError):
            cores = os.cpu_count()
        out["cpu_baseline"] = {"value": 1.0 / tc, "unit": "RG steps/s", "cores": cores, "kind": "port",
                               "sample": f"one steady-state chi={chi} {dd} step of one instance (beta=0.44), "
                                         f"NumPy/OpenBLAS float64 oracle, {tc:.1f} s"}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
