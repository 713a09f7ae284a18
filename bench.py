"""
Benchmark: 2-population split_mig integrations per second at pts=100, FP64,
batched on one MI355X, beside the plain one-integration-per-call GPU path and
the reference C path on the host.

Workload (seeded, numpy.random.default_rng): nu1, nu2 ~ logU[0.1, 10],
T ~ U[0.01, 2], m ~ U[0, 5]; phi0 = phi_1D_to_2D(phi_1D) on default_grid(100);
one unit = one two_pops(phi0, xx, T, nu1, nu2, m12=m, m21=m) integration (the
constant-parameter path, bit-identical to dadi's CPU result).

A step = one launch integrating --batch instances (resident in HBM, reset
from phi0 by a device copy before each launch): the register-resident kernel
(each integration's phi held in a workgroup's registers from start to end)
where the grid size has one, else the packed streaming kernel.  value =
batch * steps / wall time of the timed steps.  The packed kernel is also
timed on the same batch for comparison (and must give the same bits).

Prints one JSON line.  --gpus > 1 is refused: this build uses one GPU.
"""
import argparse
import json
import multiprocessing
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12


def draws(rng, k):
    nu1 = np.exp(rng.uniform(np.log(0.1), np.log(10), k))
    nu2 = np.exp(rng.uniform(np.log(0.1), np.log(10), k))
    T = rng.uniform(0.01, 2, k)
    m = rng.uniform(0, 5, k)
    return list(zip(nu1, nu2, T, m))


def build_jobs(dadi_amd, params, pts):
    from dadi_amd import schedule as S, _lib
    xx = dadi_amd.Numerics.default_grid(pts)
    phi0 = dadi_amd.PhiManip.phi_1D_to_2D(xx, dadi_amd.PhiManip.phi_1D(xx))
    segs = []
    for nu1, nu2, T, m in params:
        dt = min(S.compute_dt(nu1, [m], 0, 0.5), S.compute_dt(nu2, [m], 0, 0.5))
        segs.append(S.const_segments(2, xx, 0.0, T, dt, [nu1, nu2], [[m], [m]], [0, 0],
                                     [0.5, 0.5], 1.0))
    return xx, phi0, segs


def _cpu_share(task):
    """One core's share of the CPU baseline: integrate draws until the
    deadline (time.monotonic, shared by the processes)."""
    params, pts, deadline = task
    from oracle import dadi_oracle as O
    backend = "ref" if O.ref_available() else "oracle"
    xx = O.default_grid(pts)
    phi0 = O.phi_1D_to_2D(xx, O.phi_1D(xx))
    done = steps = 0
    t0 = time.monotonic()
    for nu1, nu2, T, m in params:
        if time.monotonic() >= deadline:
            break
        O.integrate_const(phi0, xx, T, [nu1, nu2], [[m], [m]], [0, 0], [0.5, 0.5],
                          backend=backend)
        dt = min(O.compute_dt(nu1, [m], 0, 0.5), O.compute_dt(nu2, [m], 0, 0.5))
        steps += int(np.ceil(T / dt))
        done += 1
    return done, steps, time.monotonic() - t0


#: host cores a one-GPU job may keep busy on the GPU boxes (their CPU share per GPU)
CPU_SHARE_PER_GPU = 16


def host_cores():
    """What the host offers this process: os.cpu_count(), the scheduler
    affinity mask, the cgroup CPU quota (cores, None without one), and the
    processes the CPU baseline runs: one per core it may use, i.e. the
    smallest of the affinity, the quota and the per-GPU CPU share."""
    count = os.cpu_count() or 1
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = count
    quota = None
    for path, parse in (("/sys/fs/cgroup/cpu.max", lambda t: t.split()),
                        ("/sys/fs/cgroup/cpu/cpu.cfs_quota_us", None)):
        try:
            txt = open(path).read().strip()
        except OSError:
            continue
        if parse:
            q, per = parse(txt)
            if q != "max":
                quota = float(q) / float(per)
        else:
            q = float(txt)
            per = float(open("/sys/fs/cgroup/cpu/cpu.cfs_period_us").read())
            if q > 0:
                quota = q / per
        break
    use = min(aff, CPU_SHARE_PER_GPU)
    if quota:
        use = min(use, max(1, int(quota)))
    return {"os_cpu_count": count, "affinity": aff, "cgroup_quota_cores": quota,
            "share_per_gpu": CPU_SHARE_PER_GPU, "used": use}


def cpu_baseline(params, pts, budget_s, procs):
    """The reference's own C kernels (oracle/_ref, compiled from dadi's
    sources) driven by a restatement of dadi's constant-parameter Python loop,
    one process per core over disjoint shares of the draws (the oracle's C
    restatement when _ref is absent).  Runs before the GPU is initialised:
    the processes are forked."""
    from oracle import dadi_oracle as O
    kind = "reference" if O.ref_available() else "port"
    procs = max(1, int(procs))
    t0 = time.monotonic()
    tasks = [(params[i::procs], pts, t0 + budget_s) for i in range(procs)]
    if procs == 1:
        res = [_cpu_share(tasks[0])]
    else:
        with multiprocessing.get_context("fork").Pool(procs) as pool:
            res = pool.map(_cpu_share, tasks)
    el = time.monotonic() - t0
    done = sum(r[0] for r in res)
    steps = sum(r[1] for r in res)
    per_core = [r[0] / r[2] for r in res if r[0]]
    return {"value": done / el, "unit": "integrations/s", "cores": procs, "kind": kind,
            "single_core": float(np.mean(per_core)) if per_core else None,
            "sample": "%d draws of the same seeded split_mig distribution at pts=%d (%d time "
                      "steps), full two_pops integrations, %d processes x 1 thread, %.1f s"
                      % (done, pts, steps, procs, el)}


def model_workloads(args, rng):
    """Whole model evaluations through the public API, many in flight at once
    (engine.map_batched coalesces their integrations into shared launches):
    split_mig and the 3-population chain extrapolated over three grids, and
    the DFE caches (Cache1D of two_epoch_sel; optionally Cache2D of
    split_mig_sel); optionally one 3- or 4-population integration's step time
    on the multi-workgroup sweeps."""
    import torch
    from dadi_amd import engine, Numerics, Demographics2D, Demographics3D
    from dadi_amd.DFE import DemogSelModels, Cache1D_mod, Cache2D_mod
    out = {}

    def timed(fn):
        t0 = time.perf_counter()
        r = fn()
        return r, time.perf_counter() - t0

    if args.evals > 0:
        ex = Numerics.make_extrap_log_func(Demographics2D.split_mig)
        ps = draws(rng, args.evals)
        res, el = timed(lambda: engine.map_batched(
            lambda p: ex(p, (20, 20), [80, 90, 100]), [(p,) for p in ps], workers=args.workers))
        out["split_mig_evaluations"] = {
            "value": len(ps) / el, "unit": "evaluations/s", "evaluations": len(ps),
            "failed": sum(isinstance(r, Exception) for r in res), "wall_s": el,
            "model": "split_mig, pts (80, 90, 100) log-extrapolated, ns (20, 20)"}
    if args.evals3d > 0:
        ex3 = Numerics.make_extrap_log_func(Demographics3D.three_pop_chain)
        lu = lambda: float(np.exp(rng.uniform(np.log(0.1), np.log(10))))
        ps = [(lu(), rng.uniform(0.01, 1), lu(), lu(), rng.uniform(0.01, 1), rng.uniform(0, 5),
               lu(), lu(), lu(), rng.uniform(0.01, 1), rng.uniform(0, 5))
              for _ in range(args.evals3d)]
        res, el = timed(lambda: engine.map_batched(
            lambda p: ex3(p, (10, 10, 10), [40, 50, 60]), [(p,) for p in ps],
            workers=args.workers))
        out["chain3d_evaluations"] = {
            "value": len(ps) / el, "unit": "evaluations/s", "evaluations": len(ps),
            "failed": sum(isinstance(r, Exception) for r in res), "wall_s": el,
            "model": "1D -> 2D -> 3D chain, pts (40, 50, 60) log-extrapolated, ns (10, 10, 10)"}
    if args.opt:
        # BFGS in log(params) on a seeded Poisson sample of a split_mig model:
        # one evaluation at a time (as dadi) and with each point's
        # finite-difference stencil evaluated as one batch (same trajectory)
        from dadi_amd import Inference, Spectrum
        ex = Numerics.make_extrap_log_func(Demographics2D.split_mig)
        grids, ns = [40, 50, 60], (20, 20)
        truth = [1.8, 0.6, 0.4, 1.5]
        model = ex(truth, ns, grids)
        data = Spectrum(rng.poisson(5000 * np.maximum(model.data, 0)).astype(float))
        data.mask = model.mask
        kw = dict(lower_bound=[1e-2, 1e-2, 1e-3, 0], upper_bound=[20, 20, 5, 10],
                  maxiter=args.opt, full_output=True)
        p0 = [1.0, 1.0, 0.5, 1.0]
        ref, el_s = timed(lambda: Inference.optimize_log(p0, data, ex, grids, batch=False, **kw))
        got, el_b = timed(lambda: Inference.optimize_log(p0, data, ex, grids, batch=True, **kw))
        out["optimize_log"] = {
            "value": el_s / el_b, "unit": "x (serial time / batched time)",
            "serial_s": el_s, "batched_s": el_b, "iterations": args.opt,
            "evaluations": Inference.last_model_cache.batched,
            "same_result": bool(np.array_equal(ref[0], got[0]) and ref[1] == got[1]),
            "model": "split_mig, pts (40, 50, 60) log-extrapolated, ns (20, 20), 4 parameters"}
    if args.dfe1d:
        c1, el = timed(lambda: Cache1D_mod.Cache1D((2.0, 0.1), (50,),
                                                   DemogSelModels.two_epoch_sel,
                                                   (200, 220, 240), gamma_pts=500))
        out["dfe_cache1d"] = {"value": el, "unit": "s", "spectra": len(c1.gammas) + 1,
                              "model": "two_epoch_sel nu=2 T=0.1, 500 gammas, pts (200, 220, "
                                       "240), ns 50"}
    if args.sweep:
        # one 3- or 4-population integration on the multi-workgroup path
        # (dh_integrate_batch), constant parameters, the serving loop's and
        # the chain evaluations' epochs at one size: microseconds per step
        from dadi_amd import schedule as S, _lib
        D, pts, nst = (int(v) for v in args.sweep.split(","))
        ctx = engine.context()
        xx = Numerics.default_grid(pts)
        seg = S.const_segments(D, xx, 0.0, 0.05, 0.05 / nst, [1.3] * D, [[0.7] * (D - 1)] * D,
                               [0.0] * D, [0.5] * D, 1.0)
        order = _lib.ORDER_PRE if D <= 3 else _lib.ORDER_TV
        init = np.random.default_rng(0).uniform(0.5, 1.0, (pts,) * D)
        phi = ctx.empty_phi(init.shape)
        best = None
        for _ in range(2):
            phi.copy_(torch.from_numpy(init))
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            engine.PackedBatch(ctx, [engine.Job(phi, xx, seg, order, 0, 1)]).launch_spread(
                ctx.stream())
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            best = el if best is None else min(best, el)
        out["sweep_steps"] = {"value": 1e6 * best / nst, "unit": "us/step", "steps": nst,
                              "model": "one %d-population integration, pts %d, constant "
                                       "parameters, multi-workgroup sweeps" % (D, pts)}
    if args.dfe2d:
        nu1, nu2, T, m = draws(rng, 1)[0]
        c2, el = timed(lambda: Cache2D_mod.Cache2D((nu1, nu2, T, m), (20, 20),
                                                   DemogSelModels.split_mig_sel, (60, 70, 80),
                                                   gamma_pts=10))
        out["dfe_cache2d"] = {"value": el, "unit": "s", "spectra": 100,
                              "model": "split_mig_sel nu1=%.4g nu2=%.4g T=%.4g m=%.4g, 10x10 "
                                       "gammas, pts (60, 70, 80), ns (20, 20)"
                                       % (nu1, nu2, T, m)}
    return out


def measured_traffic(pts, batch, seed, kernel):
    """HBM bytes per launch of this exact workload and kernel, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summaries
    (profiles/traffic_*.json, each counter scaled by what it reads for a
    device copy of known size); None when no summary matches."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "traffic_*.json"))):
        try:
            rec = json.load(open(path))
        except (OSError, ValueError):
            continue
        if rec.get("pts") == pts and rec.get("batch") == batch and rec.get("seed") == seed \
                and rec.get("kernel") == kernel:
            return (rec["fetch_kib"] * rec["fetch_scale"] +
                    rec["write_kib"] * rec["write_scale"]) * 1024
    return None


def kernel_of(pb, packed=False):
    """Name (as rocprofv3 lists it, without arguments) of the kernel that
    pb.launch() runs for a 2-population PRE-order batch: the
    register-resident kernel where it applies, else the factored kernel for
    few integrations, else the packed one; packed=True: the packed kernel."""
    if not packed and pb.resident_ok():
        ctx = pb.ctx
        rows = ctx.lib.dh_resident_rows(pb.nmax)
        if len(pb.jobs) <= ctx.resident_latency_max(pb.nmax) or not rows:
            return "k_integrate_rf<%d, 1, true>" % ctx.lib.dh_resident_latency_rows(pb.nmax)
        return "k_integrate_rr<%d, 1, true>" % rows
    if not packed and pb.latency_bound():
        return "k_integrate_fz<2, 1>"
    return "k_integrate_packed<2, 1>"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--pts", type=int, default=100)
    ap.add_argument("--plain", type=int, default=12)
    ap.add_argument("--waves", type=int, default=0)
    ap.add_argument("--chunk", type=int, default=-1)
    ap.add_argument("--recompute", type=int, default=-1)
    ap.add_argument("--gamma-global", type=int, default=-1)
    ap.add_argument("--slots", type=int, default=-1,
                    help="packed kernel: instances per workgroup (1 = unpacked kernel; "
                         "-1 = automatic)")
    ap.add_argument("--kernel", choices=("auto", "resident", "packed"), default="auto",
                    help="batched launch: the register-resident kernel where the grid size "
                         "has one (auto), always it, or the packed streaming kernel")
    ap.add_argument("--compare-packed", type=int, default=1,
                    help="also time one packed-kernel launch over the same batch")
    ap.add_argument("--dense", action="store_true",
                    help="store phi rows unpadded (default: rows padded to 64-byte multiples)")
    ap.add_argument("--groups", type=int, default=0,
                    help="persistent workgroups (0 = as many as fit)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-procs", type=int, default=0,
                    help="CPU baseline processes (0 = one per core this job may use: the "
                         "affinity mask, capped by the cgroup quota and the per-GPU CPU share)")
    ap.add_argument("--evals", type=int, default=512,
                    help="split_mig model evaluations (3 grids each) through the public API")
    ap.add_argument("--evals3d", type=int, default=96,
                    help="3-population chain evaluations (3 grids each)")
    ap.add_argument("--workers", type=int, default=256,
                    help="threads evaluating models concurrently (each runs the grids of "
                         "its extrapolated evaluation in two more threads)")
    ap.add_argument("--dfe1d", type=int, default=1, help="time a Cache1D build (500 gammas)")
    ap.add_argument("--opt", type=int, default=4,
                    help="BFGS iterations of the optimiser comparison (0 = skip)")
    ap.add_argument("--dfe2d", type=int, default=1,
                    help="time a Cache2D build (10x10 gammas; long: up to ~5e5*T steps)")
    ap.add_argument("--sweep", default="",
                    help="D,pts,steps: time one D-population integration (D = 3, 4) on the "
                         "multi-workgroup sweeps; '' = skip")
    ap.add_argument("--routing", default="512,2048",
                    help="batch sizes (comma-separated, from the same draws) timed with the "
                         "automatic kernel choice and with the packed kernel; '' = skip")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-warmup", type=int, default=20,
                    help="requests of a short untimed mix served first (kernel loading, "
                         "allocator growth)")
    ap.add_argument("--seed", type=int, default=20250101)
    ap.add_argument("--calibrate-copy", type=float, default=0.0,
                    help="only run 3 device copies of this many GiB (rocprofv3 counter "
                         "calibration of FETCH_SIZE/WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.calibrate_copy > 0:
        import torch
        n = int(args.calibrate_copy * (1 << 30)) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_each": n * 8, "copies": 3}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: every integration is independent and this "
              "build runs on a single MI355X; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    rng = np.random.default_rng(args.seed)
    params = draws(rng, args.batch)
    cpu = None
    if not args.no_cpu:
        # before anything initialises the GPU: the baseline forks its processes
        hc = host_cores()
        procs = args.cpu_procs or hc["used"]
        cpu = cpu_baseline(params, args.pts, args.cpu_seconds, procs)
        cpu["host"] = hc

    import torch
    import dadi_amd
    from dadi_amd import engine, _lib
    if not dadi_amd.cuda_enabled(True):
        raise SystemExit("HIP backend unavailable")
    ctx = engine.context()
    xx, phi0, segs = build_jobs(dadi_amd, params, args.pts)
    n = args.pts
    phi0d = ctx.to_dev(phi0)
    # phi rows padded to a multiple of 8 doubles (64-byte aligned rows)
    if args.dense:
        arena = torch.empty((args.batch, n, n), dtype=torch.float64, device=ctx.device)
    else:
        arena = ctx.empty_phi((n, n), batch=args.batch)
    arena.copy_(phi0d.expand(args.batch, n, n))
    jobs = [engine.Job(arena[i], xx, segs[i], _lib.ORDER_PRE, 0, 3) for i in range(args.batch)]
    order = sorted(range(args.batch), key=lambda i: -jobs[i].work())
    pb = engine.PackedBatch(ctx, [jobs[i] for i in order])
    steps_total = sum(int(np.sum(s["nsteps"])) for s in segs)
    alg_bytes = 16.0 * n * n * 2 * steps_total   # read+write phi per point per axis sweep

    def opt(v):
        return None if v < 0 else v

    shape = ctx.shape(2, n, args.waves or None, opt(args.chunk), opt(args.recompute),
                      opt(args.gamma_global), opt(args.slots), packable=pb.packable)
    resident = args.kernel == "resident" or (args.kernel == "auto" and pb.resident_ok())
    if resident and not pb.resident_ok():
        raise SystemExit("no register-resident kernel for pts=%d" % n)

    def launch_packed():
        pb.launch(waves=shape.waves, groups=args.groups, chunk=shape.chunk,
                  recompute=shape.recompute, gamma_global=shape.gamma_global,
                  slots=shape.slots, resident=False)

    def launch():
        if resident:
            pb.launch_resident(groups=args.groups)
        else:
            launch_packed()

    def one_step():
        arena.copy_(phi0d.expand(args.batch, n, n))
        pb.restart()
        launch()

    for _ in range(args.warmup):
        one_step()
    torch.cuda.synchronize()
    if not args.no_check:
        # correctness spot check of the batch against the plain path (bit-exact)
        spot = order[0]
        # (another kernel: the factored one)
        solo = engine.run_single(phi0.copy(), xx, segs[spot], _lib.ORDER_PRE, 0, 3,
                                 mode="factored")
        one_step()
        torch.cuda.synchronize()
        assert np.array_equal(arena[spot].cpu().numpy(), solo), "batched != plain result"

    kern_ms = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        arena.copy_(phi0d.expand(args.batch, n, n))
        pb.restart()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        launch()
        e1.record()
        kern_ms.append((e0, e1))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern = [a.elapsed_time(b) for a, b in kern_ms]
    kavg = float(np.mean(kern)) / 1e3
    value = args.batch * args.steps / wall
    headline_phi = arena.cpu().numpy() if resident and args.compare_packed else None
    if resident:
        flagged = int(np.sum(pb.tab.cpu().numpy().view(engine.INST_DTYPE)["reserved"] & 1))

    # the packed streaming kernel over the same batch, for comparison
    packed = None
    if resident and args.compare_packed:
        arena.copy_(phi0d.expand(args.batch, n, n))
        pb.restart()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        launch_packed()
        e1.record()
        torch.cuda.synchronize()
        pk = e0.elapsed_time(e1) / 1e3
        packed = {"kernel": kernel_of(pb, packed=True), "kernel_ms": pk * 1e3,
                  "value": args.batch / pk, "unit": "integrations/s",
                  "roofline_frac": alg_bytes / pk / PEAK_HBM,
                  "same_result": bool(np.array_equal(arena.cpu().numpy(), headline_phi))}
        del headline_phi

    # plain path: one integration per call (same draws), host copies included
    lat = []
    for i in range(args.plain):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        engine.run_single(phi0.copy(), xx, segs[i], _lib.ORDER_PRE, 0, 3)
        lat.append(time.perf_counter() - t1)
    plain_rate = len(lat) / sum(lat) if lat else None
    plain_steps = sum(int(np.sum(segs[i]["nsteps"])) for i in range(args.plain))

    # kernel choice for smaller batches: the automatic route (the factored
    # kernel up to engine.Context.factored_max_inst integrations) against the
    # packed kernel on the same integrations
    routing = []
    for B in [int(v) for v in args.routing.split(",") if v.strip()]:
        B = min(B, args.batch)
        ar = ctx.empty_phi((n, n), batch=B)
        rjobs = [engine.Job(ar[i], xx, segs[i], _lib.ORDER_PRE, 0, 3) for i in range(B)]
        rpb = engine.PackedBatch(ctx, sorted(rjobs, key=lambda j: -j.work()))
        rates, outs = {}, {}
        for mode in ("auto", "packed"):
            def go():
                ar.copy_(phi0d.expand(B, n, n))
                rpb.restart()
                if mode == "auto":
                    rpb.launch()
                else:
                    rpb.launch(waves=shape.waves, chunk=shape.chunk, recompute=shape.recompute,
                               gamma_global=shape.gamma_global, slots=shape.slots,
                               resident=False)
            go()
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            go()
            torch.cuda.synchronize()
            rates[mode] = B / (time.perf_counter() - t1)
            outs[mode] = ar.cpu().numpy()
        routing.append({"batch": B, "unit": "integrations/s",
                        "auto_kernel": kernel_of(rpb),
                        "auto": rates["auto"], "packed": rates["packed"],
                        "same_result": bool(np.array_equal(outs["auto"], outs["packed"]))})
        del rpb, rjobs, ar

    workloads = model_workloads(args, np.random.default_rng(args.seed + 2))

    # serving last: a Server run left the host-bound model workloads above
    # slower when they ran after it (DFE Cache2D 7.6 -> 12.1 s after two runs)
    serving = None
    if args.serve_requests > 0:
        from dadi_amd import serve
        if args.serve_warmup > 0:   # untimed: the same mix with epochs 20x shorter
            serve.Server().run(serve.mixed_requests(np.random.default_rng(args.seed + 3),
                                                    args.serve_warmup, 200.0, scale=0.05),
                               realtime=False)
        reqs = serve.mixed_requests(np.random.default_rng(args.seed + 1), args.serve_requests,
                                    args.serve_rate)
        serving = serve.Server().run(reqs)
        serving.update({"warmup_requests": args.serve_warmup})
        serving.update({"rate_req_s": args.serve_rate, "mix": "55% 2D split_mig pts 40-120, "
                        "30% 1D three_epoch pts 100-600, 12% 3D chain pts 30-60, "
                        "3% 4D chain pts 16-24; Poisson arrivals, seeded"})

    achieved = alg_bytes / kavg
    if resident:
        kname = kernel_of(pb)
    elif shape.slots >= 2:
        kname = "k_integrate_packed<2, 1>"
    else:
        kname = "k_integrate_nd<2, 1, %d, %d, %d>" % (shape.chunk, shape.recompute,
                                                     shape.gamma_global)
    traffic = None if args.dense else measured_traffic(n, args.batch, args.seed, kname)
    if resident:
        launch_cfg = {"parallelism": "single GPU, one workgroup of %d waves per integration "
                                     "at a time (persistent), phi held in registers"
                                     % ((n + 63) // 64),
                      "instances_per_workgroup": 1,
                      "integrations_on_exact_division_kernel": flagged}
    else:
        launch_cfg = {"parallelism": "single GPU, persistent workgroups, %d integration(s) "
                                     "each" % shape.slots,
                      "waves_per_workgroup": shape.waves, "thomas_chunk": shape.chunk,
                      "thomas_recompute": shape.recompute,
                      "thomas_gamma_global": shape.gamma_global,
                      "instances_per_workgroup": shape.slots}
    out = {
        "metric": "2D split_mig integrations/s at pts=100 FP64, batched vs plain; p99 req latency",
        "value": value,
        "unit": "integrations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded parameter draws; phi from phi_1D_to_2D(phi_1D))",
        "config": dict({"workload": "2D split_mig two_pops integrations, pts=%d, batch of %d "
                                    "seeded draws per launch" % (n, args.batch),
                        "pts": n, "batch": args.batch, "time_steps_per_batch": steps_total,
                        "row_pitch": engine.pitch_of(arena[0]) or n}, **launch_cfg),
        "plain": {"value": plain_rate, "unit": "integrations/s", "calls": len(lat),
                  "time_steps": plain_steps,
                  "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3) if lat else None,
                                 "p99": float(np.percentile(lat, 99) * 1e3) if lat else None}},
        "batched_over_plain": value / plain_rate if plain_rate else None,
        "packed_kernel": packed,
        "routing": routing,
        "serving": serving,
        "model_workloads": workloads,
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9,
                     "unit": "GB/s", "frac": achieved / PEAK_HBM, "traffic": traffic,
                     "kernel": kname, "kernel_ms_avg": kavg * 1e3,
                     "algorithmic_bytes_per_launch": alg_bytes},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
