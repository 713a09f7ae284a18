"""
Continuous-batching request server for model evaluations on one GPU.

A request is one model evaluation at one grid size: a short program of
operations on a phi that lives on the device from start to finish -- its
initial density, integration epochs (1-5 populations), population splits
between epochs, and the final sampling into a spectrum.  The server admits
requests as they arrive, as long as the HBM footprint of everything in flight
stays within a budget, and advances all in-flight integration epochs
together: every quantum is one persistent launch per (populations, order)
class, each on its own stream so the classes overlap on the device, and each
instance advances up to ``quantum`` time steps (its progress is kept in the
instance table between launches).  A request whose epoch finishes runs its
next operations (device-side splits, sampling) and either starts its next
epoch in the following quantum or completes.  New arrivals join at the next
quantum -- they do not wait for a batch to drain.

Quanta are sized in time, not steps: in a launch every instance advances as
many steps as fit in about ``quantum_ms`` (a wall-clock budget checked by
the kernel after each step), so a small integration is not held to the step
count of the largest one it shares the launch with, and classes of cheap
steps (1D, small 2D) are not held back behind classes of expensive ones.
Where a kernel has no time budget (the lane-per-instance 1D kernel, the
multi-workgroup path) the class's step count per launch adapts toward the
same duration.
Epochs whose sweeps have many lines (3D and 4D phi) would run their steps
serially on one workgroup in the batched launch; they take the
multi-workgroup path instead (every sweep of a step spread over the whole
GPU, steps replayed from a captured hipGraph), advanced one quantum at a time
on their own streams, so their latency is that of the per-step sweeps rather
than of one compute unit.

Every request's arithmetic is independent of what it shares a launch with,
so a result is bit-identical alone and in any mix.
"""
import math
import time
from dataclasses import dataclass

import numpy as np

from . import _lib, engine, Numerics, PhiManip, Integration
from . import schedule as S
from ._lib import INST_DTYPE


# ---------------------------------------------------------------------------
# request programs
# ---------------------------------------------------------------------------
@dataclass
class Epoch:
    D: int
    seg: np.ndarray
    order: int
    frozen: int
    inject: int


@dataclass
class Program:
    """init -> [epoch | split]* -> sample; ops are ('epoch', Epoch),
    ('split', callable(phi_dev) -> phi_dev)."""
    pts: int
    xx: np.ndarray
    phi0: np.ndarray
    ops: list
    ns: tuple
    kind: str = ""

    def max_points(self):
        D = 1
        best = self.pts
        for op, arg in self.ops:
            if op == "epoch":
                D = max(D, arg.D)
        return self.pts ** D


def _const_epoch(D, xx, T, nus, migs, gammas=None, hs=None, theta0=1.0, frozen=None):
    gammas = gammas or [0.0] * D
    hs = hs or [0.5] * D
    frozen = frozen or [False] * D
    dt = min(S.compute_dt(nus[k], migs[k] if D > 1 else [0], gammas[k], hs[k],
                          Integration.timescale_factor) for k in range(D))
    seg = S.const_segments(D, xx, 0.0, T, dt, nus, migs, gammas, hs, theta0)
    fz, inj = S.masks(D, frozen, [False] * D)
    order = _lib.ORDER_PRE1 if D == 1 else (_lib.ORDER_PRE if D <= 3 else _lib.ORDER_TV)
    return Epoch(D, seg, order, fz, inj)


def prog_split_mig(params, ns, pts):
    """Demographics2D.split_mig (dadi/Demographics2D.py:133-167)."""
    nu1, nu2, T, m = params
    xx = Numerics.default_grid(pts)
    phi0 = PhiManip.phi_1D(xx)
    ops = [("split", lambda phi: PhiManip.phi_1D_to_2D_device(xx, phi)),
           ("epoch", _const_epoch(2, xx, T, [nu1, nu2], [[m], [m]]))]
    return Program(pts, xx, phi0, ops, tuple(ns), "split_mig")


def prog_three_epoch(params, ns, pts):
    """Demographics1D.three_epoch (dadi/Demographics1D.py:128-161)."""
    nuB, nuF, TB, TF = params
    xx = Numerics.default_grid(pts)
    phi0 = PhiManip.phi_1D(xx)
    ops = [("epoch", _const_epoch(1, xx, TB, [nuB], [[]])),
           ("epoch", _const_epoch(1, xx, TF, [nuF], [[]]))]
    return Program(pts, xx, phi0, ops, tuple(ns), "three_epoch")


def prog_three_pop_chain(params, ns, pts):
    """Demographics3D.three_pop_chain: 1D epoch -> split -> 2D epoch -> split ->
    3D epoch with symmetric migration."""
    nuA, TA, nu1, nu2, T2, m2, nu1b, nu2b, nu3, T3, m3 = params
    xx = Numerics.default_grid(pts)
    phi0 = PhiManip.phi_1D(xx)
    ops = [("epoch", _const_epoch(1, xx, TA, [nuA], [[]])),
           ("split", lambda phi: PhiManip.phi_1D_to_2D_device(xx, phi)),
           ("epoch", _const_epoch(2, xx, T2, [nu1, nu2], [[m2], [m2]])),
           ("split", lambda phi: PhiManip.phi_2D_to_3D_split_2_device(xx, phi)),
           ("epoch", _const_epoch(3, xx, T3, [nu1b, nu2b, nu3], [[m3, m3]] * 3))]
    return Program(pts, xx, phi0, ops, tuple(ns), "three_pop_chain")


def prog_four_pop_chain(params, ns, pts):
    """Demographics3D.four_pop_chain: three_pop_chain, a split of population 3
    into 3 and 4, and a 4-population epoch."""
    prog = prog_three_pop_chain(params[:11], ns, pts)
    nu1c, nu2c, nu3c, nu4, T4, m4 = params[11:]
    xx = prog.xx
    prog.ops = prog.ops + [
        ("split", lambda phi: PhiManip.phi_3D_to_4D_device(phi, 0, 0, xx, xx, xx, xx)),
        ("epoch", _const_epoch(4, xx, T4, [nu1c, nu2c, nu3c, nu4], [[m4] * 3] * 4))]
    prog.kind = "four_pop_chain"
    return prog


# ---------------------------------------------------------------------------
# server
# ---------------------------------------------------------------------------
@dataclass
class Request:
    rid: int
    prog: Program
    arrival: float = 0.0
    state: int = 0
    phi: object = None
    rec: np.ndarray = None
    segd: object = None
    result: np.ndarray = None
    error: str = None
    t_admit: float = None
    t_done: float = None
    footprint: int = 0
    wide: object = None
    seg: object = None


class Server:
    """Continuous batching of model-evaluation requests.

    budget_bytes: HBM the in-flight requests may occupy (phi, sampling work
    and schedules); requests that do not fit wait in arrival order.
    quantum: time steps each in-flight integration advances per launch.
    wide_streams: streams the multi-workgroup (3-5 population) epochs are
    spread over, by request; measured on MI355X with the seeded 200-request
    mix, p99 0.43-0.56 s (median of five runs 0.44 s) with 4 against
    0.47-0.52 s with 2 and 0.54-0.58 s with 6 or 8.
    """

    def __init__(self, budget_bytes=8 << 30, quantum=200, ctx=None, quantum_ms=8.0,
                 wide_lines=512, wide_streams=4, steps_per_graph=16, wide_waves=2,
                 factored=True, wide_mode="graph", resident=True):
        import torch
        self.torch = torch
        self.ctx = ctx or engine.context()
        self.budget = int(budget_bytes)
        self.quantum = int(quantum)
        self.quantum_ms = float(quantum_ms)
        self.wide_lines = int(wide_lines)
        self.wide_streams = max(1, int(wide_streams))
        self.spg = max(1, int(steps_per_graph))
        self.wide_waves = int(wide_waves)
        self.factored = bool(factored)
        self.resident = bool(resident)
        if wide_mode not in ("spread", "graph"):
            raise ValueError("wide_mode is 'spread' or 'graph'")
        self.wide_mode = wide_mode
        self.steps = {}       # class -> time steps per launch (adapted)
        self.streams = {}
        self.scratch = {}
        self.used = 0
        self.launches = 0
        self.wide_launches = 0

    def _steps_for(self, key):
        q = self.steps.get(key)
        if q is None:
            q = self.quantum if key[0] not in ("wide", "spread") else 4 * self.spg
        return q

    def _adapt(self, key, q, ms):
        """Steer the class's steps per launch toward quantum_ms per launch."""
        if ms <= 0:
            return
        want = q * self.quantum_ms / ms
        want = min(max(want, 0.5 * q), 2.0 * q)
        if key[0] in ("wide", "spread"):
            want = max(self.spg, int(round(want / self.spg)) * self.spg)
        else:
            want = max(1, min(int(round(want)), 50 * self.quantum))
        self.steps[key] = want

    def is_wide(self, D, n):
        """Epochs of D >= 3 populations whose sweeps have at least wide_lines
        lines take the multi-workgroup path."""
        return D >= 3 and n ** (D - 1) >= self.wide_lines

    @staticmethod
    def footprint(prog):
        """Bytes a request holds while in flight: phi at its largest
        dimensionality plus the split/sampling temporaries (about two more
        copies) and its schedules."""
        seg_bytes = sum(arg.seg.nbytes for op, arg in prog.ops if op == "epoch")
        return 3 * 8 * prog.max_points() + seg_bytes + 4096

    def _stream(self, key):
        st = self.streams.get(key)
        if st is None:
            st = self.torch.cuda.Stream(device=self.ctx.device)
            self.streams[key] = st
        return st

    # -- per-request state machine ------------------------------------------
    def _advance(self, r):
        """Run host/device operations until the request reaches an epoch or
        completes."""
        t = self.torch
        while r.state < len(r.prog.ops):
            op, arg = r.prog.ops[r.state]
            if op == "split":
                r.phi = arg(r.phi).contiguous()
                r.state += 1
                continue
            ep = arg
            if r.wide is not None:
                if r.wide["cur"] >= r.wide["total"]:
                    r.wide = None
                    r.state += 1
                    continue
                return "wide"
            if r.rec is None:
                if len(ep.seg) == 0 or ep.frozen == (1 << ep.D) - 1:
                    r.state += 1
                    continue
                if r.phi.dim() != ep.D:
                    raise ValueError("epoch dimensionality does not match phi")
                if self.is_wide(ep.D, r.prog.pts):
                    r.wide = self._wide_state(r, ep)
                    return "wide"
                r.segd = self.ctx.to_dev(ep.seg.view(np.uint8))
                xxd = self.ctx.grid(r.prog.xx)
                r.rec = engine._inst_record(r.phi.data_ptr(), xxd.data_ptr(), r.segd.data_ptr(),
                                            len(ep.seg), ep.D, r.prog.pts, ep.order, ep.frozen,
                                            ep.inject, engine.pitch_of(r.phi))
                r.seg = ep.seg
                return "epoch"
            if r.rec["status"][0] == 1:
                r.rec = None
                r.segd = None
                r.state += 1
                continue
            return "epoch"
        from .Spectrum_mod import sample_device
        r.result = sample_device(r.phi, r.prog.ns, r.prog.xx, self.ctx).cpu().numpy()
        return "done"

    def _wide_state(self, r, ep):
        """Device state of a multi-workgroup epoch: the schedule, the step of
        each time step's segment and the step counter of the captured graph."""
        t = self.torch
        ctx = self.ctx
        segd = ctx.to_dev(ep.seg.view(np.uint8))
        total = int(np.sum(ep.seg["nsteps"]))
        sos = np.repeat(np.arange(len(ep.seg), dtype=np.int32), ep.seg["nsteps"])
        sosd = ctx.to_dev(sos)
        base = t.zeros(1, dtype=t.int64, device=ctx.device)
        xxd = ctx.grid(r.prog.xx)
        inst = _lib.DhInstance(r.phi.data_ptr(), xxd.data_ptr(), segd.data_ptr(), len(ep.seg), 0,
                               0, ep.D, r.prog.pts, ep.order, ep.frozen, ep.inject, 0,
                               engine.pitch_of(r.phi))
        return {"inst": inst, "segd": segd, "sosd": sosd, "base": base, "total": total,
                "cur": 0, "keep": xxd}

    def _launch_wide(self, r):
        """One quantum of one multi-workgroup epoch (steps from w['cur'])."""
        t = self.torch
        ctx = self.ctx
        w = r.wide
        key = ("wide", r.prog.pts ** w["inst"].D)
        q = self._steps_for(key)
        st = self._stream(("wide", r.rid % self.wide_streams))
        st.wait_stream(t.cuda.current_stream(ctx.device))
        ev0 = t.cuda.Event(enable_timing=True)
        ev0.record(st)
        _lib.check(ctx.lib.dh_integrate_graph_range(w["inst"], w["base"].data_ptr(),
                                                    w["sosd"].data_ptr(), w["total"], w["cur"],
                                                    q, self.spg, self.wide_waves,
                                                    st.cuda_stream),
                   "dh_integrate_graph_range")
        w["pending"] = min(q, w["total"] - w["cur"])
        self.launches += 1
        self.wide_launches += 1
        return st, key, q, ev0

    SPREAD_CAP = 128      # multi-workgroup epochs per spread launch
    SPREAD_WGS = 2048     # workgroups per spread sweep launch
    SPREAD_WAVES = 4

    def _launch_spread(self, key, group):
        """One quantum of every in-flight multi-workgroup epoch of one
        (D, order) class in shared launches (dh_integrate_batch): each sweep
        launch covers all of them, each epoch at its own step (step_shift).
        The item and workgroup tables live in fixed buffers and the launch
        shape is fixed, so one captured graph serves every quantum."""
        t = self.torch
        ctx = self.ctx
        _, D, order = key
        for r in group:                # beyond the capacity: no steps this quantum
            r.wide["pending"] = 0
        group = group[:self.SPREAD_CAP]
        cap, nwg, waves = self.SPREAD_CAP, self.SPREAD_WGS, self.SPREAD_WAVES
        bufs = self.scratch.setdefault(key, {})
        if "items" not in bufs:
            bufs["items"] = t.zeros((cap + 1) * _lib.ITEM_DTYPE.itemsize, dtype=t.uint8,
                                    device=ctx.device)
            bufs["wg"] = t.zeros(nwg, dtype=t.int32, device=ctx.device)
            bufs["base"] = t.zeros(1, dtype=t.int64, device=ctx.device)
        q = self._steps_for(("spread", D, order))
        items = np.zeros(cap + 1, dtype=_lib.ITEM_DTYPE)   # the last one is empty
        tiles = [-(-(r.prog.pts ** (D - 1)) // 64) for r in group]
        want = [max(1, -(-nt // waves)) for nt in tiles]
        scale = min(1.0, nwg / max(sum(want), 1))
        wg_item = np.full(nwg, cap, dtype=np.int32)
        wg = 0
        longest = 0
        for i, (r, w0) in enumerate(zip(group, want)):
            w = r.wide
            items[i]["inst"] = np.frombuffer(bytes(w["inst"]), dtype=_lib.INST_DTYPE)[0]
            items[i]["seg_of_step"] = w["sosd"].data_ptr()
            items[i]["total_steps"] = w["total"]
            items[i]["first_axis"] = next(k for k in range(D) if not (w["inst"].frozen >> k) & 1)
            items[i]["step_shift"] = w["cur"]
            nw_i = max(1, int(w0 * scale))
            items[i]["wg_begin"] = wg
            items[i]["wgs"] = nw_i
            wg_item[wg:wg + nw_i] = i
            wg += nw_i
            w["pending"] = min(q, w["total"] - w["cur"])
            longest = max(longest, w["total"] - w["cur"])
        st = self._stream(key)
        st.wait_stream(t.cuda.current_stream(ctx.device))
        nmax = -(-max(r.prog.pts for r in group) // 16) * 16
        ev0 = t.cuda.Event(enable_timing=True)
        with t.cuda.stream(st):
            bufs["items"].copy_(t.from_numpy(items.view(np.uint8)))
            bufs["wg"].copy_(t.from_numpy(wg_item))
            ev0.record(st)
            _lib.check(ctx.lib.dh_integrate_batch(
                bufs["items"].data_ptr(), bufs["wg"].data_ptr(), cap + 1, nwg, int(D), int(order),
                nmax, waves, bufs["base"].data_ptr(), longest, 0, q, self.spg, st.cuda_stream),
                "dh_integrate_batch")
        self.launches += 1
        self.wide_launches += 1
        return st, ("spread", D, order), q, ev0

    def _class_of(self, r):
        """The launch class of a request's current epoch: (populations,
        order), and for 2 populations whose grid has a register-resident
        kernel also that kernel's row count (one launch serves one such
        kernel)."""
        D, order = int(r.rec["D"][0]), int(r.rec["order"][0])
        if D == 2 and self.resident and order in (_lib.ORDER_TV, _lib.ORDER_PRE):
            rows = self.ctx.lib.dh_resident_latency_rows(int(r.rec["n"][0]))
            if rows:
                return (D, order, rows)
        return (D, order)

    def _launch(self, group):
        """One quantum for every in-flight epoch of one class (_class_of)."""
        t = self.torch
        ctx = self.ctx
        D, order = group[0].rec["D"][0], group[0].rec["order"][0]
        tab = np.concatenate([r.rec for r in group])
        key = self._class_of(group[0])
        q = self._steps_for(key)
        st = self._stream(key)
        # splits and uploads of this quantum were issued on the current stream
        st.wait_stream(t.cuda.current_stream(ctx.device))
        ev0 = t.cuda.Event(enable_timing=True)
        with t.cuda.stream(st):
            tabd = t.from_numpy(tab.view(np.uint8).copy()).to(ctx.device, non_blocking=False)
            ev0.record(st)
            counter = t.zeros(1, dtype=t.int32, device=ctx.device)
            nmax = int(tab["n"].max())
            h = st.cuda_stream
            budget_us = int(self.quantum_ms * 1000)
            if D == 1 and ctx.one_d_wave(len(group), nmax):
                cfg = _lib.DhLaunch(1, 0, 0, 0, 0)   # a time budget, not a step count
                cfg.max_us = budget_us
                _lib.check(ctx.lib.dh_integrate_1d_wave(tabd.data_ptr(), len(group), int(order),
                                                        nmax, counter.data_ptr(), cfg, h),
                           "dh_integrate_1d_wave")
            elif D == 1:
                thr = 64
                g = max(1, math.ceil(len(group) / thr))
                need = 6 * nmax * g * thr
                ws = self.scratch.setdefault(key, {})
                scr = ws.get("lanes")
                if scr is None or scr.numel() < need:
                    scr = t.empty(need, dtype=t.float64, device=ctx.device)
                    ws["lanes"] = scr
                cfg = _lib.DhLaunch(1, g, q, 0, 0)
                _lib.check(ctx.lib.dh_integrate_1d(tabd.data_ptr(), len(group), int(order), nmax,
                                                   counter.data_ptr(), scr.data_ptr(), scr.numel(),
                                                   cfg, h), "dh_integrate_1d")
            elif len(key) == 3:
                # a time budget, not a step count; the kernel's latency form
                # while the class holds few integrations
                ws = self.scratch.setdefault(key, {})
                ctx.resident_launch(tabd.data_ptr(), len(group), int(order), nmax,
                                    counter.data_ptr(), h, ws, max_us=budget_us)
            elif self.factored and ctx.factorable([r.seg for r in group]):
                ws = self.scratch.setdefault(key, {})
                ctx.factored_launch(tabd.data_ptr(), len(group), int(D), int(order), nmax,
                                    counter.data_ptr(), h, ws, max_us=budget_us)
            else:
                ws = self.scratch.setdefault(key, {})
                ctx.nd_launch(tabd.data_ptr(), len(group), int(D), int(order), nmax,
                              counter.data_ptr(), h, ws, max_us=budget_us)
        self.launches += 1
        return st, tabd, counter, key, q, ev0

    def run(self, requests, realtime=True, clock=time.perf_counter):
        """Serve ``requests`` (arrival times in seconds from the start).
        realtime=False admits every request at time 0.  Returns a stats dict;
        results are left on the requests.

        Each (populations, order) class runs its own sequence of quanta on its
        own stream: as soon as a class's launch finishes, its requests advance
        and the class is relaunched with whatever is in flight then, including
        requests admitted meanwhile -- a slow class never holds back a fast
        one."""
        t = self.torch
        pending = sorted(requests, key=lambda r: r.arrival)
        for r in pending:
            r.footprint = self.footprint(r.prog)
        waiting = {}      # class -> requests ready for their next quantum
        inflight = {}     # class -> (requests, stream, table, counter, event)
        done = []
        t0 = clock()

        def enqueue(r):
            try:
                status = self._advance(r)
            except Exception as e:  # a bad request fails alone
                r.error = repr(e)
                status = "done"
            if status == "done":
                self._retire(r, done, t0, clock)
            elif status == "wide" and self.wide_mode == "spread":
                inst = r.wide["inst"]
                waiting.setdefault(("spread", inst.D, inst.order), []).append(r)
            elif status == "wide":
                waiting.setdefault(("wide", r.rid), []).append(r)
            else:
                waiting.setdefault(self._class_of(r), []).append(r)

        while pending or waiting or inflight:
            now = clock() - t0
            while pending and (not realtime or pending[0].arrival <= now):
                r = pending[0]
                if r.footprint > self.budget:
                    pending.pop(0)
                    r.error = "request needs %d bytes, over the %d-byte budget" % (
                        r.footprint, self.budget)
                    r.t_done = clock() - t0
                    done.append(r)
                    continue
                if self.used + r.footprint > self.budget:
                    break
                pending.pop(0)
                self.used += r.footprint
                r.t_admit = now
                try:
                    r.phi = self.ctx.to_dev(np.ascontiguousarray(r.prog.phi0, dtype=np.float64))
                except Exception as e:
                    r.error = repr(e)
                    self._retire(r, done, t0, clock)
                    continue
                enqueue(r)
            for key in sorted(waiting, key=repr):
                if key in inflight or not waiting[key]:
                    continue
                grp = waiting.pop(key)
                try:
                    if key[0] == "wide":
                        st, ckey, q, ev0 = self._launch_wide(grp[0])
                        tabd = None
                    elif key[0] == "spread":
                        st, ckey, q, ev0 = self._launch_spread(key, grp)
                        tabd = None
                    else:
                        st, tabd, counter, ckey, q, ev0 = self._launch(grp)
                except Exception as e:  # the class's requests fail, the server goes on
                    for r in grp:
                        r.error = repr(e)
                        self._retire(r, done, t0, clock)
                    continue
                ev = t.cuda.Event(enable_timing=True)
                ev.record(st)
                inflight[key] = (grp, tabd, ev, ev0, ckey, q)
            progressed = False
            for key in list(inflight):
                grp, tabd, ev, ev0, ckey, q = inflight[key]
                if not ev.query():
                    continue
                progressed = True
                del inflight[key]
                ms = ev0.elapsed_time(ev)
                if tabd is None:
                    if any(r.wide["pending"] == q for r in grp):
                        self._adapt(ckey, q, ms)
                    for r in grp:
                        r.wide["cur"] += r.wide["pending"]
                        enqueue(r)
                    continue
                back = tabd.cpu().numpy().view(INST_DTYPE)
                for r, rec in zip(grp, back):
                    r.rec["cur_seg"] = rec["cur_seg"]
                    r.rec["cur_step"] = rec["cur_step"]
                    r.rec["status"] = rec["status"]
                if np.any(back["status"] != 1):   # someone ran the whole quantum
                    self._adapt(ckey, q, ms)
                for r in grp:
                    enqueue(r)
            waiting = {k: v for k, v in waiting.items() if v}
            if not progressed:
                if inflight:
                    time.sleep(0.0002)
                elif pending and realtime and not waiting:
                    time.sleep(max(0.0, min(pending[0].arrival - (clock() - t0), 0.002)))
        wall = clock() - t0
        lat = np.array([r.t_done - r.arrival for r in done if r.error is None])
        return {
            "requests": len(done),
            "failed": sum(1 for r in done if r.error is not None),
            "wall_s": wall,
            "throughput_req_s": len(lat) / wall if wall > 0 else 0.0,
            "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3) if len(lat) else None,
                           "p99": float(np.percentile(lat, 99) * 1e3) if len(lat) else None},
            "launches": self.launches,
            "wide_launches": self.wide_launches,
        }

    def _retire(self, r, done, t0, clock):
        r.t_done = clock() - t0
        r.phi = None
        r.rec = None
        r.segd = None
        r.wide = None
        self.used -= r.footprint
        done.append(r)


def poisson_arrivals(rng, n, rate):
    """n arrival times of a Poisson process of the given rate (per second)."""
    return np.cumsum(rng.exponential(1.0 / rate, n))


def mixed_requests(rng, n, rate, scale=1.0):
    """The serving mix: 55% 2D split_mig (pts U{40..120}), 30% 1D three_epoch
    (pts U{100..600}), 12% 3D chain (pts U{30..60}), 3% 4D chain
    (pts U{16..24}); sample sizes U{10..40} per population; sizes
    logU[0.1,10], migration U[0,5], times U[0.01,2] (2D) or U[0.01,1] (chains).
    ``scale`` multiplies every epoch time (a shorter mix for tests)."""
    arr = poisson_arrivals(rng, n, rate)
    out = []
    lu = lambda: float(np.exp(rng.uniform(np.log(0.1), np.log(10))))
    for i in range(n):
        u = rng.uniform()
        if u < 0.55:
            pts = int(rng.integers(40, 121))
            ns = (int(rng.integers(10, 41)), int(rng.integers(10, 41)))
            prog = prog_split_mig((lu(), lu(), scale * rng.uniform(0.01, 2), rng.uniform(0, 5)),
                                  ns, pts)
        elif u < 0.85:
            pts = int(rng.integers(100, 601))
            ns = (int(rng.integers(10, 41)),)
            prog = prog_three_epoch((lu(), lu(), scale * rng.uniform(0.01, 2),
                                     scale * rng.uniform(0.01, 2)), ns, pts)
        elif u < 0.97:
            pts = int(rng.integers(30, 61))
            ns = tuple(int(rng.integers(10, 41)) for _ in range(3))
            p = (lu(), scale * rng.uniform(0.01, 1), lu(), lu(), scale * rng.uniform(0.01, 1),
                 rng.uniform(0, 5), lu(), lu(), lu(), scale * rng.uniform(0.01, 1),
                 rng.uniform(0, 5))
            prog = prog_three_pop_chain(p, ns, pts)
        else:
            pts = int(rng.integers(16, 25))
            ns = tuple(int(rng.integers(10, 41)) for _ in range(4))
            p = (lu(), scale * rng.uniform(0.01, 1), lu(), lu(), scale * rng.uniform(0.01, 1),
                 rng.uniform(0, 5), lu(), lu(), lu(), scale * rng.uniform(0.01, 1),
                 rng.uniform(0, 5), lu(), lu(), lu(), lu(), scale * rng.uniform(0.01, 1),
                 rng.uniform(0, 5))
            prog = prog_four_pop_chain(p, ns, pts)
        out.append(Request(i, prog, float(arr[i])))
    return out
