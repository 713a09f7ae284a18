"""
Device side of the integration: packs schedules and instance descriptors into
device buffers (torch tensors), picks a launch shape and calls the C ABI.

Three execution modes, all computing bit-identical results:

* ``run_single``  one integration per call (the "plain" path): a single
  persistent workgroup when the swept axis has few lines, otherwise one
  multi-workgroup launch per axis per step replayed from a hipGraph.
* ``run_batch``   many independent integrations in one persistent launch per
  (D, order) class, workgroups (D >= 2) or lanes (D == 1) pulling instances
  from a device-side queue.
* quanta (``max_steps``) let a caller advance a resident batch a few steps at
  a time; the serving loop in :mod:`dadi_amd.serve` builds on that.
"""
import math
import threading

import numpy as np

from . import _lib
from ._lib import ITEM_DTYPE, INST_DTYPE, SEG_DTYPE, DhInstance, DhLaunch, check

_ctx_lock = threading.Lock()
_ctx = {}
_local = threading.local()


def torch():
    import torch as _t
    return _t


class Context:
    """Per-device state: library handle, device, grid cache, tuning table."""

    def __init__(self, device=0):
        t = torch()
        self.lib = _lib.init(device)
        self.device = t.device("cuda", device)
        props = t.cuda.get_device_properties(self.device)
        self.num_cu = props.multi_processor_count
        self.max_lds = int(self.lib.dh_max_lds())
        self._xx_cache = {}
        from .autotune import Tuning
        self.tuning = Tuning.load()

    #: 1D batches: "auto" (wave per integration while the batch is small
    #: enough for the GPU to hold one wave per integration), "wave" or "lane"
    one_d_kernel = "auto"
    one_d_wave_max = 16384
    #: 2D batches of at most this many integrations run on the factored
    #: kernel when their segments are long; 0 disables it.  Measured on
    #: MI355X (split_mig draws, n = 40..100): 1.2-3.1x the packed kernel's
    #: integrations/s for 512-2048 integrations, 0.68x at 8192 (n = 100)
    factored_max_inst = 2048
    #: total factorisation workspace up to which a large factored batch runs
    #: two workgroups per compute unit (more stays out of the last-level cache)
    factored_ws_bytes = 160 << 20
    #: batches of 3-5 populations of at most this many integrations spread
    #: each integration's sweeps over several workgroups; 0 disables it
    spread_max_inst = 512
    spread_waves = 0          # 0: chosen per batch
    #: spread batches run in phases that drop finished integrations
    spread_phases = True
    #: 2D batches whose grid size has a register-resident kernel use it when
    #: no other launch shape is asked for; False keeps them on the others
    resident_auto = True
    #: register-resident batches take the kernel's latency form
    #: (dh_integrate_resident_latency: the Thomas factorisation kept per
    #: segment, a step runs only the right-hand-side recurrences) up to this
    #: many integrations per workgroup the latency form keeps resident, the
    #: throughput form above.  Measured on MI355X with split_mig draws
    #: (latency / throughput form, integrations/s): n = 100, 32..512
    #: integrations 1.34..1.08x, 1024 0.85x; n = 60, 32..1024 1.35..1.06x.
    resident_latency_per_group = 2

    def resident_latency_max(self, nmax):
        """Largest batch on the latency form: resident_latency_per_group
        integrations per workgroup it keeps resident (one per compute unit
        when its n x (n+1) LDS tile exceeds half the LDS, else two)."""
        return self.resident_latency_per_group * self.resident_latency_groups(nmax)

    def resident_latency_groups(self, nmax):
        """Workgroups of the latency form resident at once: one or two per
        compute unit (its n x (n+1) LDS tile), 0 without a kernel."""
        rows = int(self.lib.dh_resident_latency_rows(int(nmax)))
        if rows == 0:
            return 0
        return max(1, min(2, self.max_lds // (8 * rows * (rows + 1)))) * self.num_cu

    def one_d_wave(self, ninst, nmax):
        """Whether a 1D batch runs one wave per integration (LDS-resident,
        latency-bound steps) rather than one lane per integration."""
        if self.lib.dh_1d_wave_lds_bytes(int(nmax)) > self.lib.dh_max_lds():
            return False
        if self.one_d_kernel != "auto":
            return self.one_d_kernel == "wave"
        return ninst <= self.one_d_wave_max

    def factored_launch(self, tab_ptr, ninst, D, order, nmax, counter_ptr, stream, workspace,
                        waves=None, groups=0, max_steps=0, max_us=0):
        """One dh_integrate_factored call (latency path for D >= 2: the Thomas
        factorisation of every line kept per segment).  ``workspace`` (a
        dict owned by the caller, one per stream) keeps the factorisation
        workspace between calls."""
        tiles = -(-(nmax ** (D - 1)) // 64)
        w = waves or max(1, min(8, tiles))
        slice_d = int(self.lib.dh_factored_slice_doubles(D, nmax))
        g = groups if groups > 0 else self.factored_groups(ninst, slice_d)
        need = g * slice_d
        ws = workspace.get("factored")
        if ws is None or ws.numel() < need:
            ws = torch().empty(need, dtype=torch().float64, device=self.device)
            workspace["factored"] = ws
        cfg = DhLaunch(w, g, max_steps, 0, 0)
        cfg.gamma_dev = ws.data_ptr()
        cfg.gamma_doubles = ws.numel()
        cfg.max_us = int(max_us)
        check(self.lib.dh_integrate_factored(tab_ptr, ninst, D, order, nmax, counter_ptr, cfg,
                                             stream), "dh_integrate_factored")

    def resident_launch(self, tab_ptr, ninst, order, nmax, counter_ptr, stream, workspace,
                        groups=0, max_steps=0, max_us=0, latency=None):
        """One register-resident 2D launch: the latency form
        (dh_integrate_resident_latency, one or two workgroups per compute
        unit, its factorisation workspace kept in ``workspace``, a dict owned by the
        caller, one per stream) when ``latency`` is True or, for None, the
        batch has at most resident_latency_max(nmax) integrations; else
        dh_integrate_resident."""
        if latency is None:
            latency = ninst <= self.resident_latency_max(nmax)
        if not self.lib.dh_resident_rows(int(nmax)):
            latency = True    # above 104 points only the latency form exists
        if not latency:
            cfg = DhLaunch(1, groups, max_steps, 0, 0)
            cfg.max_us = int(max_us)
            check(self.lib.dh_integrate_resident(tab_ptr, ninst, order, nmax, counter_ptr, cfg,
                                                 stream), "dh_integrate_resident")
            return
        slice_d = int(self.lib.dh_resident_latency_slice_doubles(int(nmax)))
        g = groups if groups > 0 else min(ninst, self.resident_latency_groups(nmax))
        need = g * slice_d
        ws = workspace.get("resident")
        if ws is None or ws.numel() < need:
            ws = torch().empty(need, dtype=torch().float64, device=self.device)
            workspace["resident"] = ws
        cfg = DhLaunch(1, g, max_steps, 0, 0)
        cfg.gamma_dev = ws.data_ptr()
        cfg.gamma_doubles = ws.numel()
        cfg.max_us = int(max_us)
        check(self.lib.dh_integrate_resident_latency(tab_ptr, ninst, order, nmax, counter_ptr,
                                                     cfg, stream),
              "dh_integrate_resident_latency")

    def factored_groups(self, ninst, slice_doubles):
        """Workgroups of a factored launch: one per compute unit, two for
        batches of at least eight integrations per compute unit whose
        workspaces then stay within factored_ws_bytes (measured: n = 40 and
        60, 2048 integrations, 1.6x and 1.35x; at n = 100 two per compute
        unit are slower)."""
        if ninst >= 8 * self.num_cu and \
                2 * self.num_cu * slice_doubles * 8 <= self.factored_ws_bytes:
            return 2 * self.num_cu
        return min(ninst, self.num_cu)

    @staticmethod
    def factorable(segs, min_steps=4):
        """Whether schedules have segments long enough for the factorisation
        (formed once per segment) to pay: at least ``min_steps`` steps per
        segment on average."""
        nst = sum(int(np.sum(sg["nsteps"])) for sg in segs)
        nseg = sum(len(sg) for sg in segs)
        return nseg > 0 and nst >= min_steps * nseg

    # -- helpers ---------------------------------------------------------
    def stream(self):
        return torch().cuda.current_stream(self.device).cuda_stream

    def to_dev(self, arr, dtype=None):
        t = torch()
        a = np.ascontiguousarray(arr)
        return t.from_numpy(a).to(self.device, non_blocking=False)

    def grid(self, xx):
        key = (len(xx), xx.tobytes())
        g = self._xx_cache.get(key)
        if g is None:
            g = self.to_dev(np.asarray(xx, dtype=np.float64))
            if len(self._xx_cache) > 256:
                self._xx_cache.clear()
            self._xx_cache[key] = g
        return g

    def empty_phi(self, shape, batch=None):
        """Uninitialised device phi of ``shape`` (all axes of equal length),
        stored with its last axis padded to row_pitch(n) when D >= 2; with
        ``batch``, a (batch,) + shape arena of such phis."""
        t = torch()
        shape = tuple(int(v) for v in shape)
        n = shape[-1]
        ld = row_pitch(n) if len(shape) >= 2 else n
        full = ((batch,) if batch is not None else ()) + shape[:-1] + (ld,)
        return t.empty(full, dtype=t.float64, device=self.device)[..., :n]

    def side_streams(self, k):
        """k streams (default priority) for launches that run side by side."""
        t = torch()
        pool = self.__dict__.setdefault("_side", [])
        while len(pool) < k:
            pool.append(t.cuda.Stream(self.device))
        return pool[:k]

    def max_waves(self, D, n, chunk=0, recompute=0, gamma_global=0, delj=0):
        """Largest waves-per-workgroup whose LDS fits (mirrors dh_lds_bytes)."""
        L = self.lib
        fixed = L.dh_lds_bytes(D, n, 0, chunk, recompute, gamma_global, delj)
        per_wave = L.dh_lds_bytes(D, n, 1, chunk, recompute, gamma_global, delj) - fixed
        return max(0, min(8, (self.max_lds - fixed) // per_wave))

    #: resident waves per compute unit the nD kernels reach (register-limited:
    #: 3 per SIMD at their register budget)
    RESIDENT_WAVES = 12

    def _occupancy(self, lds, waves):
        per_cu = min(self.max_lds // max(lds, 1), 16 // waves + 1)
        return min(per_cu * waves, self.RESIDENT_WAVES)

    def pack_plan(self, D, n, waves=None, slots=None):
        """(waves, slots, score) of the packed kernel for instances of D
        populations on n points, maximising busy lanes x resident waves per
        compute unit; None when nothing fits."""
        L = n ** (D - 1)
        Lp = -(-L // 8) * 8
        best = None
        for W in ([waves] if waves else range(1, 9)):
            S = slots if slots else min(16, (64 * W) // Lp)
            if S < 1:
                continue
            lds = self.lib.dh_packed_lds_bytes(D, n, W, S)
            if lds > self.max_lds:
                continue
            tiles = -(-(S * Lp) // 64)
            rounds = -(-tiles // W)
            util = S * L / (64.0 * W * rounds)
            score = util * self._occupancy(lds, W)
            if best is None or score > best[2] + 1e-9:
                best = (W, S, score)
        return best

    def _plain_score(self, D, n, sh):
        L = n ** (D - 1)
        lds = self.lib.dh_lds_bytes(D, n, sh.waves, sh.chunk, sh.recompute, sh.gamma_global, 0)
        util = L / (64.0 * sh.waves * -(-L // (64 * sh.waves)))
        return util * self._occupancy(lds, sh.waves)

    def shape(self, D, n, waves=None, chunk=None, recompute=None, gamma_global=None,
              slots=None, packable=False, ninst=None):
        """Launch shape of the nD kernels for instances of D populations on n
        points: the autotuned choice when one is recorded, else a rule of
        thumb.  ``packable``: the instances share n and grid, so several may
        share a workgroup (slots >= 2, the packed kernel).  ``ninst``: number
        of instances in the launch; too few to fill the GPU at the tuned
        shape (tuned on full batches) and each instance gets as many waves as
        it has 64-line tiles, for the shortest time per step.
        Returns Shape(waves, chunk, recompute, gamma_global, slots)."""
        sh = self._shape(D, n, waves, chunk, recompute, gamma_global, slots, packable)
        if ninst is None or waves is not None or slots is not None or D < 2:
            return sh
        tiles = -(-(n ** (D - 1)) // 64)
        if -(-ninst // sh.slots) * sh.waves >= 4 * self.num_cu or \
                (sh.slots < 2 and sh.waves >= min(8, tiles)):
            return sh
        ck = 8 if sh.slots >= 2 else sh.chunk
        ru = 0 if sh.slots >= 2 else sh.recompute
        gg = 0 if sh.slots >= 2 else sh.gamma_global
        wmax = self.max_waves(D, n, ck, ru, gg)
        w = max(sh.waves if sh.slots < 2 else 1, min(8, tiles, wmax))
        return Shape(w, ck, ru, gg, 1)

    def _shape(self, D, n, waves, chunk, recompute, gamma_global, slots, packable):
        if slots is not None and slots >= 2:
            if not packable:
                raise ValueError("packing needs instances that share n and the grid")
            plan = self.pack_plan(D, n, waves, slots)
            if plan is None:
                raise ValueError("%d slots of n=%d do not fit a workgroup" % (slots, n))
            return Shape(plan[0], 8, 0, 0, plan[1])
        free = waves is None and chunk is None and recompute is None and gamma_global is None
        if None in (waves, chunk, recompute, gamma_global) and self.tuning is not None:
            hit = self.tuning.lookup(D, n, packable=packable and free and slots is None)
            if hit is not None and hit.slots >= 2:
                return self._shape(D, n, hit.waves, None, None, None, hit.slots, True)
            if hit is not None:
                waves = hit.waves if waves is None else waves
                chunk = hit.chunk if chunk is None else chunk
                recompute = hit.recompute if recompute is None else recompute
                gamma_global = hit.gamma_global if gamma_global is None else gamma_global
                slots = 1
        gamma_global = int(bool(gamma_global))
        if gamma_global:
            chunk, recompute = 0, 0
        if chunk is None:
            chunk = 8 if (D >= 2 and n >= 40) else 0
        if recompute is None:
            recompute = 0
        if chunk == 0:
            recompute = 0
        nsys = n ** (D - 1)
        wmax = self.max_waves(D, n, chunk, recompute, gamma_global)
        if wmax < 1 and chunk == 0 and not gamma_global:
            chunk, recompute = 8, 0
            wmax = self.max_waves(D, n, chunk, recompute)
        if wmax < 1:
            raise ValueError("n=%d too large for the workgroup kernel at D=%d" % (n, D))
        auto_waves = waves is None
        if waves is None:
            # large instances get up to 8 waves (lower latency per step);
            # small ones few, so more instances share a compute unit
            cap = 8 if nsys >= 1024 else 4
            waves = max(1, min(wmax, cap, math.ceil(nsys / 64)))
        sh = Shape(min(waves, wmax), chunk, recompute, gamma_global, 1)
        if packable and slots is None and auto_waves and D >= 2 and chunk == 8 and \
                not recompute and not gamma_global:
            plan = self.pack_plan(D, n)
            if plan is not None and plan[1] >= 2 and plan[2] > 1.05 * self._plain_score(D, n, sh):
                return Shape(plan[0], 8, 0, 0, plan[1])
        return sh

    def nd_launch(self, tab_ptr, ninst, D, order, nmax, counter_ptr, stream, workspace,
                  waves=None, groups=0, max_steps=0, chunk=None, recompute=None,
                  gamma_global=None, slots=None, packable=False, delj=False, max_us=0):
        """One dh_integrate_nd (or, with slots >= 2, dh_integrate_packed)
        call.  ``workspace`` is a dict owned by the caller (one per stream:
        launches that may overlap must not share it) that keeps the gamma
        workspace between calls.  ``packable``: every instance has n = nmax
        and the same grid.  ``max_us`` > 0: time budget per instance (the
        unpacked kernel only, so it forbids packing)."""
        if max_us > 0:
            packable = False
        if delj:   # the delta trick runs on the plain Thomas path only
            tiles = -(-(nmax ** (D - 1)) // 64)
            w = waves or max(1, min(8, tiles, self.max_waves(D, nmax, 0, 0, 0, 1)))
            if self.max_waves(D, nmax, 0, 0, 0, 1) < 1:
                raise NotImplementedError("the delta trick needs the full gamma column in LDS, "
                                          "which n=%d at D=%d exceeds" % (nmax, D))
            cfg = DhLaunch(w, groups, max_steps, 0, 0)
            cfg.delj = 1
            cfg.max_us = int(max_us)
            check(self.lib.dh_integrate_nd(tab_ptr, ninst, D, order, nmax, counter_ptr, cfg,
                                           stream), "dh_integrate_nd")
            return Shape(w, 0, 0, 0, 1)
        sh = self.shape(D, nmax, waves, chunk, recompute, gamma_global, slots, packable, ninst)
        cfg = DhLaunch(sh.waves, groups, max_steps, sh.chunk, sh.recompute)
        cfg.max_us = int(max_us)
        if sh.slots >= 2:
            check(self.lib.dh_integrate_packed(tab_ptr, ninst, D, order, nmax, sh.slots,
                                               counter_ptr, cfg, stream), "dh_integrate_packed")
            return sh
        if sh.gamma_global:
            g = self.lib.dh_nd_groups(D, nmax, sh.waves, sh.chunk, sh.recompute, 1, groups, ninst)
            check(min(g, 0), "dh_nd_groups")
            need = g * sh.waves * 64 * nmax
            ws = workspace.get("gamma")
            if ws is None or ws.numel() < need:
                ws = torch().empty(need, dtype=torch().float64, device=self.device)
                workspace["gamma"] = ws
            cfg.gamma_dev = ws.data_ptr()
            cfg.gamma_doubles = ws.numel()
        check(self.lib.dh_integrate_nd(tab_ptr, ninst, D, order, nmax, counter_ptr, cfg, stream),
              "dh_integrate_nd")
        return sh


class Shape(tuple):
    """Launch shape of the nD kernels: waves per workgroup, Thomas chunk,
    recompute flag, gamma columns in a global workspace, instances per
    workgroup (slots >= 2: the packed kernel)."""
    __slots__ = ()
    _fields = ("waves", "chunk", "recompute", "gamma_global", "slots")

    def __new__(cls, waves, chunk, recompute, gamma_global=0, slots=1):
        return tuple.__new__(cls, (int(waves), int(chunk), int(recompute), int(gamma_global),
                                   int(slots)))

    waves = property(lambda self: self[0])
    chunk = property(lambda self: self[1])
    recompute = property(lambda self: self[2])
    gamma_global = property(lambda self: self[3])
    slots = property(lambda self: self[4])


def context(device=0):
    """The process-wide Context of ``device``."""
    with _ctx_lock:
        ctx = _ctx.get(device)
        if ctx is None:
            ctx = Context(device)
            _ctx[device] = ctx
        return ctx


def _lines(D, n):
    return n ** (D - 1)


def pitch_of(phi):
    """Row pitch (in doubles) of a device phi stored in C order with at most
    its last axis padded; 0 when it is dense.  ValueError for other layouts."""
    D = phi.dim()
    st, sh = phi.stride(), tuple(phi.shape)
    if D <= 1:
        if D == 1 and st[0] != 1:
            raise ValueError("phi must be C-ordered")
        return 0
    ld = st[-2]
    ok = st[-1] == 1 and ld >= sh[-1]
    expect = ld
    for k in range(D - 2, -1, -1):
        ok = ok and st[k] == expect
        expect *= sh[k]
    if not ok:
        raise ValueError("phi must be C-ordered with at most its last axis padded")
    return 0 if ld == sh[-1] else int(ld)


def row_pitch(n):
    """Row pitch the package allocates for phi on n points: a multiple of 8
    doubles, so that every row of the last axis starts 64-byte aligned."""
    return -(-n // 8) * 8


def _inst_record(phi_ptr, xx_ptr, seg_ptr, nseg, D, n, order, frozen, inject, ld=0):
    rec = np.zeros(1, dtype=INST_DTYPE)
    rec["ld"] = ld
    rec["phi"] = phi_ptr
    rec["xx"] = xx_ptr
    rec["seg"] = seg_ptr
    rec["nseg"] = nseg
    rec["D"] = D
    rec["n"] = n
    rec["order"] = order
    rec["frozen"] = frozen
    rec["inject"] = inject
    return rec


def run_single(phi, xx, seg, order, frozen, inject, ctx=None, mode="auto", waves=None,
               steps_per_graph=32, delj=False):
    """Integrate one phi in place on the device.

    phi: torch float64 tensor on the device (C order, shape (n,)*D, last axis
    possibly padded), or a numpy array (copied in and out; the numpy result
    is returned).  delj: Chang and Cooper's delta (dadi's use_delj_trick).
    """
    t = torch()
    ctx = ctx or context()
    host = isinstance(phi, np.ndarray)
    co = getattr(_local, "coalescer", None)
    if host:
        src = np.ascontiguousarray(phi, dtype=np.float64)
        if src.shape[0] < 3:
            raise ValueError("grids need at least 3 points")
        if co is not None and len(seg) and frozen != (1 << src.ndim) - 1:
            # the launch uploads every waiting host phi at once
            job = Job(None, xx, seg, order, frozen, inject, delj=delj)
            job.host = src
            co.submit(job)
            return job.result
        phid = ctx.empty_phi(src.shape)
        phid.copy_(t.from_numpy(src))
    else:
        phid = phi
    if co is not None and len(seg) and frozen != (1 << phid.dim()) - 1:
        job = Job(phid, xx, seg, order, frozen, inject, delj=delj)
        co.submit(job)
        return phid.cpu().numpy() if host else phid
    if not (phid.is_cuda and phid.dtype == t.float64):
        raise ValueError("phi must be a float64 device tensor")
    ld = pitch_of(phid)
    D = phid.dim()
    n = int(phid.shape[0])
    if n < 3:
        raise ValueError("grids need at least 3 points")
    if len(seg) == 0 or frozen == (1 << D) - 1:
        return phid.cpu().numpy() if host else phid
    xxd = ctx.grid(np.asarray(xx, dtype=np.float64))
    segd = ctx.to_dev(seg.view(np.uint8))
    nsys = _lines(D, n)
    if delj:
        # the delta trick runs on the plain Thomas path of the workgroup kernel
        if ctx.max_waves(D, n, 0, 0, 0, 1) < 1:
            raise NotImplementedError("the delta trick needs the full gamma column in LDS, "
                                      "which n=%d at D=%d exceeds" % (n, D))
        mode = "wg"
    if mode == "auto":
        wmax = max(ctx.max_waves(D, n, 0), ctx.max_waves(D, n, 8))
        mode = "wg" if (D == 1 or nsys <= 64 * max(wmax, 1)) and wmax >= 1 else "graph"
        if mode == "wg" and D == 2 and not delj and ctx.factored_max_inst > 0 and \
                ctx.factorable([seg]):
            mode = "factored"
        if D == 2 and not delj and ctx.resident_auto and \
                order in (_lib.ORDER_TV, _lib.ORDER_PRE) and ctx.lib.dh_resident_latency_rows(n):
            mode = "resident"
    stream = ctx.stream()
    if mode == "wg":
        # one instance: the full gamma column (less VALU) when it fits
        dj = 1 if delj else 0
        ck = 0 if (dj or ctx.max_waves(D, n, 0) >= min(8, math.ceil(nsys / 64))) else 8
        w = waves or max(1, min(ctx.max_waves(D, n, ck, 0, 0, dj), math.ceil(nsys / 64)))
        rec = _inst_record(phid.data_ptr(), xxd.data_ptr(), segd.data_ptr(), len(seg), D, n,
                           order, frozen, inject, ld)
        tab = ctx.to_dev(rec.view(np.uint8))
        counter = t.zeros(1, dtype=t.int32, device=ctx.device)
        cfg = DhLaunch(w, 1, 0, ck, 0)   # one instance: full gamma column when it fits
        cfg.delj = dj
        check(ctx.lib.dh_integrate_nd(tab.data_ptr(), 1, D, order, n, counter.data_ptr(),
                                      cfg, stream), "dh_integrate_nd")
        keep = (tab, counter, segd)
    elif mode in ("resident", "resident_throughput"):
        # "resident": the latency form (per-segment factorisation)
        rec = _inst_record(phid.data_ptr(), xxd.data_ptr(), segd.data_ptr(), len(seg), D, n,
                           order, frozen, inject, ld)
        tab = ctx.to_dev(rec.view(np.uint8))
        counter = t.zeros(1, dtype=t.int32, device=ctx.device)
        ws = {}
        ctx.resident_launch(tab.data_ptr(), 1, order, n, counter.data_ptr(), stream, ws,
                            groups=1, latency=(mode == "resident"))
        keep = (tab, counter, segd, ws)
    elif mode == "factored":
        rec = _inst_record(phid.data_ptr(), xxd.data_ptr(), segd.data_ptr(), len(seg), D, n,
                           order, frozen, inject, ld)
        tab = ctx.to_dev(rec.view(np.uint8))
        counter = t.zeros(1, dtype=t.int32, device=ctx.device)
        ws = {}
        ctx.factored_launch(tab.data_ptr(), 1, D, order, n, counter.data_ptr(), stream, ws,
                            waves=waves)
        keep = (tab, counter, segd, ws)
    elif mode == "graph":
        nst = int(np.sum(seg["nsteps"]))
        seg_of_step = np.repeat(np.arange(len(seg), dtype=np.int32), seg["nsteps"])
        sosd = ctx.to_dev(seg_of_step)
        base = t.zeros(1, dtype=t.int64, device=ctx.device)
        inst = DhInstance(phid.data_ptr(), xxd.data_ptr(), segd.data_ptr(), len(seg), 0, 0,
                          D, n, order, frozen, inject, 0, ld)
        # one wave per workgroup: 3-6% shorter steps than two for a single 3D/4D
        # integration alone (3D n = 40: 40.0 against 42.5 us, 60: 54.6 / 58.1;
        # 4D n = 24: 43.4 / 45.2)
        w = waves or 1
        check(ctx.lib.dh_integrate_graph(inst, base.data_ptr(), sosd.data_ptr(), nst,
                                         steps_per_graph, w, stream), "dh_integrate_graph")
        keep = (sosd, base, segd)
    else:
        raise ValueError(mode)
    if host:
        out = phid.cpu().numpy()
        del keep
        return out
    t.cuda.current_stream(ctx.device).synchronize()
    del keep
    return phid


class Job:
    """One integration of a batch: device phi (updated in place) plus schedule.
    A job submitted to a Coalescer may instead carry a host array (``host``,
    phi None): the launch uploads all such arrays at once and hands back
    ``result``, the integrated host array."""
    __slots__ = ("phi", "xx", "seg", "order", "frozen", "inject", "tag", "done", "delj",
                 "error", "host", "result")

    def __init__(self, phi, xx, seg, order, frozen, inject, tag=None, delj=False):
        self.done = False
        self.error = None
        self.host = None
        self.result = None
        self.delj = bool(delj)
        self.phi = phi
        self.xx = np.asarray(xx, dtype=np.float64)
        self.seg = seg
        self.order = order
        self.frozen = frozen
        self.inject = inject
        self.tag = tag

    @property
    def D(self):
        return (self.phi if self.phi is not None else self.host).ndim

    @property
    def n(self):
        return int((self.phi if self.phi is not None else self.host).shape[0])

    def work(self):
        act = sum(1 for k in range(self.D) if not (self.frozen >> k) & 1)
        return int(np.sum(self.seg["nsteps"])) * (self.n ** self.D) * max(act, 1)


class PackedBatch:
    """Device-resident instance table of one (D, order) class."""

    def __init__(self, ctx, jobs):
        t = torch()
        self.ctx = ctx
        self.jobs = jobs
        self.D = jobs[0].D
        self.order = jobs[0].order
        segs = [j.seg for j in jobs]
        offs = np.cumsum([0] + [len(s) for s in segs])
        allseg = np.concatenate(segs) if segs else np.zeros(0, dtype=SEG_DTYPE)
        self.segd = ctx.to_dev(allseg.view(np.uint8))
        base = self.segd.data_ptr()
        tab = np.zeros(len(jobs), dtype=INST_DTYPE)
        self.xxs = []
        for i, j in enumerate(jobs):
            xxd = ctx.grid(j.xx)
            self.xxs.append(xxd)
            tab[i] = _inst_record(j.phi.data_ptr(), xxd.data_ptr(),
                                  base + int(offs[i]) * SEG_DTYPE.itemsize, len(j.seg), j.D, j.n,
                                  j.order, j.frozen, j.inject, pitch_of(j.phi))[0]
        self.host_tab = tab
        self.tab0 = ctx.to_dev(tab.view(np.uint8))
        self.tab = self.tab0.clone()
        self.counter = t.zeros(1, dtype=t.int32, device=ctx.device)
        self.nmax = max(j.n for j in jobs)
        self.workspace = {}
        self.delj = bool(jobs[0].delj)
        if any(j.delj != self.delj for j in jobs):
            raise ValueError("a batch runs with or without the delta trick, not both")
        # several instances may share a workgroup when they share n and grid
        self.packable = self.D >= 2 and len({int(r) for r in tab["xx"]}) == 1 and \
            len({int(r) for r in tab["ld"]}) == 1 and all(j.n == self.nmax for j in jobs) and \
            not self.delj

    def restart(self):
        """Reset every instance's progress to the start of its schedule
        (device copy; the caller resets phi)."""
        self.tab.copy_(self.tab0)

    def resident_ok(self):
        """Whether the register-resident 2D kernels can run this batch: two
        populations, grid sizes within one kernel's range (33..104 points for
        dh_integrate_resident, up to 128 for its latency form), the TV or PRE
        coefficient order, no delta trick."""
        if self.D != 2 or self.delj or self.order not in (_lib.ORDER_TV, _lib.ORDER_PRE):
            return False
        rows = self.ctx.lib.dh_resident_latency_rows(self.nmax)
        return rows > 0 and all(self.ctx.lib.dh_resident_latency_rows(j.n) == rows
                                for j in self.jobs)

    def launch_resident(self, groups=0, max_steps=0, max_us=0, stream=None, latency=None):
        """One register-resident launch over the table (Context.resident_launch:
        latency True / False picks the form, None by batch size)."""
        ctx = self.ctx
        stream = ctx.stream() if stream is None else stream
        ctx.resident_launch(self.tab.data_ptr(), len(self.jobs), self.order, self.nmax,
                            self.counter.data_ptr(), stream, self.workspace, groups, max_steps,
                            max_us, latency)

    def launch(self, waves=None, groups=0, max_steps=0, stream=None, chunk=None,
               recompute=None, gamma_global=None, slots=None, resident=None, latency=None):
        """One persistent launch over the table (max_steps > 0: at most that
        many steps per instance).  Shape arguments left None are autotuned /
        chosen by rule of thumb; slots=1 forbids packing.  resident: True runs
        the register-resident 2D kernel, False never does; None uses it when
        it can run the batch and no other launch shape is asked for; latency
        picks its form (launch_resident)."""
        ctx = self.ctx
        stream = ctx.stream() if stream is None else stream
        if resident or (resident is None and ctx.resident_auto and self.resident_ok() and
                        all(v is None for v in (waves, chunk, recompute, gamma_global, slots))):
            if not self.resident_ok():
                raise ValueError("the register-resident kernel cannot run this batch")
            return self.launch_resident(groups, max_steps, 0, stream, latency)
        if self.D == 1:
            if ctx.one_d_wave(len(self.jobs), self.nmax):
                cfg = DhLaunch(1, groups, max_steps, 0, 0)
                cfg.delj = 1 if self.delj else 0
                check(ctx.lib.dh_integrate_1d_wave(self.tab.data_ptr(), len(self.jobs),
                                                   self.order, self.nmax,
                                                   self.counter.data_ptr(), cfg, stream),
                      "dh_integrate_1d_wave")
                return
            w = waves or 1
            thr = 64 * w
            g = groups or max(1, math.ceil(len(self.jobs) / thr))
            nl = g * thr
            need = (8 if self.delj else 6) * self.nmax * nl
            scr = self.workspace.get("lanes")
            if scr is None or scr.numel() < need:
                scr = torch().empty(need, dtype=torch().float64, device=ctx.device)
                self.workspace["lanes"] = scr
            cfg = DhLaunch(w, g, max_steps, 0, 0)
            cfg.delj = 1 if self.delj else 0
            check(ctx.lib.dh_integrate_1d(self.tab.data_ptr(), len(self.jobs), self.order,
                                          self.nmax, self.counter.data_ptr(), scr.data_ptr(),
                                          scr.numel(), cfg, stream), "dh_integrate_1d")
        elif max_steps <= 0 and self.spread_bound(waves, chunk, recompute, gamma_global, slots):
            self.launch_spread(stream)
        elif self.latency_bound(waves, chunk, recompute, gamma_global, slots):
            ctx.factored_launch(self.tab.data_ptr(), len(self.jobs), self.D, self.order,
                                self.nmax, self.counter.data_ptr(), stream, self.workspace,
                                waves, groups, max_steps)
        else:
            return ctx.nd_launch(self.tab.data_ptr(), len(self.jobs), self.D, self.order,
                                 self.nmax, self.counter.data_ptr(), stream, self.workspace,
                                 waves, groups, max_steps, chunk, recompute, gamma_global,
                                 slots, self.packable, self.delj)

    def spread_bound(self, waves=None, chunk=None, recompute=None, gamma_global=None,
                     slots=None):
        """Whether this batch runs with each integration's sweeps spread over
        several workgroups (dh_integrate_batch): three or more populations,
        fewer integrations than twice the compute units (one workgroup each
        would leave the GPU idle while each integration's thousands of lines
        go through one compute unit), and no launch shape asked for."""
        ctx = self.ctx
        if self.D < 3 or self.delj or ctx.spread_max_inst <= 0:
            return False
        if any(v is not None for v in (waves, chunk, recompute, gamma_global, slots)):
            return False
        return len(self.jobs) <= ctx.spread_max_inst

    def launch_spread(self, stream, waves=None, steps_per_graph=8, max_wgs=4096):
        """Run every job to completion with dh_integrate_batch, in phases:
        each phase launches only the integrations still running, and ends
        (on a multiple of steps_per_graph) when half of them have finished,
        so the long tail of a batch of unequal integrations is not swept by
        the workgroups of finished ones (Context.spread_phases False: one
        phase).  Each phase picks its own launch shape."""
        ctx = self.ctx
        t = torch()
        raw = self.tab.cpu().numpy().view(INST_DTYPE)
        keep = []
        totals, sos_ptr, first_axis = [], [], []
        for j in self.jobs:
            act = [k for k in range(j.D) if not (j.frozen >> k) & 1]
            total = int(np.sum(j.seg["nsteps"])) if act else 0
            sos = ctx.to_dev(np.repeat(np.arange(len(j.seg), dtype=np.int32), j.seg["nsteps"])
                             if total else np.zeros(1, dtype=np.int32))
            keep.append(sos)
            totals.append(total)
            sos_ptr.append(sos.data_ptr())
            first_axis.append(act[0] if act else 0)
        longest = max(totals) if totals else 0
        if longest == 0:
            return
        spg = steps_per_graph
        for start, end, live in spread_phases(totals, spg, ctx.spread_phases):
            tiles = [-(-(self.jobs[i].n ** (self.D - 1)) // 64) for i in live]
            total_tiles = sum(tiles)
            w_per = waves
            if w_per is None:
                # two waves per workgroup (each lane's whole gamma column in
                # LDS) while that leaves at most one workgroup per compute
                # unit, else four (more workgroups resident per compute unit)
                w_per = ctx.spread_waves or (2 if total_tiles <= 2 * ctx.num_cu else 4)
            # workgroups per integration: one wave per 64-line tile while the
            # launch stays within max_wgs workgroups
            scale = min(1.0, max_wgs * w_per / max(total_tiles, 1))
            items = np.zeros(len(live), dtype=ITEM_DTYPE)
            wg = 0
            wg_item = []
            for k, (i, nt) in enumerate(zip(live, tiles)):
                items[k]["inst"] = raw[i]
                items[k]["seg_of_step"] = sos_ptr[i]
                items[k]["total_steps"] = totals[i]
                items[k]["first_axis"] = first_axis[i]
                w = max(1, int(-(-nt * scale // w_per)))
                items[k]["wg_begin"] = wg
                items[k]["wgs"] = w
                wg_item.extend([k] * w)
                wg += w
            itd = ctx.to_dev(items.view(np.uint8))
            wgd = ctx.to_dev(np.asarray(wg_item, dtype=np.int32))
            base = t.zeros(1, dtype=t.int64, device=ctx.device)
            nmax = max(self.jobs[i].n for i in live)
            check(ctx.lib.dh_integrate_batch(itd.data_ptr(), wgd.data_ptr(), len(live), wg,
                                             self.D, self.order, nmax, w_per, base.data_ptr(),
                                             longest, start, end - start, spg, stream),
                  "dh_integrate_batch")
            keep.append((itd, wgd, base))
        # the buffers must outlive the queued work
        self.workspace["spread_keep"] = keep

    def latency_bound(self, waves=None, chunk=None, recompute=None, gamma_global=None,
                      slots=None):
        """Whether this batch goes to the factored kernel: two populations,
        at most factored_max_inst integrations (up to there each one's step
        latency, not the batch's bandwidth, sets the time, and the factored
        kernel's shorter steps win), long segments, and no launch shape asked
        for explicitly."""
        ctx = self.ctx
        if self.D != 2 or self.delj or ctx.factored_max_inst <= 0:
            return False
        if any(v is not None for v in (waves, chunk, recompute, gamma_global, slots)):
            return False
        return (len(self.jobs) <= ctx.factored_max_inst and
                ctx.factorable([j.seg for j in self.jobs]))

    def progress(self):
        raw = self.tab.cpu().numpy().view(INST_DTYPE)
        return raw


def spread_phases(totals, spg, phased=True):
    """Phases (start, end, live) of a spread batch whose integrations take
    ``totals`` steps: steps start .. end-1 run the integrations in ``live``
    (those with more than ``start`` steps), each phase ending on a multiple
    of ``spg`` steps after its start (or at the longest schedule) once half
    of its integrations have finished; phased False: one phase."""
    longest = max(totals) if len(totals) else 0
    out = []
    start = 0
    while start < longest:
        live = [i for i, tot in enumerate(totals) if tot > start]
        end = longest
        if phased and len(live) > 1:
            half = sorted(totals[i] for i in live)[len(live) // 2 - 1]
            end = min(longest, start + max(1, -(-(half - start) // spg)) * spg)
        out.append((start, end, live))
        start = end
    return out


def run_batch(jobs, ctx=None, waves=None, groups=0, sort=True, chunk=None, recompute=None,
              gamma_global=None, slots=None):
    """Run every job to completion.  Jobs are grouped by (D, order); a group
    of several populations is split further by (n, grid, row pitch) so that each part
    can pack several instances per workgroup, the parts running side by side
    on their own streams.  Each part is one persistent launch, longest job
    first."""
    t = torch()
    ctx = ctx or context()
    groups_by = {}
    for j in jobs:
        groups_by.setdefault((j.D, j.order, j.delj), []).append(j)
    parts = []
    explicit = any(v is not None for v in (waves, chunk, recompute, gamma_global, slots))
    for key in sorted(groups_by):
        js = groups_by[key]
        D, order, delj = key
        rows = ctx.lib.dh_resident_latency_rows
        if (D == 2 and not explicit and not delj and ctx.resident_auto and
                order in (_lib.ORDER_TV, _lib.ORDER_PRE) and all(rows(j.n) for j in js)):
            # the register-resident kernels take mixed grids of one row
            # bucket in a launch: one part per bucket, side by side
            by = {}
            for j in js:
                by.setdefault(rows(j.n), []).append(j)
            parts.extend(by[k] for k in sorted(by))
            continue
        # the latency paths take mixed grids in one launch: one launch, not
        # one per grid size on streams that may share a hardware queue
        one_launch = not explicit and not delj and (
            (D == 2 and len(js) <= ctx.factored_max_inst and
             ctx.factorable([j.seg for j in js])) or
            (D >= 3 and len(js) <= ctx.spread_max_inst))
        if key[0] >= 2 and slots != 1 and not one_launch:
            by = {}
            for j in js:
                by.setdefault((j.n, ctx.grid(j.xx).data_ptr(), pitch_of(j.phi)), []).append(j)
            parts.extend(by[k] for k in sorted(by))
        else:
            parts.append(js)
    cur = t.cuda.current_stream(ctx.device)
    streams = ctx.side_streams(len(parts)) if len(parts) > 1 else [cur]
    packs = []
    for js, st in zip(parts, streams):
        if sort:
            js = sorted(js, key=lambda j: -j.work())
        pb = PackedBatch(ctx, js)
        if st is not cur:
            st.wait_stream(cur)
        pb.launch(waves=waves, groups=groups, chunk=chunk, recompute=recompute,
                  gamma_global=gamma_global, slots=slots, stream=st.cuda_stream)
        packs.append(pb)
    for st in streams:
        st.synchronize()
    return packs


class Coalescer:
    """Gathers the integrations that worker threads issue concurrently and
    runs them as one batched persistent launch whenever every live worker is
    waiting for one.  Model functions written against the ordinary
    Integration API thereby batch across evaluations without changes."""

    def __init__(self, ctx=None):
        self._ctx = ctx          # resolved at the first flush: no GPU needed until then
        self.cv = threading.Condition()
        self.waiting = []
        self.live = 0
        self.launches = 0

    def register(self, counted=False):
        """Make the calling thread a worker.  counted: its place in ``live``
        was already taken by expect()."""
        if not counted:
            with self.cv:
                self.live += 1
        _local.coalescer = self

    def expect(self, k):
        """Count k workers as live before they start, so that the first
        ones to submit do not flush a batch without the rest."""
        with self.cv:
            self.live += k

    def unregister(self):
        _local.coalescer = None
        with self.cv:
            self.live -= 1
            self._maybe_flush()

    def suspend(self):
        """The calling worker stops counting as live (it is about to wait for
        something other than a launch, e.g. threads of its own)."""
        with self.cv:
            self.live -= 1
            self._maybe_flush()

    def resume(self):
        with self.cv:
            self.live += 1

    def submit(self, job):
        with self.cv:
            self.waiting.append(job)
            self._maybe_flush()
            while not job.done:
                self.cv.wait()
        if job.error is not None:
            raise job.error

    def _maybe_flush(self):
        """Runs the batch once every live worker waits; a launch that fails
        raises its error in every worker whose job it held."""
        if self.waiting and len(self.waiting) >= self.live:
            jobs, self.waiting = self.waiting, []
            err = None
            try:
                if self._ctx is None:
                    self._ctx = context()
                samp = [j for j in jobs if isinstance(j, SampleJob)]
                ints = [j for j in jobs if not isinstance(j, SampleJob)]
                if ints:
                    staged = _upload_host_jobs(self._ctx, ints)
                    run_batch(ints, self._ctx)
                    _download_host_jobs(staged)
                if samp:
                    from .Spectrum_mod import sample_jobs
                    sample_jobs(samp, self._ctx)
            except BaseException as e:  # noqa: BLE001  (raised in every worker)
                err = e
            for j in jobs:
                j.error = err
                j.done = True
            self.launches += 1
            self.cv.notify_all()


class SampleJob:
    """A phi -> spectrum sampling submitted to a Coalescer: host phi in, host
    spectrum data (``result``) out; the launch samples all waiting phis of
    one shape, sample sizes and grid together."""
    __slots__ = ("phi", "ns", "xx", "done", "error", "result")

    def __init__(self, phi, ns, xx):
        self.phi, self.ns, self.xx = phi, tuple(int(v) for v in ns), xx
        self.done = False
        self.error = None
        self.result = None


def _upload_host_jobs(ctx, jobs):
    """Device phis for the jobs that carry host arrays: one arena and one
    host-to-device copy per shape."""
    t = torch()
    groups = {}
    for j in jobs:
        if getattr(j, "host", None) is not None and j.phi is None:
            groups.setdefault(j.host.shape, []).append(j)
    staged = []
    for shape, grp in groups.items():
        arena = ctx.empty_phi(shape, batch=len(grp))
        arena.copy_(t.from_numpy(np.stack([j.host for j in grp])))
        for i, j in enumerate(grp):
            j.phi = arena[i]
        staged.append((arena, grp))
    return staged


def _download_host_jobs(staged):
    """The integrated arrays back to the host, one copy per arena."""
    for arena, grp in staged:
        res = arena.cpu().numpy()
        for i, j in enumerate(grp):
            j.result = np.array(res[i])
            j.phi = None


#: map_in_worker runs its items one after another once a coalescer has this
#: many workers (bounds the threads of one process)
max_worker_threads = 768


def current_coalescer():
    """The Coalescer the calling thread works for, or None."""
    return getattr(_local, "coalescer", None)


def map_in_worker(func, items):
    """[func(x) for x in items].  Called from a Coalescer's worker thread, the
    items after the first run in threads of their own that join the same
    coalescer, so that the integrations of all items go into the same batched
    launches instead of one launch round per item (the grids of an
    extrapolated model evaluation); elsewhere a plain loop.  An exception of
    any item is raised after all have finished (the first in item order)."""
    co = current_coalescer()
    if co is None or len(items) < 2 or co.live + len(items) - 1 > max_worker_threads:
        return [func(x) for x in items]
    n = len(items)
    out = [None] * n
    err = [None] * n

    def run(i):
        co.register(counted=True)
        try:
            out[i] = func(items[i])
        except BaseException as e:  # noqa: BLE001  (re-raised by the caller)
            err[i] = e
        finally:
            co.unregister()

    threads = [threading.Thread(target=run, args=(i,), daemon=True) for i in range(1, n)]
    co.expect(n - 1)
    for th in threads:
        th.start()
    try:
        out[0] = func(items[0])
    except BaseException as e:  # noqa: BLE001
        err[0] = e
    try:
        co.suspend()
    finally:
        for th in threads:
            th.join()
        co.resume()
    for e in err:
        if e is not None:
            raise e
    return out


def map_batched(func, args_list, workers=256, ctx=None):
    """[func(*a) for a in args_list], evaluated by ``workers`` threads whose
    integrations are coalesced into batched launches.  Exceptions are
    returned in place of results (one failing evaluation does not stop the
    others)."""
    co = Coalescer(ctx)
    results = [None] * len(args_list)
    nxt = [0]
    lock = threading.Lock()

    def worker():
        co.register(counted=True)
        try:
            while True:
                with lock:
                    i = nxt[0]
                    nxt[0] += 1
                if i >= len(args_list):
                    return
                try:
                    results[i] = func(*args_list[i])
                except Exception as e:  # noqa: BLE001  (reported to the caller)
                    results[i] = e
        finally:
            co.unregister()

    nthreads = max(1, min(workers, len(args_list)))
    threads = [threading.Thread(target=worker, daemon=True) for _ in range(nthreads)]
    co.expect(nthreads)   # every worker counts from the start: one batch, not several
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    return results
