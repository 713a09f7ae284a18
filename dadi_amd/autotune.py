"""
Autotuner for the persistent workgroup kernel.

Per (populations D, grid size n) it times the launch shapes the kernel
supports on a synthetic batch that fills the GPU and records the fastest:

* waves per workgroup (64 tridiagonal systems per wave): the systems in
  flight per instance, traded against instances per compute unit;
* the Thomas scratch layout: the whole gamma column of each lane in LDS
  (chunk 0: least VALU, most LDS) or a bet checkpoint every 8 rows with the
  gammas recomputed in registers (chunk 8: ~8x less LDS, more VALU, and the
  contiguous axis staged through LDS tiles);
* with chunk 8, whether u' is written over phi between the two Thomas passes
  or recomputed from checkpoints (recompute 1: a quarter less phi traffic,
  half the writes, more VALU);
* the gamma columns in a global workspace instead (gamma_global 1: no
  recomputation and no gamma LDS, 16 more bytes of memory traffic per point
  and sweep);
* instances per workgroup (slots >= 2: the packed kernel, whose lockstep
  instances fill the lanes a lone instance leaves idle).

The axis order of the sweeps is not a tunable: it is the operator splitting
of the reference (x, then y, ...) and changing it would change results.

Results live in a JSON table (``tuning_gfx950.json`` beside this file by
default, or $DADI_AMD_TUNING) that engine.Context consults.
"""
import json
import os
import time

import numpy as np

from . import _lib, engine
from . import schedule as S

DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_gfx950.json")


class Tuning:
    def __init__(self, table=None, path=None):
        self.table = table or {}
        self.path = path

    @classmethod
    def load(cls, path=None):
        path = path or os.environ.get("DADI_AMD_TUNING") or DEFAULT_PATH
        try:
            with open(path) as f:
                table = json.load(f)
        except (OSError, ValueError):
            table = {}
        return cls(table, path)

    def lookup(self, D, n, packable=True):
        """engine.Shape tuned for (D, n), or for the nearest tuned n of the
        same D within 25%; None if nothing applies.  packable=False: the
        fastest recorded shape with one instance per workgroup."""
        hit = self.table.get("%d:%d" % (D, n))
        if hit is None:
            best = None
            for k, v in self.table.items():
                d, m = (int(x) for x in k.split(":"))
                if d == D and abs(m - n) <= 0.25 * n and (best is None or abs(m - n) < best[0]):
                    best = (abs(m - n), v)
            hit = best[1] if best else None
        if hit is None:
            return None
        pool = [hit] + list(hit.get("candidates", []))
        if not packable:
            pool = [c for c in pool if c.get("slots", 1) < 2]
        if not pool:
            return None
        top = max(pool, key=lambda c: c.get("points_per_s", 0.0)) if len(pool) > 1 else pool[0]
        return _shape_of(top)

    def record(self, D, n, shape, rate, candidates):
        self.table["%d:%d" % (D, n)] = dict(_fields_of(shape), points_per_s=float(rate),
                                            candidates=candidates)

    def save(self, path=None):
        path = path or self.path or DEFAULT_PATH
        with open(path, "w") as f:
            json.dump(self.table, f, indent=1, sort_keys=True)


def _fields_of(shape):
    return {"waves": shape.waves, "chunk": shape.chunk, "recompute": shape.recompute,
            "gamma_global": shape.gamma_global, "slots": shape.slots}


def _shape_of(rec):
    return engine.Shape(rec["waves"], rec["chunk"], rec.get("recompute", 0),
                        rec.get("gamma_global", 0), rec.get("slots", 1))


def _synthetic_jobs(ctx, D, n, count, steps, rng):
    import torch
    xx = np.asarray(__import__("dadi_amd").Numerics.default_grid(n))
    order = _lib.ORDER_PRE if 2 <= D <= 3 else (_lib.ORDER_PRE1 if D == 1 else _lib.ORDER_TV)
    arena = ctx.empty_phi((n,) * D, batch=count)
    arena.uniform_(0.5, 1.5)
    jobs = []
    for i in range(count):
        nus = list(np.exp(rng.uniform(np.log(0.2), np.log(5), D)))
        migs = [list(rng.uniform(0, 3, D - 1)) for _ in range(D)]
        gammas = list(rng.uniform(-5, 5, D))
        hs = [0.5] * D
        seg = S.const_segments(D, xx, 0.0, steps * 1e-4 * (1 - 1e-9), 1e-4, nus, migs, gammas,
                               hs, 1.0)
        jobs.append(engine.Job(arena[i], xx, seg, order, 0, (1 << D) - 1))
    return arena, jobs


def candidates(ctx, D, n):
    nsys = n ** (D - 1)
    out = []
    for chunk, ru, gg in ((0, 0, 0), (8, 0, 0), (8, 1, 0), (0, 0, 1)):
        wmax = ctx.max_waves(D, n, chunk, ru, gg)
        for w in (1, 2, 4, 8):
            if w <= wmax and (w == 1 or 64 * (w // 2) < nsys):
                out.append(engine.Shape(w, chunk, ru, gg))
    if D >= 2:
        # several instances per workgroup: every slot count that fills 1-8 waves
        Lp = -(-nsys // 8) * 8
        seen = set()
        for w in range(1, 9):
            S = min(16, (64 * w) // Lp)
            if S >= 2 and S not in seen and ctx.pack_plan(D, n, w, S) is not None:
                seen.add(S)
                out.append(engine.Shape(w, 8, 0, 0, S))
    return out


def tune(D, n, ctx=None, steps=30, count=None, reps=2, tuning=None, save=False):
    """Time every candidate shape for (D, n); record and return the best."""
    import torch
    ctx = ctx or engine.context()
    tuning = tuning or ctx.tuning or Tuning.load()
    rng = np.random.default_rng(1000 * D + n)
    count = count or 12 * ctx.num_cu   # enough to fill the GPU with packed workgroups too
    arena, jobs = _synthetic_jobs(ctx, D, n, count, steps, rng)
    pb = engine.PackedBatch(ctx, jobs)
    work = sum(int(np.sum(j.seg["nsteps"])) for j in jobs) * (n ** D) * D
    results = []
    for sh in candidates(ctx, D, n):
        best = None
        for _ in range(reps + 1):
            pb.restart()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            pb.launch(waves=sh.waves, chunk=sh.chunk, recompute=sh.recompute,
                      gamma_global=sh.gamma_global, slots=sh.slots)
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            best = el if best is None else min(best, el)   # first run warms up
        results.append(dict(_fields_of(sh), points_per_s=work / best))
    top = max(results, key=lambda r: r["points_per_s"])
    shape = _shape_of(top)
    tuning.record(D, n, shape, top["points_per_s"], results)
    if save:
        tuning.save()
    ctx.tuning = tuning
    return shape
