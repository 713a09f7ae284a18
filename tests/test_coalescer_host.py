"""
CPU tests of the coalescer that merges the integrations of concurrent model
evaluations into shared launches (the launch itself is replaced by a
recorder: no GPU), and of the phase plan of spread batches.
"""
import random
import time

import numpy as np
import pytest

from dadi_amd import engine


class _Job:
    def __init__(self, k):
        self.k = k
        self.done = False


def test_map_batched_gives_one_batch_however_threads_start(monkeypatch):
    """Every evaluation submits one integration; however late some worker
    threads start, the first ones to submit must not flush a partial batch:
    all integrations go to one launch."""
    batches = []
    monkeypatch.setattr(engine, "run_batch", lambda jobs, ctx=None: batches.append(len(jobs)))
    rng = random.Random(4)
    # the first evaluations submit at once, while later threads are still
    # being started; the rest after uneven delays
    delays = [0.0 if k < 8 else rng.random() * 0.02 for k in range(300)]

    def model(k):
        time.sleep(delays[k])              # workers reach their submission unevenly
        job = _Job(k)
        engine._local.coalescer.submit(job)
        assert job.done
        return k

    out = engine.map_batched(model, [(k,) for k in range(300)], workers=512, ctx=object())
    assert out == list(range(300))
    assert batches == [300]


def test_map_batched_second_round_batches_again(monkeypatch):
    """Evaluations that integrate twice (two epochs) give two full batches."""
    batches = []
    monkeypatch.setattr(engine, "run_batch", lambda jobs, ctx=None: batches.append(len(jobs)))

    def model(k):
        for _ in range(2):
            job = _Job(k)
            engine._local.coalescer.submit(job)
        return k

    engine.map_batched(model, [(k,) for k in range(64)], workers=16, ctx=object())
    assert sum(batches) == 128 and len(batches) <= 16


def test_grids_of_an_extrapolated_evaluation_share_one_launch(monkeypatch):
    """Inside map_batched the grids of make_extrap_log_func run side by side:
    every evaluation's three integrations go into the first launch (one
    round, not three), and results come back in grid order."""
    from dadi_amd import Numerics
    batches = []
    monkeypatch.setattr(engine, "run_batch", lambda jobs, ctx=None: batches.append(len(jobs)))
    monkeypatch.setattr(Numerics, "parallel_grids", True)

    def model(k, pts):
        time.sleep(0.001 * (pts % 7))
        job = _Job(k)
        engine._local.coalescer.submit(job)
        assert job.done
        return (k, pts)

    out = engine.map_batched(lambda k: Numerics.make_extrap_func(model)(k, [40, 50, 60],
                                                                         no_extrap=True),
                             [(k,) for k in range(40)], workers=40, ctx=object())
    assert out == [[(k, 40), (k, 50), (k, 60)] for k in range(40)]
    assert batches == [120]


def test_grid_threads_report_errors_and_release_the_batch(monkeypatch):
    """A grid that raises is reported to its evaluation; the other
    evaluations' integrations still run (no worker is left waiting)."""
    batches = []
    monkeypatch.setattr(engine, "run_batch", lambda jobs, ctx=None: batches.append(len(jobs)))

    def model(k, pts):
        if k == 3 and pts == 50:
            raise ValueError("bad grid")
        job = _Job(k)
        engine._local.coalescer.submit(job)
        return pts

    out = engine.map_batched(lambda k: engine.map_in_worker(lambda p: model(k, p), [40, 50, 60]),
                             [(k,) for k in range(8)], workers=8, ctx=object())
    assert isinstance(out[3], ValueError)
    assert all(out[k] == [40, 50, 60] for k in range(8) if k != 3)
    assert sum(batches) == 7 * 3 + 2


def test_map_in_worker_outside_a_coalescer_is_a_loop():
    assert engine.map_in_worker(lambda x: x * x, [1, 2, 3]) == [1, 4, 9]


def test_failed_launch_raises_in_every_worker_of_it(monkeypatch):
    """A launch that fails raises its error in each evaluation whose
    integration it held, not only in the thread that ran it."""
    def launch(jobs, ctx=None):
        raise RuntimeError("launch failed")

    monkeypatch.setattr(engine, "run_batch", launch)

    def model(k):
        job = _Job(k)
        engine._local.coalescer.submit(job)
        return k

    out = engine.map_batched(model, [(k,) for k in range(16)], workers=16, ctx=object())
    assert all(isinstance(r, RuntimeError) for r in out)


@pytest.mark.parametrize("seed", range(6))
def test_spread_phases_cover_every_step_once(seed):
    """The phases of a spread batch cover steps 0 .. longest-1 once, in
    order; each runs exactly the integrations that still have steps then;
    all but the last end a multiple of the graph's steps after they start;
    each drops at least half of its integrations before the next; without
    phasing there is one phase with everything."""
    from dadi_amd.engine import spread_phases
    rng = np.random.default_rng(seed)
    totals = list(rng.integers(0, 5000, rng.integers(1, 300)))
    totals[0] = max(totals[0], 1)
    spg = int(rng.choice([1, 4, 8, 16]))
    ph = spread_phases(totals, spg)
    longest = max(totals)
    assert ph[0][0] == 0 and ph[-1][1] == longest
    for (s0, e0, live), nxt in zip(ph, ph[1:] + [None]):
        assert s0 < e0
        assert live == [i for i, t in enumerate(totals) if t > s0]
        if nxt is not None:
            assert nxt[0] == e0
            assert (e0 - s0) % spg == 0
            assert len(nxt[2]) <= len(live) - len(live) // 2
    one = spread_phases(totals, spg, phased=False)
    assert one == [(0, longest, [i for i, t in enumerate(totals) if t > 0])]
    assert spread_phases([0, 0], spg) == []
