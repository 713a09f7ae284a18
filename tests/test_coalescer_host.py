"""
CPU tests of the coalescer that merges the integrations of concurrent model
evaluations into shared launches (the launch itself is replaced by a
recorder: no GPU).
"""
import random
import time

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
