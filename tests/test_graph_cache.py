"""
The step-graph cache of dh_integrate_graph / dh_integrate_batch under
concurrent use: two host threads capture more distinct graphs than the cache
holds (64), so entries are evicted while the other thread is between looking
up (or inserting) its own graph and launching it.  Every integration must
still give the oracle's bits.
"""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_two_threads_overflowing_the_graph_cache_stay_exact(oracle):
    import dadi_amd
    from dadi_amd import engine, schedule as S, _lib
    assert dadi_amd.cuda_enabled(True)
    rng = np.random.default_rng(31)
    cases = []
    for i in range(2 * 48):
        n = int(rng.choice([12, 14, 16]))
        xx = oracle.default_grid(n)
        phi0 = rng.uniform(0.1, 1.0, (n, n))
        nus = list(rng.uniform(0.3, 2.0, 2))
        migs = [[float(rng.uniform(0, 2))], [float(rng.uniform(0, 2))]]
        T = float(rng.uniform(0.001, 0.004))
        dt = min(S.compute_dt(nus[k], migs[k], 0.0, 0.5) for k in range(2))
        seg = S.const_segments(2, xx, 0.0, T, dt, nus, migs, [0.0, 0.0], [0.5, 0.5], 1.0)
        want = oracle.integrate_tv(phi0, xx, T, nus, migs, [0.0, 0.0], [0.5, 0.5])
        cases.append((phi0, xx, seg, want))
    fz, inj = S.masks(2, [False, False], [False, False])
    got = [None] * len(cases)
    errors = []

    def worker(idx):
        try:
            for i in idx:
                phi0, xx, seg, _ = cases[i]
                # each call has its own buffers, hence its own graph
                got[i] = engine.run_single(phi0.copy(), xx, seg, _lib.ORDER_TV, fz, inj,
                                           mode="graph", steps_per_graph=3)
        except Exception as e:   # reported by the main thread
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(range(k, len(cases), 2),))
               for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for i, (_, _, _, want) in enumerate(cases):
        assert np.array_equal(got[i], want), i
