"""
dh_integrate_factored_static, the entry point that allocates nothing (its
factorisation workspace and queue counter are the library's static device
scratch), against dh_integrate_factored with a caller-provided workspace:
bit-identical results on two calls in a row, the second on a smaller input
(the scratch keeps the first call's contents), and the size checks.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(ctx, rng, count, n, D=2):
    """count integrations of D populations on n points with constant
    parameters: device phi, grid, schedules and an instance table."""
    import torch
    from dadi_amd import engine, schedule as S, _lib
    xx = ctx.grid(np.asarray(engine_grid(n), dtype=np.float64))
    phis, segs, recs = [], [], []
    for _ in range(count):
        phi0 = rng.uniform(0.1, 1.0, (n,) * D)
        nus = list(rng.uniform(0.3, 3.0, D))
        migs = [list(rng.uniform(0, 2, D - 1)) for _ in range(D)]
        gammas = list(rng.uniform(-4, 4, D))
        hs = list(rng.uniform(0.2, 0.8, D))
        T = float(rng.uniform(0.002, 0.01))
        dt = min(S.compute_dt(nus[k], migs[k], gammas[k], hs[k]) for k in range(D))
        seg = S.const_segments(D, engine_grid(n), 0.0, T, dt, nus, migs, gammas, hs, 1.0)
        phis.append(phi0)
        segs.append(ctx.to_dev(seg.view(np.uint8)))
        recs.append((seg, len(seg)))
    return xx, phis, segs, recs


def engine_grid(n):
    from dadi_amd import Numerics
    return Numerics.default_grid(n)


def _run(ctx, xx, phis, segs, recs, D, n, static):
    """One factored call over the batch; returns the integrated phis."""
    import torch
    from dadi_amd import engine, _lib
    dev = [ctx.to_dev(p.copy()) for p in phis]
    tab = np.concatenate([engine._inst_record(d.data_ptr(), xx.data_ptr(), s.data_ptr(), ns, D,
                                              n, _lib.ORDER_PRE, 0, (1 << D) - 1)
                          for d, s, (_, ns) in zip(dev, segs, recs)])
    tabd = ctx.to_dev(tab.view(np.uint8))
    cfg = _lib.DhLaunch(2, 0, 0, 0, 0)
    st = ctx.stream()
    if static:
        rc = ctx.lib.dh_integrate_factored_static(tabd.data_ptr(), len(dev), D, _lib.ORDER_PRE,
                                                  n, cfg, st)
        _lib.check(rc, "dh_integrate_factored_static")
    else:
        slice_d = int(ctx.lib.dh_factored_slice_doubles(D, n))
        groups = min(len(dev), ctx.num_cu)
        ws = torch.empty(groups * slice_d, dtype=torch.float64, device=ctx.device)
        counter = torch.zeros(1, dtype=torch.int32, device=ctx.device)
        cfg.gamma_dev = ws.data_ptr()
        cfg.gamma_doubles = ws.numel()
        rc = ctx.lib.dh_integrate_factored(tabd.data_ptr(), len(dev), D, _lib.ORDER_PRE, n,
                                           counter.data_ptr(), cfg, st)
        _lib.check(rc, "dh_integrate_factored")
    torch.cuda.synchronize()
    return [d.cpu().numpy() for d in dev]


def test_static_scratch_entry_point_equals_workspace_entry_point():
    import dadi_amd
    from dadi_amd import engine, _lib
    assert dadi_amd.cuda_enabled(True)
    ctx = engine.context()
    rng = np.random.default_rng(5)
    # first call: 12 integrations on 60 points; second: 3 on 24 points
    for count, n in ((12, 60), (3, 24)):
        print("static scratch call: %d integrations, n = %d" % (count, n), flush=True)
        xx, phis, segs, recs = _batch(ctx, rng, count, n)
        want = _run(ctx, xx, phis, segs, recs, 2, n, static=False)
        got = _run(ctx, xx, phis, segs, recs, 2, n, static=True)
        for a, b in zip(got, want):
            assert np.array_equal(a, b)
        assert not np.array_equal(got[0], phis[0])   # something was integrated


def test_static_scratch_rejects_oversize_requests():
    """Host-side checks only: nothing is launched."""
    import dadi_amd
    from dadi_amd import engine, _lib
    assert dadi_amd.cuda_enabled(True)
    ctx = engine.context()
    cap = int(ctx.lib.dh_static_scratch_doubles())
    assert cap == 4 << 20
    one = int(ctx.lib.dh_factored_slice_doubles(4, 30))   # 17 * 30^4 > cap
    assert one > cap
    cfg = _lib.DhLaunch(2, 0, 0, 0, 0)
    fake_table = 1 << 40        # never dereferenced: the size check comes first
    rc = ctx.lib.dh_integrate_factored_static(fake_table, 1, 4, _lib.ORDER_TV, 30, cfg, None)
    assert rc == -3 and "static scratch" in _lib.last_error()
    slice2 = int(ctx.lib.dh_factored_slice_doubles(2, 100))
    cfg.groups = cap // slice2 + 1
    rc = ctx.lib.dh_integrate_factored_static(fake_table, 1000, 2, _lib.ORDER_PRE, 100, cfg,
                                              None)
    assert rc == -3 and "workgroups" in _lib.last_error()
    cfg.groups = 0
    cfg.gamma_dev = 1 << 40
    rc = ctx.lib.dh_integrate_factored_static(fake_table, 1, 2, _lib.ORDER_PRE, 100, cfg, None)
    assert rc == -1
