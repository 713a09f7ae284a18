"""
Serving loop: continuous batching of mixed-dimensionality model evaluations.
GPU tests check that every result is bit-identical alone and in the mix (and
equal to the sequential dadi-style model functions), that admission by HBM
footprint queues instead of failing, and that a poisoned request fails alone.
The CPU test checks the host-side request construction.
"""
import numpy as np
import pytest


def test_request_mix_host_side():
    from dadi_amd import serve
    rng = np.random.default_rng(3)
    reqs = serve.mixed_requests(rng, 400, rate=100.0, scale=0.01)
    kinds = [r.prog.kind for r in reqs]
    frac = {k: kinds.count(k) / len(kinds) for k in set(kinds)}
    assert abs(frac["split_mig"] - 0.55) < 0.08
    assert abs(frac["three_epoch"] - 0.30) < 0.08
    assert frac.get("four_pop_chain", 0) < 0.08
    arr = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(arr) >= 0)
    assert abs(len(arr) / arr[-1] - 100.0) < 20
    for r in reqs:
        fp = serve.Server.footprint(r.prog)
        assert fp >= 3 * 8 * r.prog.pts ** {"split_mig": 2, "three_epoch": 1,
                                            "three_pop_chain": 3,
                                            "four_pop_chain": 4}[r.prog.kind]


@pytest.fixture(scope="module")
def dadi_amd():
    import dadi_amd as d
    assert d.cuda_enabled(True)
    return d


def _small_mix(seed, n):
    from dadi_amd import serve
    rng = np.random.default_rng(seed)
    reqs = serve.mixed_requests(rng, n, rate=1e6, scale=0.01)
    return reqs


@pytest.mark.gpu
def test_results_identical_alone_and_in_mix(dadi_amd):
    from dadi_amd import serve
    mix = _small_mix(11, 24)
    srv = serve.Server(quantum=7)
    stats = srv.run(mix, realtime=False)
    assert stats["failed"] == 0 and stats["requests"] == 24
    for r in mix[:8]:
        alone = serve.Request(0, r.prog, 0.0)
        serve.Server(quantum=1000).run([alone], realtime=False)
        assert np.array_equal(alone.result, r.result), r.prog.kind


@pytest.mark.gpu
def test_served_split_mig_equals_model_function(dadi_amd):
    """A served split_mig request equals Demographics2D.split_mig called
    directly (same integration bits, same device sampling)."""
    from dadi_amd import serve
    params, ns, pts = (2.0, 0.4, 0.03, 1.5), (12, 15), 37
    req = serve.Request(0, serve.prog_split_mig(params, ns, pts), 0.0)
    serve.Server().run([req], realtime=False)
    fs = dadi_amd.Demographics2D.split_mig(params, ns, pts)
    assert np.array_equal(req.result, fs.data)


@pytest.mark.gpu
def test_budget_queues_and_oversize_fails_alone(dadi_amd):
    from dadi_amd import serve
    mix = _small_mix(5, 10)
    biggest = max(serve.Server.footprint(r.prog) for r in mix)
    srv = serve.Server(budget_bytes=biggest, quantum=50)   # one at a time at worst
    stats = srv.run(mix, realtime=False)
    assert stats["failed"] == 0 and stats["requests"] == 10
    big = serve.Request(99, serve.prog_split_mig((1, 1, 0.01, 1), (10, 10), 120), 0.0)
    small = serve.Request(100, serve.prog_split_mig((1, 1, 0.01, 1), (10, 10), 40), 0.0)
    srv2 = serve.Server(budget_bytes=serve.Server.footprint(small.prog), quantum=50)
    st2 = srv2.run([big, small], realtime=False)
    assert big.error is not None and small.error is None
    assert st2["failed"] == 1


@pytest.mark.gpu
def test_poisoned_request_fails_alone(dadi_amd):
    from dadi_amd import serve
    mix = _small_mix(8, 12)
    bad = serve.Request(50, serve.prog_split_mig((float("nan"), 1.0, 0.02, 1.0), (10, 10), 50),
                        0.0)
    solo = []
    for r in mix[:4]:
        a = serve.Request(0, r.prog, 0.0)
        serve.Server().run([a], realtime=False)
        solo.append(a.result)
    serve.Server(quantum=9).run(mix + [bad], realtime=False)
    assert not np.all(np.isfinite(bad.result))
    for r, s in zip(mix[:4], solo):
        assert np.array_equal(r.result, s)


@pytest.mark.gpu
def test_wide_path_equals_batched_path(dadi_amd):
    """3D and 4D epochs served on the multi-workgroup paths (every sweep spread
    over the GPU, steps replayed from a captured graph in quanta of a few
    steps: one graph per request, or one shared launch per sweep for all
    in-flight epochs of a class at their own steps) give the bits of the
    one-workgroup batched path."""
    from dadi_amd import serve
    mix = _small_mix(21, 120)
    progs = ([r.prog for r in mix if r.prog.kind == "three_pop_chain"][:4]
             + [r.prog for r in mix if r.prog.kind == "four_pop_chain"][:2])
    assert len(progs) == 6
    narrow = [serve.Request(i, p, 0.0) for i, p in enumerate(progs)]
    st2 = serve.Server(wide_lines=1 << 60, quantum=5).run(narrow, realtime=False)
    assert st2["failed"] == 0 and st2["wide_launches"] == 0
    for mode in ("graph", "spread"):
        wide = [serve.Request(i, p, 0.0) for i, p in enumerate(progs)]
        srv = serve.Server(wide_lines=1, steps_per_graph=4, quantum_ms=0.5, wide_mode=mode)
        st = srv.run(wide, realtime=False)
        assert st["failed"] == 0 and st["wide_launches"] > 2, (mode, st)
        for a, b in zip(wide, narrow):
            assert np.array_equal(a.result, b.result), (mode, a.prog.kind)


@pytest.mark.gpu
def test_graph_range_rejects_ragged_quanta(dadi_amd):
    from dadi_amd import _lib, engine
    import torch
    ctx = engine.context()
    inst = _lib.DhInstance(0, 0, 0, 1, 0, 0, 3, 10, 0, 0, 0, 0, 0)
    base = torch.zeros(1, dtype=torch.int64, device=ctx.device)
    rc = ctx.lib.dh_integrate_graph_range(inst, base.data_ptr(), base.data_ptr(), 100, 0, 10,
                                          4, 2, ctx.stream())
    assert rc == -1   # DH_EINVAL


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["three_epoch", "three_pop_chain", "four_pop_chain"])
def test_served_programs_equal_model_functions(dadi_amd, kind):
    """Served 1D, 3D-chain and 4D-chain requests (phi resident on the device
    across epochs and device-side splits) equal the sequential model functions
    (Demographics1D.three_epoch, Demographics3D.three_pop_chain /
    four_pop_chain), alone and inside a mix of other requests."""
    from dadi_amd import serve
    if kind == "three_epoch":
        params, ns, pts = (0.4, 2.5, 0.05, 0.02), (17,), 150
        prog = serve.prog_three_epoch(params, ns, pts)
        fs = dadi_amd.Demographics1D.three_epoch(params, ns, pts)
    elif kind == "three_pop_chain":
        params = (1.5, 0.02, 0.7, 2.2, 0.01, 1.3, 0.5, 1.9, 3.1, 0.004, 0.8)
        ns, pts = (6, 7, 5), 24
        prog = serve.prog_three_pop_chain(params, ns, pts)
        fs = dadi_amd.Demographics3D.three_pop_chain(params, ns, pts)
    else:
        params = (1.5, 0.02, 0.7, 2.2, 0.01, 1.3, 0.5, 1.9, 3.1, 0.004, 0.8,
                  0.9, 1.4, 0.6, 2.0, 0.003, 0.5)
        ns, pts = (4, 5, 3, 4), 14
        prog = serve.prog_four_pop_chain(params, ns, pts)
        fs = dadi_amd.Demographics3D.four_pop_chain(params, ns, pts)
    alone = serve.Request(0, prog, 0.0)
    serve.Server().run([alone], realtime=False)
    assert np.array_equal(alone.result, np.asarray(fs.data))
    mixed = serve.Request(1, prog, 0.0)
    st = serve.Server(quantum=5).run(_small_mix(13, 10) + [mixed], realtime=False)
    assert st["failed"] == 0
    assert np.array_equal(mixed.result, np.asarray(fs.data))
