"""
GPU parity: the HIP path through the C ABI against the reference's own C
kernels' outputs (golden vectors) and the oracle.

Integration results must be bit-identical (same operations, same order, no
FMA contraction, IEEE division).  The device's sampling sums are compensated
(exactly rounded up to eps^2-level terms), numpy's BLAS dot sums in its own
order: spectra are compared with rtol 1e-12 against the reference-order
oracle and 1e-15 against its exactly rounded sampler; extrapolated spectra
against the stated 1e-10 on entries above 1e-14 of the total mass, and the
Poisson log-likelihood to 1e-9 absolute.
"""
import numpy as np
import pytest

from tests import golden_cases as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dadi_amd():
    import dadi_amd as d
    assert d.cuda_enabled(True), "HIP backend failed to initialise"
    return d


_FUN = {1: "one_pop", 2: "two_pops", 3: "three_pops", 4: "four_pops", 5: "five_pops"}


def _api_run(d, c, phi0, xx, force_tv):
    kw = G.api_kwargs(c, force_tv)
    T = kw.pop("T")
    f = getattr(d.Integration, _FUN[c["D"]])
    arg = phi0.copy()
    return f(arg, xx, T, **kw)


@pytest.mark.parametrize("case", range(len(G.meta())))
def test_integration_bit_exact_vs_reference(dadi_amd, case):
    c, phi0, want = G.case(case)
    xx = G.grid_of(c)
    print("case", case, "D", c["D"], "n", c["n"], flush=True)
    got = _api_run(dadi_amd, c, phi0, xx, force_tv=True)
    assert np.array_equal(got, want["tv"]), np.abs(got - want["tv"]).max()
    if "const" in want:
        got = _api_run(dadi_amd, c, phi0, xx, force_tv=False)
        assert np.array_equal(got, want["const"]), np.abs(got - want["const"]).max()


def test_four_five_pops_update_in_place(dadi_amd):
    c, phi0, want = G.case([i for i, m in enumerate(G.meta()) if m["D"] == 4][0])
    xx = G.grid_of(c)
    kw = G.api_kwargs(c, force_tv=False)
    T = kw.pop("T")
    arr = phi0.copy()
    out = dadi_amd.Integration.four_pops(arr, xx, T, **kw)
    assert out is arr
    assert np.array_equal(arr, want["tv"])


def test_graph_and_workgroup_paths_agree(dadi_amd, oracle):
    """The multi-workgroup per-step path (hipGraph) and the persistent
    workgroup path give identical bits, and both match the oracle."""
    from dadi_amd import engine, schedule as S, _lib
    rng = np.random.default_rng(7)
    for D, n in [(2, 70), (3, 21), (4, 11)]:
        xx = oracle.default_grid(n)
        phi0 = rng.uniform(0.1, 1, (n,) * D)
        nus = list(rng.uniform(0.3, 2, D))
        migs = [list(rng.uniform(0, 2, D - 1)) for _ in range(D)]
        gammas = list(rng.uniform(-3, 3, D))
        hs = list(rng.uniform(0.1, 0.9, D))
        dt = min(S.compute_dt(nus[k], migs[k], gammas[k], hs[k]) for k in range(D))
        seg = S.const_segments(D, xx, 0.0, 0.004, dt, nus, migs, gammas, hs, 1.0)
        fz, inj = S.masks(D, [False] * D, [False] * D)
        a = engine.run_single(phi0.copy(), xx, seg, _lib.ORDER_TV, fz, inj, mode="wg")
        b = engine.run_single(phi0.copy(), xx, seg, _lib.ORDER_TV, fz, inj, mode="graph",
                              steps_per_graph=5)
        ref = oracle.integrate_tv(phi0, xx, 0.004, nus, migs, gammas, hs)
        assert np.array_equal(a, ref)
        assert np.array_equal(b, ref)


def test_batch_equals_solo_and_isolates_bad_instances(dadi_amd, oracle):
    """Many integrations in one persistent launch give exactly the solo
    results; a NaN-parameter instance fails alone."""
    import torch
    from dadi_amd import engine, schedule as S, _lib
    ctx = engine.context()
    rng = np.random.default_rng(9)
    jobs, refs = [], []
    for i in range(40):
        n = int(rng.choice([20, 24, 30]))
        xx = oracle.default_grid(n)
        phi0 = oracle.phi_1D_to_2D(xx, oracle.phi_1D(xx))
        nu1, nu2 = np.exp(rng.uniform(np.log(0.1), np.log(10), 2))
        T = rng.uniform(0.001, 0.05)
        m = rng.uniform(0, 5)
        if i == 13:
            nu1 = float("nan")
        dt = min(S.compute_dt(nu1, [m], 0, 0.5), S.compute_dt(nu2, [m], 0, 0.5))
        seg = S.const_segments(2, xx, 0.0, T, dt, [nu1, nu2], [[m], [m]], [0, 0], [0.5, 0.5],
                               1.0)
        jobs.append(engine.Job(ctx.to_dev(phi0), xx, seg, _lib.ORDER_PRE, 0, 3))
        if i != 13:
            refs.append((i, oracle.integrate_const(phi0, xx, T, [nu1, nu2], [[m], [m]], [0, 0],
                                                   [0.5, 0.5])))
    engine.run_batch(jobs)
    for i, ref in refs:
        assert np.array_equal(jobs[i].phi.cpu().numpy(), ref), i
    assert not np.all(np.isfinite(jobs[13].phi.cpu().numpy()))


@pytest.mark.parametrize("kernel", ["lane", "wave"])
def test_1d_kernels_match_oracle(dadi_amd, oracle, kernel):
    """Both batched 1D kernels are exact: one lane per integration, and one
    wave per integration (Thomas factorisation formed once per segment,
    divisions through the known reciprocals).  DFE-style gamma sweep down to
    gamma = -2000, whose densities underflow at high frequency (the slow
    division path); time-varying sizes (one step per segment); a frozen
    population; several quanta per integration."""
    from dadi_amd import engine, schedule as S, _lib
    ctx = engine.context()
    gammas = np.concatenate([-np.logspace(np.log10(2000), -3, 40), [0.0, 3.0, 25.0]])
    jobs, refs = [], []
    for i, g in enumerate(gammas):
        n = [40, 50, 60, 130][i % 4]
        xx = oracle.default_grid(n)
        phi0 = oracle.phi_1D(xx, gamma=g)
        T = 0.01 if abs(g) < 500 else 0.002
        dt = S.compute_dt(2.0, [0], g, 0.5)
        seg = S.const_segments(1, xx, 0.0, T, dt, [2.0], [[]], [g], [0.5], 1.0)
        jobs.append(engine.Job(ctx.to_dev(phi0), xx, seg, _lib.ORDER_PRE1, 0, 1))
        refs.append(oracle.integrate_const(phi0, xx, T, [2.0], [[]], [g], [0.5]))
    xx = oracle.default_grid(70)
    phi0 = oracle.phi_1D(xx)
    nu = lambda t: 0.5 + 20 * t
    c = lambda v: (lambda t: v)
    seg = S.tv_segments(1, xx, 0.0, 0.03, [nu], [[]], [c(-3.0)], [c(0.3)], c(1.0), c(1.0))
    jobs.append(engine.Job(ctx.to_dev(phi0), xx, seg, _lib.ORDER_TV, 0, 1))
    refs.append(oracle.integrate_tv(phi0, xx, 0.03, [nu], [[]], [-3.0], [0.3]))
    seg = S.const_segments(1, xx, 0.0, 0.05, 1e-3, [1.0], [[]], [0.0], [0.5], 1.0)
    jobs.append(engine.Job(ctx.to_dev(phi0), xx, seg, _lib.ORDER_PRE1, 1, 1))
    refs.append(phi0)
    old = ctx.one_d_kernel
    ctx.one_d_kernel = kernel
    try:
        for order in (_lib.ORDER_PRE1, _lib.ORDER_TV):   # a batch shares one order
            b = engine.PackedBatch(ctx, [j for j in jobs if j.order == order])
            for _ in range(200):
                b.launch(max_steps=97)
                if np.all(b.progress()["status"] == 1):
                    break
            assert np.all(b.progress()["status"] == 1)
    finally:
        ctx.one_d_kernel = old
    for k, (j, r) in enumerate(zip(jobs, refs)):
        assert np.array_equal(j.phi.cpu().numpy(), r), k


@pytest.mark.parametrize("D,n,ns", [(1, 60, [20]), (2, 40, [10, 14]), (3, 16, [6, 5, 7]),
                                    (4, 9, [4, 3, 5, 4]), (5, 7, [3, 3, 2, 4, 3])])
def test_sampling_matches_oracle(dadi_amd, oracle, D, n, ns):
    rng = np.random.default_rng(D)
    xx = oracle.default_grid(n)
    phi = rng.uniform(0.1, 2.0, (n,) * D)
    fs = dadi_amd.Spectrum.from_phi(phi, ns, [xx] * D)
    want = oracle.from_phi(phi, ns, xx)
    np.testing.assert_allclose(fs.data, want, rtol=1e-12, atol=1e-300)
    # the device sums are compensated: the exactly rounded values, to rounding
    np.testing.assert_allclose(fs.data, oracle.from_phi_exact(phi, ns, xx), rtol=1e-15,
                               atol=1e-300)
    assert fs.mask[(0,) * D] and fs.mask[tuple(ns)]
    assert fs.extrap_x == xx[1]


def test_split_mig_extrapolated_spectrum_and_ll(dadi_amd, oracle):
    """End to end on the configuration's model shape at small grids:
    extrapolated spectrum within 1e-10 relative (entries above 1e-14 of the
    mass) and Poisson ll within 1e-9 absolute of the oracle (its exactly
    rounded sampling sums)."""
    params = (1.7, 0.35, 0.21, 1.3)
    ns = (12, 12)
    pts_l = (30, 36, 42)
    func = dadi_amd.Numerics.make_extrap_log_func(dadi_amd.Demographics2D.split_mig)
    model = func(params, ns, pts_l)
    res, ref, xs = [], [], []
    for pts in pts_l:
        xx = oracle.default_grid(pts)
        phi = oracle.phi_1D_to_2D(xx, oracle.phi_1D(xx))
        phi = oracle.integrate_const(phi, xx, params[2], [params[0], params[1]],
                                     [[params[3]], [params[3]]], [0, 0], [0.5, 0.5])
        res.append(oracle.from_phi_exact(phi, ns, xx))
        ref.append(oracle.from_phi(phi, ns, xx))
        xs.append(xx[1])
    want = oracle.extrapolate(res, xs, log=True)
    mask = oracle.corner_mask(want.shape)
    big = (want > 1e-14 * want[~mask].sum()) & ~mask
    rel = np.abs(model.data - want)[big] / np.abs(want[big])
    assert rel.max() <= 1e-10
    # numpy's summation order (dadi's own) differs from the exactly rounded
    # sums by its rounding, amplified by the extrapolation
    want_ref = oracle.extrapolate(ref, xs, log=True)
    assert (np.abs(model.data - want_ref)[big] / np.abs(want_ref[big])).max() <= 2e-9
    rng = np.random.default_rng(1234)
    data = rng.poisson(1000 * want)
    data_fs = dadi_amd.Spectrum(data)
    ll_gpu = dadi_amd.Inference.ll(1000 * model, data_fs)
    ll_ref = oracle.ll(1000 * want, data)
    assert abs(ll_gpu - ll_ref) <= 1e-9


@pytest.mark.parametrize("chunk,recompute,gg,slots", [(0, 0, 0, 1), (8, 0, 0, 1), (8, 1, 0, 1),
                                                     (0, 0, 1, 1), (8, 0, 0, 2), (8, 0, 0, 3)])
@pytest.mark.parametrize("waves", [1, 2])
def test_batched_variants_bit_exact_on_golden(dadi_amd, chunk, recompute, gg, slots, waves):
    """Every golden case in one batch per (D, order), with the full gamma
    column in LDS (chunk 0), checkpointed gammas (chunk 8, with u' stored or
    recomputed), gamma columns in a global workspace, and 2 or 3 instances
    per workgroup (packed kernel), at 1 and 2 waves per workgroup: identical
    bits to the reference kernels."""
    from dadi_amd import engine
    ctx = engine.context()
    jobs, want = [], []
    for i, c in enumerate(G.meta()):
        if c["D"] < 2:
            continue
        _, phi0, w = G.case(i)
        xx = G.grid_of(c)
        for which in (["tv", "const"] if "const" in w else ["tv"]):
            seg, order, fz, inj = G.segments(c, xx, which)
            jobs.append(engine.Job(ctx.to_dev(phi0), xx, seg, order, fz, inj))
            want.append(w[which])
    if slots >= 2:
        engine.run_batch(jobs, waves=waves, slots=slots)
    else:
        engine.run_batch(jobs, waves=waves, chunk=chunk, recompute=recompute, gamma_global=gg,
                         slots=1)
    for j, w in zip(jobs, want):
        assert np.array_equal(j.phi.cpu().numpy(), w)


@pytest.mark.parametrize("D,n,order", [(2, 40, 1), (2, 100, 1), (2, 45, 0), (3, 12, 1)])
def test_packed_slots_refill_and_quanta(dadi_amd, oracle, D, n, order):
    """The packed kernel (several instances per workgroup in lockstep, each on
    its own schedule, slots refilled from the queue) on row-padded phi gives
    the unpacked kernel's bits on dense phi, also when run in quanta of a few
    steps; spot checks against the oracle.  Some instances have a frozen population, step counts
    differ by two orders of magnitude."""
    import torch
    from dadi_amd import engine, schedule as S, _lib
    from dadi_amd.Misc import ensure_1arg_func as f1
    ctx = engine.context()
    rng = np.random.default_rng(100 * D + n)
    xx = oracle.default_grid(n)
    phi0 = oracle.phi_1D(xx)
    for _ in range(D - 1):
        phi0 = oracle.phi_1D_to_2D(xx, phi0) if phi0.ndim == 1 else \
            oracle.phi_2D_to_3D_split_2(xx, phi0)
    specs = []
    for i in range(37):
        nus = list(np.exp(rng.uniform(np.log(0.2), np.log(5), D)))
        migs = [list(rng.uniform(0, 3, D - 1)) for _ in range(D)]
        gammas = list(rng.uniform(-3, 3, D))
        T = float(np.exp(rng.uniform(np.log(0.002), np.log(0.3))))
        frozen = [False] * D
        if i % 5 == 2:
            frozen[0] = True
            for k in range(D):
                migs[k] = [0.0] * (D - 1)
        fz, inj = S.masks(D, frozen, [False] * D)
        if order == _lib.ORDER_TV:
            seg = S.tv_segments(D, xx, 0.0, T, [f1(v) for v in nus],
                                [[f1(m) for m in ms] for ms in migs], [f1(g) for g in gammas],
                                [f1(0.5)] * D, f1(1.0), None)
        else:
            dt = min(S.compute_dt(nus[k], migs[k], gammas[k], 0.5) for k in range(D))
            seg = S.const_segments(D, xx, 0.0, T, dt, nus, migs, gammas, [0.5] * D, 1.0)
        specs.append((seg, fz, inj, (T, nus, migs, gammas, frozen)))

    def batch(padded=True):
        jobs = []
        for seg, fz, inj, _ in specs:
            if padded:   # rows of the last axis padded to a multiple of 8
                phi = ctx.empty_phi(phi0.shape)
                phi.copy_(torch.from_numpy(phi0))
                assert engine.pitch_of(phi) == engine.row_pitch(n) or n % 8 == 0
            else:
                phi = ctx.to_dev(phi0)
            jobs.append(engine.Job(phi, xx, seg, order, fz, inj))
        return jobs

    plain = batch(padded=False)
    engine.run_batch(plain, slots=1)
    want = [j.phi.cpu().numpy() for j in plain]
    plan = ctx.pack_plan(D, n)
    assert plan is not None
    W, Smax = plan[0], max(2, plan[1])
    for slots in sorted({2, Smax}):
        packed = batch()
        engine.run_batch(packed, waves=W, slots=slots)
        for i, (j, w) in enumerate(zip(packed, want)):
            assert np.array_equal(j.phi.cpu().numpy(), w), (slots, i)
    # quanta: at most 23 steps per instance per launch until every one is done
    jobs = batch()
    pb = engine.PackedBatch(ctx, jobs)
    for _ in range(100000):
        pb.launch(waves=W, slots=Smax, max_steps=23)
        torch.cuda.synchronize()
        if np.all(pb.progress()["status"] == 1):
            break
    for i, (j, w) in enumerate(zip(jobs, want)):
        assert np.array_equal(j.phi.cpu().numpy(), w), ("quanta", i)
    # the unpacked result itself against the oracle, one free and one frozen
    if order == _lib.ORDER_PRE:
        for i in (0, 2):
            T, nus, migs, gammas, frozen = specs[i][3]
            ref = oracle.integrate_const(phi0, xx, T, nus, migs, gammas, [0.5] * D,
                                         frozen=frozen)
            assert np.array_equal(want[i], ref), i


@pytest.mark.parametrize("D,n,ns,het", [(1, 60, (12,), None), (1, 60, (12,), "xx"),
                                        (2, 30, (8, 10), "yy"), (2, 25, (6, 6), None),
                                        (3, 14, (4, 5, 6), "zz"), (4, 10, (3, 4, 3, 2), "xx")])
def test_direct_sampler_matches_oracle(dadi_amd, oracle, D, n, ns, het):
    """het_ascertained / force_direct sampling on the device against the
    oracle's restatement of dadi's trapezoid samplers; grids differ between
    axes where the reference allows it."""
    rng = np.random.default_rng(D * 100 + n)
    xxs = [oracle.default_grid(n + (k if D <= 2 else 0)) for k in range(D)]
    phi = rng.uniform(0.1, 2.0, tuple(len(x) for x in xxs))
    want = oracle.from_phi_direct(phi, ns, xxs, het)
    fs = dadi_amd.Spectrum.from_phi(phi, ns, xxs, het_ascertained=het, force_direct=True)
    assert np.allclose(fs.data, want, rtol=1e-12, atol=0)


@pytest.mark.parametrize("nu,gamma,h,beta,alpha,T", [(1.0, 0.0, 0.5, 1.0, 1.0, 0.1),
                                                     (0.5, -4.0, 0.3, 2.0, 1.5, 0.2),
                                                     (2.5, 6.0, 0.9, 0.5, 0.7, 0.05)])
def test_one_pop_x_bit_exact(dadi_amd, oracle, nu, gamma, h, beta, alpha, T):
    """X-chromosome integration (one_pop_X) through the public API and through
    the batched lane kernel, bit for bit against the oracle's restatement of
    dadi/Integration.py:1256-1352."""
    from dadi_amd import engine, schedule as S, _lib
    xx = oracle.default_grid(80)
    phi0 = oracle.phi_1D(xx)
    want = oracle.one_pop_X(phi0, xx, T, nu, gamma, h, beta, alpha)
    got = dadi_amd.Integration.one_pop_X(phi0, xx, T, nu, gamma, h, beta, alpha)
    assert np.array_equal(got, want)
    ctx = engine.context()
    dt = S.compute_dt(nu, [0], gamma, h)
    seg = S.xchrom_segments(xx, 0.0, T, dt, nu, gamma, h, beta, alpha, 1.0)
    jobs = [engine.Job(ctx.to_dev(phi0), xx, seg, _lib.ORDER_PRE1, 0, 1) for _ in range(3)]
    engine.run_batch(jobs)
    for j in jobs:
        assert np.array_equal(j.phi.cpu().numpy(), want)


def test_delj_trick_matches_reference(dadi_amd, oracle):
    """use_delj_trick (Chang and Cooper's delta per interval) on the device
    against the reference paths: constant 1-3D parameters through the numpy
    coefficient arrays, time-varying and 4D through the reference C sweeps,
    one_pop_X through its restatement.  The device's exp differs from
    glibc's and numpy's in the last bit now and then, so the bar is 1e-10
    relative to the largest entry; batched and single results are identical."""
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built")
    from dadi_amd import engine
    I = dadi_amd.Integration
    tol = lambda want: dict(rtol=1e-10, atol=1e-10 * np.abs(want).max())
    I.use_delj_trick = True
    try:
        xx = oracle.default_grid(36)
        p1 = oracle.phi_1D(xx)
        want = oracle.integrate_const(p1, xx, 0.1, [0.4], [[]], [-6.0], [0.2], beta=1.7,
                                      backend="ref", delj=True)
        got = I.one_pop(p1, xx, 0.1, nu=0.4, gamma=-6.0, h=0.2, beta=1.7)
        assert np.allclose(got, want, **tol(want))
        p2 = oracle.phi_1D_to_2D(xx, p1)
        kw = (p2, xx, 0.05, [0.3, 2.0], [[4.0], [1.5]], [-8.0, 5.0], [0.3, 0.7])
        want2 = oracle.integrate_const(*kw, backend="ref", delj=True)
        got2 = I.two_pops(p2, xx, 0.05, 0.3, 2.0, m12=4.0, m21=1.5, gamma1=-8.0, gamma2=5.0,
                          h1=0.3, h2=0.7)
        assert np.allclose(got2, want2, **tol(want2))
        nu1 = lambda t: 0.3 + 2 * t
        want3 = oracle.integrate_tv(p2, xx, 0.05, [nu1, 2.0], [[4.0], [1.5]], [-8.0, 5.0],
                                    [0.3, 0.7], backend="ref", delj=True)
        got3 = I.two_pops(p2, xx, 0.05, nu1, 2.0, m12=4.0, m21=1.5, gamma1=-8.0, gamma2=5.0,
                          h1=0.3, h2=0.7)
        assert np.allclose(got3, want3, **tol(want3))
        x3 = oracle.default_grid(14)
        p3 = oracle.phi_2D_to_3D_split_2(x3, oracle.phi_1D_to_2D(x3, oracle.phi_1D(x3)))
        want4 = oracle.integrate_const(p3, x3, 0.02, [1.0, 0.5, 2.0], [[1, 2], [3, 0.5], [1, 1]],
                                       [2.0, -3.0, 0.0], [0.5, 0.5, 0.5], backend="ref", delj=True)
        got4 = I.three_pops(p3, x3, 0.02, 1.0, 0.5, 2.0, m12=1, m13=2, m21=3, m23=0.5, m31=1,
                            m32=1, gamma1=2.0, gamma2=-3.0)
        assert np.allclose(got4, want4, **tol(want4))
        x4 = oracle.default_grid(10)
        p4 = oracle.phi_3D_to_4D(oracle.phi_2D_to_3D_split_2(
            x4, oracle.phi_1D_to_2D(x4, oracle.phi_1D(x4))), 0, 0, x4, x4, x4, x4)
        want5 = oracle.integrate_tv(p4, x4, 0.01, [1.0, 0.7, 1.5, 2.0], [[1, 0, 2]] * 4,
                                    [1.0, -2.0, 0.0, 3.0], [0.5] * 4, backend="ref", delj=True)
        got5 = p4.copy()
        I.four_pops(got5, x4, 0.01, 1.0, 0.7, 1.5, 2.0, m12=1, m13=0, m14=2, m21=1, m23=0,
                    m24=2, m31=1, m32=0, m34=2, m41=1, m42=0, m43=2, gamma1=1.0, gamma2=-2.0,
                    gamma4=3.0)
        assert np.allclose(got5, want5, **tol(want5))
        wantx = oracle.one_pop_X(p1, xx, 0.1, 0.6, -5.0, 0.3, 2.0, 1.5, delj=True)
        gotx = I.one_pop_X(p1, xx, 0.1, 0.6, -5.0, 0.3, 2.0, 1.5)
        assert np.allclose(gotx, wantx, **tol(wantx))
        # batched (lane kernel for 1D, workgroup kernel for 2D) equals single
        res = engine.map_batched(
            lambda k: (I.one_pop(p1, xx, 0.1, nu=0.4, gamma=-6.0, h=0.2, beta=1.7) if k == 0
                       else I.two_pops(p2, xx, 0.05, 0.3, 2.0, m12=4.0, m21=1.5, gamma1=-8.0,
                                       gamma2=5.0, h1=0.3, h2=0.7)), [(0,), (1,), (0,), (1,)])
        assert np.array_equal(res[0], got) and np.array_equal(res[2], got)
        assert np.array_equal(res[1], got2) and np.array_equal(res[3], got2)
    finally:
        I.use_delj_trick = False


@pytest.mark.parametrize("D,n", [(1, 300), (2, 40), (3, 14)])
def test_time_budget_quanta_bit_exact(dadi_amd, oracle, D, n):
    """A launch with a wall-clock budget (dh_launch.max_us) stops taking
    steps once the budget is spent and records its progress; resuming until
    done gives the oracle's bits, whatever the number of launches."""
    from dadi_amd import engine, schedule as S, _lib
    ctx = engine.context()
    xx = oracle.default_grid(n)
    phi0 = oracle.phi_1D(xx)
    if D >= 2:
        phi0 = oracle.phi_1D_to_2D(xx, phi0)
    if D == 3:
        phi0 = oracle.phi_2D_to_3D_split_2(xx, phi0)
    nus = [0.7, 1.6, 0.9][:D]
    migs = [[1.2] * (D - 1) for _ in range(D)]
    gam = [-2.0, 1.0, 0.0][:D]
    T = 0.05
    dt = min(S.compute_dt(nus[k], migs[k] if D > 1 else [0], gam[k], 0.5) for k in range(D))
    seg = S.const_segments(D, xx, 0.0, T, dt, nus, migs if D > 1 else [[]], gam, [0.5] * D, 1.0)
    want = oracle.integrate_const(phi0, xx, T, nus, migs if D > 1 else [[]], gam, [0.5] * D)
    order = _lib.ORDER_PRE1 if D == 1 else _lib.ORDER_PRE
    import torch
    phis = [ctx.to_dev(phi0) for _ in range(3)]
    segd = ctx.to_dev(seg.view(np.uint8))
    xxd = ctx.grid(xx)
    fz, inj = S.masks(D, [False] * D, [False] * D)
    tab = ctx.to_dev(np.concatenate([engine._inst_record(
        p.data_ptr(), xxd.data_ptr(), segd.data_ptr(), len(seg), D, n, order, fz, inj,
        engine.pitch_of(p)) for p in phis]).view(np.uint8))
    counter = torch.zeros(1, dtype=torch.int32, device=ctx.device)
    launches = 0
    while True:
        if D == 1:
            cfg = _lib.DhLaunch(1, 0, 0, 0, 0)
            cfg.max_us = 40
            _lib.check(ctx.lib.dh_integrate_1d_wave(tab.data_ptr(), 3, order, n,
                                                     counter.data_ptr(), cfg, ctx.stream()),
                       "dh_integrate_1d_wave")
        else:
            ctx.nd_launch(tab.data_ptr(), 3, D, order, n, counter.data_ptr(), ctx.stream(), {},
                          max_us=40)
        launches += 1
        st = tab.cpu().numpy().view(_lib.INST_DTYPE)
        if np.all(st["status"] == 1) or launches > 100000:
            break
    assert launches > 2, launches
    for p in phis:
        assert np.array_equal(p.cpu().numpy(), want)


@pytest.mark.parametrize("D,n,order,frozen0", [(2, 30, "pre", False), (2, 23, "tv", False),
                                               (2, 40, "pre", True), (3, 13, "pre", False),
                                               (3, 11, "tv", False), (4, 8, "tv", False),
                                               (5, 6, "tv", False)])
def test_factored_kernel_matches_oracle(dadi_amd, oracle, D, n, order, frozen0):
    """The latency kernel for 2-5 populations (Thomas factorisation formed
    once per segment, divisions through known reciprocals with a per-line
    fallback) gives the oracle's bits: constant and time-varying schedules,
    strong selection (densities that underflow: the fallback rows), a frozen
    population, several instances per workgroup in turn, several launches
    (max_steps quanta)."""
    from dadi_amd import engine, schedule as S, _lib
    import torch
    ctx = engine.context()
    xx = oracle.default_grid(n)
    p = oracle.phi_1D(xx, gamma=-400.0)   # subnormal and zero tails
    for d in range(2, D + 1):
        p = oracle.phi_1D_to_2D(xx, p) if d == 2 else \
            oracle.phi_2D_to_3D_split_2(xx, p) if d == 3 else \
            oracle.phi_3D_to_4D(p, 0, 0, xx, xx, xx, xx) if d == 4 else \
            oracle.phi_4D_to_5D(p, 0, 0, 0, xx, xx, xx, xx, xx)
    T = 0.03
    nus = [0.8, 1.7, 0.6, 1.2, 2.0][:D]
    migs = [[0.9 + 0.1 * k] * (D - 1) for k in range(D)]
    gams = [-400.0, 2.0, -5.0, 0.0, 1.0][:D]
    hs = [0.3, 0.5, 0.7, 0.5, 0.2][:D]
    frozen = [frozen0] + [False] * (D - 1)
    if frozen0:
        migs[0] = [0.0] * (D - 1)
    fz, inj = S.masks(D, frozen, [False] * D)
    if order == "pre":
        dt = min(S.compute_dt(nus[k], migs[k], gams[k], hs[k]) for k in range(D))
        seg = S.const_segments(D, xx, 0.0, T, dt, nus, migs, gams, hs, 1.0)
        ordc = _lib.ORDER_PRE if D <= 3 else _lib.ORDER_TV
        want = oracle.integrate_const(p, xx, T, nus, migs, gams, hs, frozen=frozen)
    else:
        c = lambda v: (lambda t: v)
        nu0 = lambda t: 0.8 + 10 * t
        seg = S.tv_segments(D, xx, 0.0, T, [nu0] + [c(v) for v in nus[1:]],
                            [[c(m) for m in ms] for ms in migs], [c(g) for g in gams],
                            [c(h) for h in hs], c(1.0), None)
        ordc = _lib.ORDER_TV
        want = oracle.integrate_tv(p, xx, T, [nu0] + nus[1:], migs, gams, hs, frozen=frozen)
    phis = [ctx.to_dev(p) for _ in range(3)]
    segd = ctx.to_dev(seg.view(np.uint8))
    xxd = ctx.grid(xx)
    tab = ctx.to_dev(np.concatenate([engine._inst_record(
        q.data_ptr(), xxd.data_ptr(), segd.data_ptr(), len(seg), D, n, ordc, fz, inj,
        engine.pitch_of(q)) for q in phis]).view(np.uint8))
    counter = torch.zeros(1, dtype=torch.int32, device=ctx.device)
    ws = {}
    for _ in range(10000):
        ctx.factored_launch(tab.data_ptr(), 3, D, ordc, n, counter.data_ptr(), ctx.stream(), ws,
                            groups=2, max_steps=7)
        if np.all(tab.cpu().numpy().view(_lib.INST_DTYPE)["status"] == 1):
            break
    for q in phis:
        got = q.cpu().numpy()
        assert np.array_equal(got, want), np.max(np.abs(got - want))


def test_spread_batch_matches_oracle(dadi_amd, oracle):
    """Small batches of 3-5 populations run with each integration's sweeps
    spread over several workgroups (dh_integrate_batch, lockstep steps from a
    captured graph): every integration gets the oracle's bits, whatever its
    step count, frozen populations included, and the same bits as the
    one-workgroup kernel."""
    from dadi_amd import engine, schedule as S, _lib
    ctx = engine.context()
    c = lambda v: (lambda t: v)
    specs = []
    for D, n, T, tv, frozen0 in [(3, 13, 0.02, False, False), (3, 11, 0.05, False, False),
                                 (3, 12, 0.03, True, False), (4, 8, 0.02, True, False),
                                 (4, 8, 0.01, True, True), (5, 6, 0.01, True, False)]:
        xx = oracle.default_grid(n)
        p = oracle.phi_1D(xx)
        for d in range(2, D + 1):
            p = oracle.phi_1D_to_2D(xx, p) if d == 2 else \
                oracle.phi_2D_to_3D_split_2(xx, p) if d == 3 else \
                oracle.phi_3D_to_4D(p, 0, 0, xx, xx, xx, xx) if d == 4 else \
                oracle.phi_4D_to_5D(p, 0, 0, 0, xx, xx, xx, xx, xx)
        nus = [0.8, 1.7, 0.6, 1.2, 2.0][:D]
        migs = [[0.9 + 0.2 * k] * (D - 1) for k in range(D)]
        gams = [-3.0, 2.0, -5.0, 0.0, 1.0][:D]
        hs = [0.3, 0.5, 0.7, 0.5, 0.2][:D]
        frozen = [frozen0] + [False] * (D - 1)
        if frozen0:
            migs[0] = [0.0] * (D - 1)
        fz, inj = S.masks(D, frozen, [False] * D)
        if tv:
            nu0 = lambda t: 0.8 + 10 * t
            seg = S.tv_segments(D, xx, 0.0, T, [nu0] + [c(v) for v in nus[1:]],
                                [[c(m) for m in ms] for ms in migs], [c(g) for g in gams],
                                [c(h) for h in hs], c(1.0), None)
            want = oracle.integrate_tv(p, xx, T, [nu0] + nus[1:], migs, gams, hs, frozen=frozen)
            order = _lib.ORDER_TV
        else:
            dt = min(S.compute_dt(nus[k], migs[k], gams[k], hs[k]) for k in range(D))
            seg = S.const_segments(D, xx, 0.0, T, dt, nus, migs, gams, hs, 1.0)
            want = oracle.integrate_const(p, xx, T, nus, migs, gams, hs, frozen=frozen)
            order = _lib.ORDER_PRE
        specs.append((p, xx, seg, order, fz, inj, want))
    old = ctx.spread_max_inst
    try:
        for spread in (512, 0):
            ctx.spread_max_inst = spread
            jobs = [engine.Job(ctx.to_dev(p), xx, seg, order, fz, inj)
                    for p, xx, seg, order, fz, inj, _ in specs]
            engine.run_batch(jobs)
            for k, (j, sp) in enumerate(zip(jobs, specs)):
                assert np.array_equal(j.phi.cpu().numpy(), sp[-1]), (spread, k)
    finally:
        ctx.spread_max_inst = old


def test_large_factored_batch_equals_packed_kernel(dadi_amd, oracle):
    """A 2D batch of eight integrations per compute unit on a small grid runs
    on the factored kernel with two workgroups per compute unit (the
    automatic choice); every result equals the packed kernel's and, spot
    checked, the oracle's."""
    import torch
    from dadi_amd import engine, schedule as S, _lib
    ctx = engine.context()
    B, n = 8 * ctx.num_cu, 20
    assert B <= ctx.factored_max_inst
    slice_d = int(ctx.lib.dh_factored_slice_doubles(2, n))
    assert ctx.factored_groups(B, slice_d) == 2 * ctx.num_cu
    rng = np.random.default_rng(77)
    xx = oracle.default_grid(n)
    phi0 = oracle.phi_1D_to_2D(xx, oracle.phi_1D(xx))
    specs = []
    for _ in range(B):
        nus = list(np.exp(rng.uniform(np.log(0.2), np.log(5), 2)))
        m = float(rng.uniform(0, 3))
        T = float(np.exp(rng.uniform(np.log(0.01), np.log(0.2))))
        dt = min(S.compute_dt(v, [m], 0, 0.5) for v in nus)
        specs.append((S.const_segments(2, xx, 0.0, T, dt, nus, [[m], [m]], [0, 0], [0.5, 0.5],
                                       1.0), (T, nus, m)))
    fz, inj = S.masks(2, [False, False], [False, False])

    def jobs():
        arena = ctx.empty_phi((n, n), batch=B)
        arena.copy_(torch.from_numpy(phi0).to(ctx.device).expand(B, n, n))
        return arena, [engine.Job(arena[i], xx, specs[i][0], _lib.ORDER_PRE, fz, inj)
                       for i in range(B)]

    a_fz, j_fz = jobs()
    pb = engine.PackedBatch(ctx, j_fz)
    assert pb.latency_bound()
    pb.launch()
    a_pk, j_pk = jobs()
    engine.run_batch(j_pk, slots=2)   # an explicit shape: the packed kernel
    torch.cuda.synchronize()
    got, want = a_fz.cpu().numpy(), a_pk.cpu().numpy()
    assert np.array_equal(got, want)
    for i in (0, B // 2, B - 1):
        T, nus, m = specs[i][1]
        ref = oracle.integrate_const(phi0, xx, T, nus, [[m], [m]], [0, 0], [0.5, 0.5])
        assert np.array_equal(got[i], ref), i


@pytest.mark.parametrize("D,n,ns", [(1, 50, (12,)), (2, 30, (8, 11)), (3, 12, (5, 4, 6))])
def test_batched_sampling_equals_single(dadi_amd, D, n, ns):
    """Phis sampled together (a leading batch axis, as the coalescer does for
    concurrent model evaluations) give each phi the bits it gets alone, and
    the coalescer's host-array path gives the bits of the device path."""
    from dadi_amd import Spectrum_mod, engine
    rng = np.random.default_rng(D * 100 + n)
    xx = dadi_amd.Numerics.default_grid(n)
    phis = rng.random((5,) + (n,) * D)
    alone = [Spectrum_mod.sample_device(p, ns, xx).cpu().numpy() for p in phis]
    together = Spectrum_mod.sample_device(phis, ns, xx, batched=True).cpu().numpy()
    for a, b in zip(alone, together):
        assert np.array_equal(a, b)
    got = engine.map_batched(lambda p: dadi_amd.Spectrum.from_phi(p, ns, [xx] * D).data,
                             [(p,) for p in phis], workers=5)
    for a, b in zip(alone, got):
        assert np.array_equal(a, b)


def test_large_spread_batch_uses_checkpoints_and_matches(dadi_amd, oracle):
    """A 3-population batch whose spread launch needs more workgroups than
    there are compute units runs the checkpointed sweep; every integration
    equals the one-workgroup kernel's, and spot checks equal the oracle."""
    import torch
    from dadi_amd import engine, schedule as S, _lib
    ctx = engine.context()
    n, B = 24, 128
    xx = oracle.default_grid(n)
    p = oracle.phi_2D_to_3D_split_2(xx, oracle.phi_1D_to_2D(xx, oracle.phi_1D(xx)))
    rng = np.random.default_rng(31)
    specs = []
    for _ in range(B):
        nus = list(np.exp(rng.uniform(np.log(0.3), np.log(3), 3)))
        migs = [list(rng.uniform(0, 2, 2)) for _ in range(3)]
        T = float(rng.uniform(0.005, 0.03))
        dt = min(S.compute_dt(nus[k], migs[k], 0, 0.5) for k in range(3))
        specs.append((S.const_segments(3, xx, 0.0, T, dt, nus, migs, [0] * 3, [0.5] * 3, 1.0),
                      (T, nus, migs)))
    fz, inj = S.masks(3, [False] * 3, [False] * 3)
    tiles = B * -(-n * n // 64)
    assert tiles > 2 * ctx.num_cu          # four waves per workgroup
    assert -(-tiles // 4) > ctx.num_cu      # more workgroups than compute units
    old = ctx.spread_max_inst
    try:
        out = {}
        for spread in (512, 0):
            ctx.spread_max_inst = spread
            jobs = [engine.Job(ctx.to_dev(p), xx, sp[0], _lib.ORDER_PRE, fz, inj) for sp in specs]
            engine.run_batch(jobs)
            torch.cuda.synchronize()
            out[spread] = [j.phi.cpu().numpy() for j in jobs]
        for a, b in zip(out[512], out[0]):
            assert np.array_equal(a, b)
        for i in (0, B - 1):
            T, nus, migs = specs[i][1]
            ref = oracle.integrate_const(p, xx, T, nus, migs, [0] * 3, [0.5] * 3)
            assert np.array_equal(out[512][i], ref), i
    finally:
        ctx.spread_max_inst = old


@pytest.mark.parametrize("order", ["pre", "tv"])
def test_mixed_factored_batch_equals_workgroup_kernel(dadi_amd, oracle, order):
    """A 2D batch of 2048 integrations on grids of 20-31 points (mixed in one
    launch), with frozen populations and populations without new mutations
    mixed in, takes the factored kernel automatically; every result equals
    the one-workgroup kernel's."""
    import torch
    from dadi_amd import engine, schedule as S, _lib
    from dadi_amd.Misc import ensure_1arg_func as f1
    ctx = engine.context()
    B = 2048
    assert B <= ctx.factored_max_inst
    rng = np.random.default_rng(2048 + (order == "tv"))
    grids = {n: oracle.default_grid(n) for n in range(20, 32)}
    phis = {n: oracle.phi_1D_to_2D(xx, oracle.phi_1D(xx)) for n, xx in grids.items()}
    specs = []
    for i in range(B):
        n = int(rng.integers(20, 32))
        xx = grids[n]
        nus = list(np.exp(rng.uniform(np.log(0.3), np.log(3), 2)))
        migs = [[float(rng.uniform(0, 2))], [float(rng.uniform(0, 2))]]
        gams = list(rng.uniform(-4, 4, 2))
        frozen = [bool(i % 7 == 3), False]
        nomut = [False, bool(i % 5 == 1)]
        if frozen[0]:
            migs = [[0.0], [0.0]]
        T = float(rng.uniform(0.01, 0.1))
        fz, inj = S.masks(2, frozen, nomut)
        if order == "tv":
            seg = S.tv_segments(2, xx, 0.0, T, [f1(v) for v in nus],
                                [[f1(m) for m in ms] for ms in migs], [f1(g) for g in gams],
                                [f1(0.5)] * 2, f1(1.0), None)
            o = _lib.ORDER_TV
        else:
            dt = min(S.compute_dt(nus[k], migs[k], gams[k], 0.5) for k in range(2))
            seg = S.const_segments(2, xx, 0.0, T, dt, nus, migs, gams, [0.5, 0.5], 1.0)
            o = _lib.ORDER_PRE
        specs.append((n, xx, seg, o, fz, inj))

    def run(**kw):
        jobs = [engine.Job(ctx.to_dev(phis[n]), xx, seg, o, fz, inj)
                for n, xx, seg, o, fz, inj in specs]
        engine.run_batch(jobs, **kw)
        torch.cuda.synchronize()
        return jobs

    auto = run()
    pb = engine.PackedBatch(ctx, [j for j in auto])
    assert pb.latency_bound()
    ref = run(slots=1)
    for i, (a, b) in enumerate(zip(auto, ref)):
        assert np.array_equal(a.phi.cpu().numpy(), b.phi.cpu().numpy()), i
