"""
The register-resident 2D kernel (dh_integrate_resident) and its latency form
(dh_integrate_resident_latency, the Thomas factorisation kept per segment):
phi held in registers for a whole integration, lines transposed through LDS
between the sweeps.  Every test runs both forms.  Its results must be the bits of the other kernels and of the oracle:
split_mig draws on the grid sizes it is built for, both coefficient orders
(constant parameters and time-varying ones), frozen populations, integrations
run in quanta of a few steps, and an instance of another shape left alone.
The fast kernel's shared-reciprocal division must hand every integration
whose quotients leave its exact range (tiny or zero phi) to the exactly
dividing kernel, and both must give the oracle's bits.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    import dadi_amd
    assert dadi_amd.cuda_enabled(True)
    from dadi_amd import engine
    return engine


def _draws(rng, k, tmax):
    out = []
    for _ in range(k):
        nu1, nu2 = np.exp(rng.uniform(np.log(0.1), np.log(10), 2))
        out.append((nu1, nu2, rng.uniform(0.001, tmax), rng.uniform(0, 5)))
    return out


def _jobs(eng, oracle, rng, n, k, tv=False, frozen_every=0, tmax=0.02):
    from dadi_amd import schedule as S, _lib
    ctx = eng.context()
    xx = oracle.default_grid(n)
    phi0 = oracle.phi_1D_to_2D(xx, oracle.phi_1D(xx)) + rng.uniform(0, 0.1, (n, n))
    jobs, refs = [], []
    for i, (nu1, nu2, T, m) in enumerate(_draws(rng, k, tmax)):
        fz = [False, False]
        if frozen_every and i % frozen_every == 1:
            fz = [i % 2 == 0, i % 2 == 1]
        migs = [[0.0 if fz[0] else m], [0.0 if fz[1] else m]]
        g = [float(rng.uniform(-3, 3)), float(rng.uniform(-3, 3))]
        h = [0.5, float(rng.uniform(0.1, 0.9))]
        if tv:
            c = lambda v: (lambda t, v=v: v)
            nu1f = (lambda a: (lambda t: a * (1 + 5 * t)))(nu1)
            seg = S.tv_segments(2, xx, 0.0, T, [nu1f, c(nu2)], [[c(v) for v in ms] for ms in migs],
                                [c(v) for v in g], [c(v) for v in h], c(1.0), None)
            order = _lib.ORDER_TV
            ref = oracle.integrate_tv(phi0, xx, T, [nu1f, nu2], migs, g, h, frozen=fz)
        else:
            dt = min(S.compute_dt(nu1, migs[0], g[0], h[0]), S.compute_dt(nu2, migs[1], g[1], h[1]))
            seg = S.const_segments(2, xx, 0.0, T, dt, [nu1, nu2], migs, g, h, 1.0)
            order = _lib.ORDER_PRE
            ref = oracle.integrate_const(phi0, xx, T, [nu1, nu2], migs, g, h, frozen=fz)
        fzm, inj = S.masks(2, fz, [False, False])
        jobs.append(eng.Job(ctx.empty_phi((n, n)), xx, seg, order, fzm, inj))
        jobs[-1].phi.copy_(ctx.to_dev(phi0))
        refs.append(ref)
    return jobs, refs, phi0


@pytest.mark.parametrize("latency", [False, True])
@pytest.mark.parametrize("n", [33, 40, 57, 64, 65, 80, 90, 97, 100, 104])
def test_resident_equals_packed_and_oracle(eng, oracle, n, latency):
    ctx = eng.context()
    assert ctx.lib.dh_resident_rows(n) == (n + 7) // 8 * 8
    rng = np.random.default_rng(n)
    jobs, refs, phi0 = _jobs(eng, oracle, rng, n, 24, frozen_every=5)
    pb = eng.PackedBatch(ctx, jobs)
    assert pb.resident_ok()
    pb.launch(resident=True, latency=latency)
    got = [j.phi.cpu().numpy() for j in jobs]
    for g, r in zip(got, refs):
        assert np.array_equal(g, r)
    # the packed kernel on the same table gives the same bits
    for j in jobs:
        j.phi.copy_(ctx.to_dev(phi0))
    pb.restart()
    pb.launch(resident=False, slots=3)
    for j, g in zip(jobs, got):
        assert np.array_equal(j.phi.cpu().numpy(), g)


@pytest.mark.parametrize("n,tv", [(105, False), (112, False), (119, False), (128, False),
                                  (112, True), (128, True)])
def test_latency_form_above_104_points(eng, oracle, n, tv):
    """The latency form serves 105..128 points (the throughput form stops at
    104), both coefficient orders: each integration gets the oracle's bits
    and the packed kernel's."""
    ctx = eng.context()
    assert ctx.lib.dh_resident_rows(n) == 0
    assert ctx.lib.dh_resident_latency_rows(n) == (n + 7) // 8 * 8
    rng = np.random.default_rng(n + tv)
    jobs, refs, phi0 = _jobs(eng, oracle, rng, n, 6, tv=tv, frozen_every=3,
                             tmax=0.004 if tv else 0.02)
    pb = eng.PackedBatch(ctx, jobs)
    assert pb.resident_ok()
    pb.launch(resident=True)
    got = [j.phi.cpu().numpy() for j in jobs]
    for g, r in zip(got, refs):
        assert np.array_equal(g, r)
    for j in jobs:
        j.phi.copy_(ctx.to_dev(phi0))
    pb.restart()
    pb.launch(resident=False)
    for j, g in zip(jobs, got):
        assert np.array_equal(j.phi.cpu().numpy(), g)


@pytest.mark.parametrize("latency", [False, True])
def test_resident_time_varying_and_quanta(eng, oracle, latency):
    ctx = eng.context()
    rng = np.random.default_rng(3)
    jobs, refs, _ = _jobs(eng, oracle, rng, 100, 6, tv=True, tmax=0.004)
    pb = eng.PackedBatch(ctx, jobs)
    # a few steps per launch, resuming from the recorded progress
    for _ in range(10000):
        pb.launch(resident=True, max_steps=7, latency=latency)
        st = pb.tab.cpu().numpy().view(eng.INST_DTYPE)["status"]
        if np.all(st == 1):
            break
    assert np.all(st == 1)
    for j, r in zip(jobs, refs):
        assert np.array_equal(j.phi.cpu().numpy(), r)


def test_resident_sizes():
    """The grid sizes with a kernel: 33 to 104 points for the throughput
    form, 33 to 128 for the latency form, in the same row classes."""
    import dadi_amd
    assert dadi_amd.cuda_enabled(True)
    from dadi_amd import engine
    ctx = engine.context()
    have = [n for n in range(3, 200) if ctx.lib.dh_resident_rows(n)]
    assert have == list(range(33, 105))
    assert {ctx.lib.dh_resident_rows(n) for n in have} == set(range(40, 105, 8))
    lat = [n for n in range(3, 200) if ctx.lib.dh_resident_latency_rows(n)]
    assert lat == list(range(33, 129))
    assert all(ctx.lib.dh_resident_latency_rows(n) == ctx.lib.dh_resident_rows(n) for n in have)


@pytest.mark.parametrize("latency", [False, True])
def test_resident_mixed_sizes_in_one_kernel(eng, oracle, latency):
    """Instances on different grid sizes served by one kernel (97 to 104
    points) run in one launch, each to the oracle's bits."""
    ctx = eng.context()
    rng = np.random.default_rng(9)
    jobs, refs = [], []
    for n in (97, 100, 104, 101):
        j, r, _ = _jobs(eng, oracle, rng, n, 3)
        jobs += j
        refs += r
    pb = eng.PackedBatch(ctx, jobs)
    assert pb.resident_ok()
    pb.launch(resident=True, latency=latency)
    for j, r in zip(jobs, refs):
        assert np.array_equal(j.phi.cpu().numpy(), r)


@pytest.mark.parametrize("latency", [False, True])
def test_resident_rejects_other_shapes(eng, oracle, latency):
    """An instance whose n the kernel is not built for is left untouched
    with status -1; the others complete."""
    from dadi_amd import _lib
    ctx = eng.context()
    rng = np.random.default_rng(4)
    jobs, refs, phi0 = _jobs(eng, oracle, rng, 100, 3)
    other, _, phi_o = _jobs(eng, oracle, rng, 60, 1)
    pb = eng.PackedBatch(ctx, jobs)
    tab = pb.tab.cpu().numpy().view(eng.INST_DTYPE).copy()
    # point instance 1 at a 60-point phi (its record says n = 60)
    ot = eng.PackedBatch(ctx, other).tab.cpu().numpy().view(eng.INST_DTYPE)
    tab[1] = ot[0]
    pb.tab.copy_(ctx.to_dev(tab.view(np.uint8)))
    pb.launch_resident(latency=latency)
    st = pb.tab.cpu().numpy().view(eng.INST_DTYPE)["status"]
    assert list(st) == [1, -1, 1]
    assert np.array_equal(other[0].phi.cpu().numpy(), phi_o)
    assert np.array_equal(jobs[0].phi.cpu().numpy(), refs[0])
    assert np.array_equal(jobs[2].phi.cpu().numpy(), refs[2])
    assert ctx.lib.dh_resident_rows(60) == 64


def _mark_exact(eng, ctx, pb, which):
    tab = pb.tab.cpu().numpy().view(eng.INST_DTYPE).copy()
    for i in which:
        tab[i]["reserved"] |= 1
    pb.tab.copy_(ctx.to_dev(tab.view(np.uint8)))


@pytest.mark.parametrize("latency", [False, True])
@pytest.mark.parametrize("tv", [False, True])
def test_resident_exact_kernel_equals_fast(eng, oracle, tv, latency):
    """Instances marked for exact division (run by the second kernel only)
    give the bits of the fast kernel and of the oracle."""
    ctx = eng.context()
    rng = np.random.default_rng(17 + tv)
    jobs, refs, phi0 = _jobs(eng, oracle, rng, 90, 8, tv=tv, frozen_every=3, tmax=0.005)
    pb = eng.PackedBatch(ctx, jobs)
    _mark_exact(eng, ctx, pb, range(0, 8, 2))
    pb.launch(resident=True, latency=latency)
    tab = pb.tab.cpu().numpy().view(eng.INST_DTYPE)
    assert list(tab["status"]) == [1] * 8
    assert list(tab["reserved"] & 1) == [1, 0] * 4
    for j, r in zip(jobs, refs):
        assert np.array_equal(j.phi.cpu().numpy(), r)
        assert np.array_equal(np.signbit(j.phi.cpu().numpy()), np.signbit(r))


@pytest.mark.parametrize("latency", [False, True])
@pytest.mark.parametrize("tv", [False, True])
def test_resident_tiny_values_go_to_exact_kernel(eng, oracle, tv, latency):
    """phi with entries far below 2^-800 makes the recurrences' numerators
    leave the fast kernel's exact range: those integrations are redone with
    exact division (the throughput form marks them for its second kernel,
    the latency form redoes them at once in the same workgroup).  Zeros of
    either sign (phi after a
    split is zero off the diagonal) and ordinary values stay on the fast
    kernel, whose quotient gives a zero numerator's sign as division does.
    Every result is the oracle's, zeros' signs included, also when run in
    quanta of a few steps."""
    from dadi_amd import schedule as S, _lib
    ctx = eng.context()
    rng = np.random.default_rng(23 + tv)
    n = 72
    xx = oracle.default_grid(n)
    base = oracle.phi_1D_to_2D(xx, oracle.phi_1D(xx))
    starts = []
    for i in range(8):
        p = base + rng.uniform(0, 0.1, (n, n))
        if i % 4 == 1:
            p[:, n // 2:] *= 1e-290          # tiny quotients
        elif i % 4 == 2:
            p = base.copy()                  # +0 off the diagonal
        elif i % 4 == 3:
            p[:, : n // 2] = -0.0            # -0 numerators
        starts.append(p)
    for quanta in (0, 40):
        jobs, refs = [], []
        for i, (nu1, nu2, T, m) in enumerate(_draws(rng, 8, 0.003)):
            migs = [[m], [m]] if i < 4 else [[0.0], [0.0]]
            g = [0.0, float(rng.uniform(-1, 1))] if i < 4 else [0.0, 0.0]
            h = [0.5, 0.5]
            if tv:
                c = lambda v: (lambda t, v=v: v)
                nu1f = (lambda a: (lambda t: a * (1 + 5 * t)))(nu1)
                seg = S.tv_segments(2, xx, 0.0, T, [nu1f, c(nu2)],
                                    [[c(v) for v in ms] for ms in migs],
                                    [c(v) for v in g], [c(v) for v in h], c(1.0), None)
                order = _lib.ORDER_TV
                ref = oracle.integrate_tv(starts[i], xx, T, [nu1f, nu2], migs, g, h)
            else:
                dt = min(S.compute_dt(nu1, migs[0], g[0], h[0]),
                         S.compute_dt(nu2, migs[1], g[1], h[1]))
                seg = S.const_segments(2, xx, 0.0, T, dt, [nu1, nu2], migs, g, h, 1.0)
                order = _lib.ORDER_PRE
                ref = oracle.integrate_const(starts[i], xx, T, [nu1, nu2], migs, g, h)
            fzm, inj = S.masks(2, [False, False], [False, False])
            jobs.append(eng.Job(ctx.empty_phi((n, n)), xx, seg, order, fzm, inj))
            jobs[-1].phi.copy_(ctx.to_dev(starts[i]))
            refs.append(ref)
        pb = eng.PackedBatch(ctx, jobs)
        for _ in range(10000):
            pb.launch(resident=True, max_steps=quanta, latency=latency)
            tab = pb.tab.cpu().numpy().view(eng.INST_DTYPE)
            if np.all(tab["status"] == 1):
                break
        assert np.all(tab["status"] == 1)
        marked = [int(v) for v in tab["reserved"] & 1]
        tiny = [0, 0] if latency else [1, 1]   # the latency form redoes them at once
        assert marked[0::4] == [0, 0] and marked[1::4] == tiny, marked
        assert marked[2::4] == [0, 0] and marked[3::4] == [0, 0], marked
        for j, r in zip(jobs, refs):
            got = j.phi.cpu().numpy()
            assert np.array_equal(got, r)
            assert np.array_equal(np.signbit(got), np.signbit(r))   # zeros' signs too


@pytest.mark.parametrize("latency", [False, True])
def test_resident_out_of_range_phi_goes_to_exact_kernel(eng, oracle, latency):
    """An instance whose phi holds a value beyond 2^600 when it is taken up
    is integrated with exact division (the fast kernel's checks assume
    finite recurrences), and so is one whose value near 2^531 drives a
    numerator above 2^500 at its first step (marked for the second kernel by
    the throughput form, redone at once by the latency form); their
    neighbours stay on the fast kernel; every result is the oracle's."""
    ctx = eng.context()
    rng = np.random.default_rng(31)
    jobs, refs, phi0 = _jobs(eng, oracle, rng, 80, 4)
    from dadi_amd import schedule as S
    big = phi0.copy()
    big[5, 7] = 1e200
    j = jobs[2]
    nu1, nu2 = 1.5, 0.7
    dt = min(S.compute_dt(nu1, [0.3], 0, 0.5), S.compute_dt(nu2, [0.3], 0, 0.5))
    xx = oracle.default_grid(80)
    j.seg = S.const_segments(2, xx, 0.0, 0.002, dt, [nu1, nu2], [[0.3], [0.3]], [0, 0],
                             [0.5, 0.5], 1.0)
    j.phi.copy_(ctx.to_dev(big))
    refs[2] = oracle.integrate_const(big, xx, 0.002, [nu1, nu2], [[0.3], [0.3]], [0, 0],
                                     [0.5, 0.5])
    large = phi0.copy()
    large[40, 41] = 1e160
    assert 2.0 ** 500 < 1e160 < 2.0 ** 600
    jobs[3].seg = j.seg
    jobs[3].phi.copy_(ctx.to_dev(large))
    refs[3] = oracle.integrate_const(large, xx, 0.002, [nu1, nu2], [[0.3], [0.3]], [0, 0],
                                     [0.5, 0.5])
    pb = eng.PackedBatch(ctx, jobs)
    pb.launch(resident=True, latency=latency)
    tab = pb.tab.cpu().numpy().view(eng.INST_DTYPE)
    assert list(tab["status"]) == [1] * 4
    assert list(tab["reserved"] & 1) == ([0, 0, 0, 0] if latency else [0, 0, 1, 1])
    for jb, r in zip(jobs, refs):
        assert np.array_equal(jb.phi.cpu().numpy(), r)
