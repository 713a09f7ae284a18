"""
CPU pins for the oracle's checker functions that the GPU parity tests lean on
for the 1e-10 spectrum and 1e-9 log-likelihood bars: the Poisson and
multinomial likelihoods (dadi/Inference.py:466-575, 668-685), Richardson
extrapolation with and without logs and its fail_mag fallback
(dadi/Numerics.py:334-501, 644-669), sampling of separable 3-5 population
densities (dadi/Spectrum_mod.py:1774-1871), and the population splits and
removals (dadi/PhiManip.py:214-495, 1036-1072).

Each check is against mathematics or a library routine the restatement does
not share (scipy.stats, exact polynomials, closed-form Beta integrals,
trapezoid mass), and each has a planted mistake that it rejects.
"""
import numpy as np
import pytest
from scipy import stats
from scipy.optimize import minimize_scalar


# ---------------------------------------------------------------------------
# likelihoods
# ---------------------------------------------------------------------------
def _model_and_data(seed, shape):
    rng = np.random.default_rng(seed)
    model = np.exp(rng.uniform(np.log(1e-3), np.log(500), shape))
    data = rng.poisson(model).astype(float)
    data.flat[::7] = 0.0          # empty bins, as real spectra have
    return model, data


@pytest.mark.parametrize("shape", [(21,), (13, 9), (6, 5, 7)])
def test_ll_equals_scipy_poisson_logpmf(oracle, shape):
    """ll = sum over unmasked bins of log P(data | Poisson(model)), with the
    corners masked by default."""
    model, data = _model_and_data(1, shape)
    mask = oracle.corner_mask(shape)
    want = stats.poisson.logpmf(data, model)[~mask].sum()
    got = oracle.ll(model, data)
    assert abs(got - want) <= 1e-9 * max(1.0, abs(want))
    # an explicit mask is honoured
    m2 = mask.copy()
    m2.flat[3] = True
    want2 = stats.poisson.logpmf(data, model)[~m2].sum()
    assert abs(oracle.ll(model, data, m2) - want2) <= 1e-9 * max(1.0, abs(want2))
    # planted mistake: dropping the lnGamma(data + 1) term is caught
    bad = np.sum((-model + data * np.log(model))[~mask])
    assert abs(bad - want) > 1.0


@pytest.mark.parametrize("shape", [(21,), (13, 9)])
def test_ll_multinom_is_ll_at_the_optimal_theta(oracle, shape):
    """ll_multinom maximises the Poisson ll over the model's scale: the
    maximiser found numerically is sum(data)/sum(model) over unmasked bins, the
    value equals ll there, and scaling the model leaves it unchanged."""
    model, data = _model_and_data(2, shape)
    mask = oracle.corner_mask(shape)
    keep = ~mask
    res = minimize_scalar(lambda lt: -stats.poisson.logpmf(data, np.exp(lt) * model)[keep].sum(),
                          bracket=(-2.0, 2.0), tol=1e-12)
    theta_num = np.exp(res.x)
    theta = data[keep].sum() / model[keep].sum()
    assert abs(theta_num / theta - 1) < 1e-6
    want = stats.poisson.logpmf(data, theta * model)[keep].sum()
    got = oracle.ll_multinom(model, data)
    assert abs(got - want) <= 1e-9 * max(1.0, abs(want))
    assert abs(oracle.ll_multinom(37.5 * model, data) - got) <= 1e-9 * max(1.0, abs(got))
    # planted mistake: theta from all bins (masked corners included) differs
    theta_all = data.sum() / model.sum()
    assert abs(stats.poisson.logpmf(data, theta_all * model)[keep].sum() - want) > 1e-6


def test_product_likelihoods_equal_oracle(oracle):
    """The product's host likelihoods (masked-array code) give the oracle's
    numbers, including on a Spectrum with an extra masked bin."""
    import dadi_amd
    model, data = _model_and_data(3, (11, 8))
    fs_m = dadi_amd.Spectrum(model)
    fs_d = dadi_amd.Spectrum(data)
    mask = oracle.corner_mask(model.shape)
    assert abs(dadi_amd.Inference.ll(fs_m, fs_d) - oracle.ll(model, data, mask)) < 1e-9
    assert abs(dadi_amd.Inference.ll_multinom(fs_m, fs_d)
               - oracle.ll_multinom(model, data, mask)) < 1e-9
    fs_d.mask[4, 4] = True
    m2 = mask.copy()
    m2[4, 4] = True
    assert abs(dadi_amd.Inference.ll(fs_m, fs_d) - oracle.ll(model, data, m2)) < 1e-9


# ---------------------------------------------------------------------------
# extrapolation
# ---------------------------------------------------------------------------
def _xs(oracle, pts_l):
    return [oracle.default_grid(p)[1] for p in pts_l]


def test_extrapolation_exact_on_polynomials(oracle):
    """Lagrange extrapolation to x = 0 through 2 (3) grids reproduces the
    intercept of any line (parabola) exactly up to rounding, entry by entry;
    with log=True the same holds for exp of them.  Grid spacings are those of
    the default grids at pts (80, 90, 100)."""
    rng = np.random.default_rng(4)
    a = rng.uniform(-3, 3, (5, 6))
    b = rng.uniform(-50, 50, (5, 6))
    c = rng.uniform(-2000, 2000, (5, 6))
    xs3 = _xs(oracle, (80, 90, 100))
    xs2 = xs3[1:]
    lin = [a + b * x for x in xs2]
    quad = [a + b * x + c * x * x for x in xs3]
    np.testing.assert_allclose(oracle.extrapolate(lin, xs2), a, rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(oracle.extrapolate(quad, xs3), a, rtol=1e-10, atol=1e-11)
    got = oracle.extrapolate([np.exp(q) for q in quad], xs3, log=True)
    np.testing.assert_allclose(got, np.exp(a), rtol=1e-10)
    got = oracle.extrapolate([np.exp(q) for q in lin], xs2, log=True)
    np.testing.assert_allclose(got, np.exp(a), rtol=1e-11)
    # one grid: the value itself
    assert np.array_equal(oracle.extrapolate([quad[0]], xs3[:1]), quad[0])
    # planted mistake: extrapolating the parabola linearly from two grids is off
    assert not np.allclose(oracle.extrapolate(quad[1:], xs2), a, rtol=1e-6)


def test_extrapolation_fail_mag_falls_back_to_finest_grid(oracle):
    """An entry whose extrapolation lands more than fail_mag decades from the
    finest grid's value takes that value; an entry moved by less keeps its
    extrapolation; the finest grid is the one with the smallest x whatever
    order the grids come in; with log=True the fallback is exp of the finest
    grid's log, as dadi/Numerics.py:455-458 forms it (not the log itself)."""
    xs = _xs(oracle, (100, 80, 90))          # finest grid first, on purpose
    w = [oracle.quadratic_extrap(list(np.eye(3)[k]), xs) for k in range(3)]
    assert abs(sum(w) - 1) < 1e-12            # Lagrange weights sum to one
    base, target = 2e-20, 1e-9
    res = [np.full(4, base) for _ in xs]
    # entry 0 extrapolates to `target`, 11 decades above the finest grid's
    # value; the coarsest grid (x of pts 80) holds another value, so a
    # fallback to the wrong grid is visible
    res[1][0] = 5 * base
    res[2][0] = (target - w[0] * base - w[1] * res[1][0]) / w[2]
    # entry 1 moves by a factor of a few (the grid with the largest weight)
    k = int(np.argmax(w))
    res[k][1] = 1.5 * base
    # entry 2 extrapolates to a negative value: on the reference's masked
    # spectra log10 masks the ratio, so it never counts as failed
    res[2][2] = (-target - (w[0] + w[1]) * base) / w[2]
    direct = [oracle.quadratic_extrap([r[j] for r in res], xs) for j in range(4)]
    assert abs(np.log10(direct[0] / base)) > 10 > abs(np.log10(direct[1] / base))
    assert direct[2] < 0
    ex = oracle.extrapolate(res, xs, fail_mag=10)
    assert ex[0] == res[0][0] != res[1][0]     # the finest grid's value
    assert ex[1] == direct[1] and ex[2] == direct[2] and ex[3] == direct[3]
    # the same rule on numpy masked arrays, as the reference evaluates it
    mex = np.ma.masked_array(np.array(direct))
    mbest = np.ma.masked_array(res[0].copy())
    failed = abs(np.log10(mex / mbest)) > 10
    mex[failed] = mbest[failed]
    assert np.array_equal(np.asarray(mex), ex)
    assert oracle.extrapolate(res, xs, fail_mag=12)[0] == direct[0]
    # log extrapolation: shift the log of one grid's entry so that the
    # extrapolated log moves by 30 (13 decades): it falls back to 1e-3
    lres = [np.full(3, 1e-3) for _ in xs]
    lres[2][0] = 1e-3 * np.exp(-30.0 / w[2])
    dl = np.exp(sum(wk * np.log(r[0]) for wk, r in zip(w, lres)))
    assert abs(np.log10(dl / 1e-3)) > 10
    exl = oracle.extrapolate(lres, xs, log=True)
    assert exl[0] == np.exp(np.log(lres[0][0]))
    np.testing.assert_allclose(exl[1:], 1e-3, rtol=1e-12)


def test_product_extrapolation_equals_oracle(oracle):
    """The product's make_extrap_func / make_extrap_log_func give the oracle's
    values on spectra, including entries that fail."""
    import dadi_amd
    xs = _xs(oracle, (40, 50, 60))
    rng = np.random.default_rng(6)
    res = [np.exp(rng.uniform(-30, 0, (7, 7))) for _ in xs]
    res[2][3, 3] = 1e-60
    for log in (False, True):
        want = oracle.extrapolate(res, xs, log=log)
        specs = []
        for r, x in zip(res, xs):
            fs = dadi_amd.Spectrum(r.copy())
            fs.extrap_x = x
            specs.append(fs)
        got = dadi_amd.Numerics.extrapolate(specs, xs, extrap_log=log)
        m = oracle.corner_mask((7, 7))
        assert np.array_equal(np.asarray(got.data)[~m], want[~m])


# ---------------------------------------------------------------------------
# sampling of separable densities in 3-5 populations
# ---------------------------------------------------------------------------
def _linear_expectation(a, b, n):
    k = np.arange(n + 1)
    return a / (n + 1) + b * (k + 1) / ((n + 1) * (n + 2))


@pytest.mark.parametrize("ns,pts", [((4, 6, 9), 23), ((3, 5, 2, 6), 13), ((2, 4, 3, 5, 3), 9)])
def test_sampler_exact_for_separable_linear_phi(oracle, ns, pts):
    """phi = prod_k (a_k + b_k x_k) with distinct a_k, b_k and distinct sample
    sizes: the analytic sampler is exact for each factor, so the spectrum is
    the outer product of the 1D expectations in population order.  An axis
    order slip changes the shape or the values."""
    D = len(ns)
    xx = oracle.default_grid(pts)
    rng = np.random.default_rng(D)
    a = rng.uniform(0.5, 2.0, D)
    b = rng.uniform(-0.4, 0.4, D)
    phi = np.ones((pts,) * D)
    want = np.ones(())
    for k in range(D):
        shp = [1] * D
        shp[k] = pts
        phi = phi * (a[k] + b[k] * xx).reshape(shp)
        want = np.multiply.outer(want, _linear_expectation(a[k], b[k], ns[k]))
    got = oracle.from_phi(phi, list(ns), xx)
    assert got.shape == tuple(n + 1 for n in ns)
    np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-15)
    # planted mistake: sample sizes in reverse population order
    bad = oracle.from_phi(np.transpose(phi), list(ns)[::-1], xx)
    assert bad.shape != want.shape or not np.allclose(bad, want, rtol=1e-6)


# ---------------------------------------------------------------------------
# splits and removals
# ---------------------------------------------------------------------------
def _trap_w(xx):
    dx = np.diff(xx)
    w = np.zeros(len(xx))
    w[:-1] += dx / 2
    w[1:] += dx / 2
    return w


def _random_phi(rng, D, n):
    xx = np.sort(np.concatenate(([0.0, 1.0], rng.uniform(0, 1, n - 2))))
    return xx, rng.uniform(0.1, 2.0, (n,) * D)


@pytest.mark.parametrize("D", [2, 3, 4])
def test_new_axis_preserves_trapezoid_mass(oracle, D):
    """Every admixture split scatters phi at each point onto the two grid
    points bracketing the admixed frequency, normalised so that the trapezoid
    integral over the new axis gives back phi exactly (the reference's
    norm = 2 phi / (f_lo dz0 + dz1 + f_up dz2)), for any proportions and grids;
    with all weight on one population the new axis copies it (the diagonal of
    phi_1D_to_2D)."""
    rng = np.random.default_rng(10 + D)
    n = 11 if D < 4 else 7
    xx, phi = _random_phi(rng, D, n)
    zz = oracle.default_grid(n)
    if D == 2:
        for f in (0.0, 0.37, 1.0):
            out = oracle.phi_2D_to_3D_admix(phi, f, xx, xx, zz)
            np.testing.assert_allclose(oracle.remove_pop(out, zz, 3), phi, rtol=1e-13)
    elif D == 3:
        for f1, f2 in ((0.0, 0.0), (0.2, 0.5), (1.0, 0.0), (0.0, 1.0)):
            out = oracle.phi_3D_to_4D(phi, f1, f2, xx, xx, xx, zz)
            np.testing.assert_allclose(oracle.remove_pop(out, zz, 4), phi, rtol=1e-13)
    else:
        for f in ((0.1, 0.2, 0.3), (0.0, 0.0, 1.0)):
            out = oracle.phi_4D_to_5D(phi, *f, xx, xx, xx, xx, zz)
            np.testing.assert_allclose(oracle.remove_pop(out, zz, 5), phi, rtol=1e-13)
    # copying one population: on the default grid the new axis is the
    # normalised diagonal of that population's axis
    xg = oracle.default_grid(n)
    phi2 = rng.uniform(0.1, 2.0, (n, n))
    for f, axis in ((1.0, 0), (0.0, 1)):
        out = oracle.phi_2D_to_3D_admix(phi2, f, xg, xg, xg)
        for i in range(1, n - 1):
            sl = [slice(None)] * 2
            sl[axis] = i
            want = phi2[tuple(sl)] * 2 / (xg[i + 1] - xg[i - 1])
            got = out[:, :, i][tuple(sl)]
            np.testing.assert_allclose(got, want, rtol=1e-13)
        # planted mistake: the other orientation does not match
        if f == 1.0:
            wrong = oracle.phi_2D_to_3D_admix(phi2, 0.0, xg, xg, xg)
            assert not np.allclose(wrong, out)


def test_splits_give_neutral_marginals(oracle):
    """Reference tests/test_4D.py:4-25: after 1D -> 2D -> 3D -> 4D splits of the
    neutral density, each population's marginal spectrum is the 1D one
    (rtol/atol 1e-3 as there)."""
    pts = 50
    xx = oracle.default_grid(pts)
    phi1 = oracle.phi_1D(xx)
    fs1 = oracle.from_phi(phi1, [10], xx)
    phi = oracle.phi_1D_to_2D(xx, phi1)
    phi = oracle.phi_2D_to_3D_admix(phi, 0, xx, xx, xx)
    phi = oracle.phi_3D_to_4D(phi, 0, 0, xx, xx, xx, xx)
    fs4 = oracle.from_phi(phi, [10] * 4, xx)
    for keep in range(4):
        fsm = oracle.marginalize(fs4, [k for k in range(4) if k != keep])
        assert np.allclose(fs1[1:-1], fsm[1:-1], rtol=1e-3, atol=1e-3)


def test_filter_pops_of_3d_run_equals_lower_runs(oracle):
    """Reference tests/test_PhiManip.py:4-32: integrating populations out of a
    3D run (no migration) gives the 1D runs, and keeping populations 1 and 3 the
    2D run with the same time steps."""
    pts = 9
    nu1, nu2, nu3, T = 0.1, 1, 10, 0.1
    xx = oracle.default_grid(pts)
    zero = [[0, 0]] * 3
    phi = oracle.phi_1D_to_2D(xx, oracle.phi_1D(xx))
    phi = oracle.phi_2D_to_3D_admix(phi, 0, xx, xx, xx)
    phi_all = oracle.integrate_const(phi, xx, T, [nu1, nu2, nu3], zero, [0] * 3, [0.5] * 3)
    phi1 = oracle.integrate_const(oracle.phi_1D(xx), xx, T, [nu1], [[]], [0], [0.5])
    phi2 = oracle.integrate_const(oracle.phi_1D(xx), xx, T, [nu2], [[]], [0], [0.5])
    phi13 = oracle.integrate_const(oracle.phi_1D_to_2D(xx, oracle.phi_1D(xx)), xx, T,
                                   [nu1, nu3], [[0], [0]], [0, 0], [0.5, 0.5])
    assert np.allclose(phi1[1:-1], oracle.filter_pops(phi_all, xx, [1])[1:-1])
    assert np.allclose(phi2[1:-1], oracle.filter_pops(phi_all, xx, [2])[1:-1])
    comp = oracle.filter_pops(phi_all, xx, [1, 3])
    phi13[0, 0] = phi13[-1, -1] = comp[0, 0] = comp[-1, -1] = 0
    assert np.allclose(phi13, comp)
    # planted mistake: keeping population 2 instead of 3 does not match
    wrong = oracle.filter_pops(phi_all, xx, [1, 2])
    wrong[0, 0] = wrong[-1, -1] = 0
    assert not np.allclose(phi13, wrong)


def test_product_host_splits_equal_oracle(oracle):
    """The product's host splits and removals give the oracle's bits on random
    densities, proportions and grids."""
    from dadi_amd import PhiManip as P
    rng = np.random.default_rng(12)
    for n in (7, 12):
        xx = oracle.default_grid(n)
        yy = np.sort(np.concatenate(([0.0, 1.0], rng.uniform(0, 1, n - 2))))
        p1 = rng.uniform(0.1, 2, n)
        assert np.array_equal(P.phi_1D_to_2D(xx, p1), oracle.phi_1D_to_2D(xx, p1))
        p2 = rng.uniform(0.1, 2, (n, n))
        for f in (0.0, 0.3, 1.0):
            assert np.array_equal(P.phi_2D_to_3D_admix(p2, f, xx, yy, xx),
                                  oracle.phi_2D_to_3D_admix(p2, f, xx, yy, xx))
        assert np.array_equal(P.phi_2D_to_3D_split_1(xx, p2), oracle.phi_2D_to_3D_split_1(xx, p2))
        assert np.array_equal(P.phi_2D_to_3D_split_2(xx, p2), oracle.phi_2D_to_3D_split_2(xx, p2))
        p3 = rng.uniform(0.1, 2, (n, n, n))
        for f1, f2 in ((0.0, 0.0), (0.25, 0.4)):
            assert np.array_equal(P.phi_3D_to_4D(p3, f1, f2, xx, yy, xx, yy),
                                  oracle.phi_3D_to_4D(p3, f1, f2, xx, yy, xx, yy))
        p4 = rng.uniform(0.1, 2, (n,) * 4)
        assert np.array_equal(P.phi_4D_to_5D(p4, 0.1, 0.2, 0.3, xx, xx, yy, yy, xx),
                              oracle.phi_4D_to_5D(p4, 0.1, 0.2, 0.3, xx, xx, yy, yy, xx))
        for pop in (1, 2, 3):
            assert np.array_equal(P.remove_pop(p3, xx, pop), oracle.remove_pop(p3, xx, pop))
        assert np.array_equal(P.filter_pops(p4, xx, [2, 4]), oracle.filter_pops(p4, xx, [2, 4]))


# ---------------------------------------------------------------------------
# exactly rounded sampling sums
# ---------------------------------------------------------------------------
def _short_split_mig_phi(oracle, pts, T=0.02):
    xx = oracle.default_grid(pts)
    phi = oracle.phi_1D_to_2D(xx, oracle.phi_1D(xx))
    return oracle.integrate_const(phi, xx, T, [0.16, 1.35], [[2.3], [2.3]], [0, 0],
                                  [0.5, 0.5]), xx


def test_compensated_sampling_equals_exact_rational_sums(oracle):
    """from_phi_exact (compensated dot products) against exact rational
    arithmetic on the same double inputs, contraction by contraction, on a
    split_mig density shortly after the split: its entries next to the
    spectrum's corners cancel by ~4 orders of magnitude.  The compensated
    sums are the exactly rounded values (within one rounding); numpy's BLAS
    order is not, and after extrapolation its error reaches the 1e-10 bar at
    pts 80-100 (here only measured, as the BLAS order depends on the host)."""
    phi, xx = _short_split_mig_phi(oracle, 30)
    ns = (6, 8)
    first = oracle._contract_last_exact(phi, ns[1], xx)            # (30, 9)
    want1 = oracle.contract_fraction(phi, ns[1], xx, range(ns[1] + 1))
    np.testing.assert_allclose(first, want1, rtol=2.3e-16, atol=0)
    mid = np.moveaxis(first, -1, 0)                                   # (9, 30)
    want2 = oracle.contract_fraction(mid, ns[0], xx, range(ns[0] + 1))
    got = oracle.from_phi_exact(phi, ns, xx)                          # (7, 9)
    np.testing.assert_allclose(got, want2.T, rtol=2.3e-16, atol=0)
    # the reference-order sampler agrees to rounding but not exactly
    ref = oracle.from_phi(phi, ns, xx)
    assert np.allclose(ref, got, rtol=1e-11, atol=0)
    # 1D and 3D go through the same contraction
    p1 = phi[:, 5]
    np.testing.assert_allclose(oracle.from_phi_exact(p1, [7], xx),
                               oracle.contract_fraction(p1, 7, xx, range(8))[0],
                               rtol=2.3e-16, atol=0)
    assert np.allclose(oracle.from_phi_exact(p1, [7], xx), oracle.from_phi(p1, [7], xx),
                       rtol=1e-12, atol=0)
    x3 = oracle.default_grid(13)
    p3 = np.random.default_rng(2).uniform(0.1, 2, (13, 13, 13))
    assert np.allclose(oracle.from_phi_exact(p3, [3, 4, 5], x3),
                       oracle.from_phi(p3, [3, 4, 5], x3), rtol=1e-13, atol=0)
    # planted mistake: the slope term without its (k+1)/((n+1)(n+2)) factor
    db1, db2 = oracle.cached_dbeta(ns[1], xx)
    s = (phi[:, 1:] - phi[:, :-1]) / (xx[1:] - xx[:-1])
    c1 = (phi[:, :-1] - s * xx[:-1]) / (ns[1] + 1)
    assert not np.allclose(c1 @ db1.T + s @ db2.T, want1, rtol=1e-3)
