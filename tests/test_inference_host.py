"""
CPU tests of the optimisation and uncertainty layer (no GPU): the restated
finite-difference formulas against closed-form derivatives, the predicted
finite-difference stencil against the points scipy's own optimisers
evaluate, and that batched prefetching leaves an optimiser's trajectory
unchanged (with a closed-form stand-in for the model spectrum).
"""
import numpy as np
import pytest
import scipy.optimize
from scipy.optimize import _numdiff

from dadi_amd import Godambe, Inference
from dadi_amd.Spectrum_mod import Spectrum


def _f(p):
    return 3.0 * p[0] ** 2 * p[1] + np.exp(p[1]) - 0.5 * p[0] * p[2] ** 3


def _grad(p):
    return np.array([6 * p[0] * p[1] - 0.5 * p[2] ** 3,
                     3 * p[0] ** 2 + np.exp(p[1]),
                     -1.5 * p[0] * p[2] ** 2])


def _hess(p):
    return np.array([[6 * p[1], 6 * p[0], -1.5 * p[2] ** 2],
                     [6 * p[0], np.exp(p[1]), 0.0],
                     [-1.5 * p[2] ** 2, 0.0, -3 * p[0] * p[2]]])


@pytest.mark.parametrize("p0", [[0.7, 1.3, -0.4], [1e-9, 0.5, 2.0], [0.0, -0.2, 0.9]])
def test_finite_difference_formulas_match_closed_forms(p0):
    """Two-sided, one-sided (parameters below 1e-6/eps) and zero-parameter
    branches of get_hess/get_grad (dadi/Godambe.py:11-177) agree with the
    analytic derivatives to the finite-difference error."""
    p0 = np.array(p0)
    eps = 1e-4
    H = Godambe.get_hess(lambda p: _f(p), p0, eps)
    g = Godambe.get_grad(lambda p: _f(p), p0, eps)[:, 0]
    assert np.allclose(H, _hess(p0), rtol=2e-3, atol=2e-3)
    assert np.allclose(g, _grad(p0), rtol=1e-3, atol=1e-3)
    assert np.array_equal(H, H.T)
    # the comparison has teeth: a term missing from the function shows
    bad = Godambe.get_hess(lambda p: _f(p) + 0.5 * p[0] ** 2 * p[2], p0, eps)
    assert not np.allclose(bad, _hess(p0), rtol=2e-3, atol=2e-3)


def test_predicted_stencil_is_what_scipy_evaluates():
    """The prefetched points are, bit for bit, the points scipy's 2-point
    finite differences (the gradient of fmin_bfgs / fmin_l_bfgs_b) visit."""
    rng = np.random.default_rng(3)
    for _ in range(50):
        x0 = rng.normal(size=5) * 10 ** rng.uniform(-3, 3, 5)
        seen = []
        _numdiff.approx_derivative(lambda x: (seen.append(x.tobytes()), 0.0)[1], x0,
                                   method='2-point', abs_step=1e-3)
        want = {p.tobytes() for p in Inference._fd_points(x0, 1e-3)}
        assert set(seen) <= want


class _FakeModel:
    """Closed-form stand-in for a model spectrum (a 1D spectrum whose shape
    depends smoothly on two parameters); counts evaluations per thread."""

    def __init__(self):
        self.calls = 0

    def __call__(self, params, ns, pts):
        self.calls += 1
        nu, T = params
        k = np.arange(ns[0] + 1, dtype=float)
        fs = Spectrum(np.exp(-k * T / nu) / (1.0 + k) * 100)
        fs.mask[0] = fs.mask[-1] = True
        return fs


def _data():
    return Spectrum(np.rint(_FakeModel()((1.7, 0.6), (20,), None)), mask_corners=True)


@pytest.mark.parametrize("opt", ["optimize_log", "optimize", "optimize_log_lbfgsb"])
def test_batched_optimisers_follow_the_serial_trajectory(opt):
    """With prefetching the optimiser returns exactly the serial result and
    asks the model for (almost) nothing beyond the batches."""
    data = _data()
    p0 = [1.0, 1.0]
    kw = dict(lower_bound=[1e-2, 1e-3], upper_bound=[10, 5], maxiter=15, full_output=True)
    if opt == "optimize_log_lbfgsb":
        kw["maxiter"] = 200
    serial_model = _FakeModel()
    ref = getattr(Inference, opt)(p0, data, serial_model, None, batch=False, **kw)
    batched_model = _FakeModel()
    got = getattr(Inference, opt)(p0, data, batched_model, None, batch=True, workers=8, **kw)
    assert np.array_equal(ref[0], got[0])
    assert ref[1] == got[1]
    cache = Inference.last_model_cache
    assert cache.batched > 0
    if opt == "optimize_log_lbfgsb":   # bounded steps may turn back: a few misses
        assert cache.serial <= 0.1 * (cache.batched + cache.serial)
    else:
        assert cache.serial == 0


def test_batched_grid_search_matches_serial():
    data = _data()
    grid = Inference.index_exp[0.5:3:6j, 0.1:1.5:5j]
    ref = Inference.optimize_grid(data, _FakeModel(), None, grid, full_output=True, batch=False)
    got = Inference.optimize_grid(data, _FakeModel(), None, grid, full_output=True, batch=True)
    assert np.array_equal(ref[0], got[0])
    assert np.array_equal(ref[3], got[3]) and np.array_equal(ref[4], got[4])


def test_godambe_batched_equals_serial():
    data = _data()
    rng = np.random.default_rng(5)
    boots = [Spectrum(rng.poisson(data.data).astype(float), mask_corners=True)
             for _ in range(4)]
    fm = _FakeModel()
    func = lambda p, ns, pts: fm(p, ns, pts)
    Godambe.cache.clear()
    ref = Godambe.GIM_uncert(func, [10], boots, [1.6, 0.55], data, batch=False)
    Godambe.cache.clear()
    got = Godambe.GIM_uncert(func, [10], boots, [1.6, 0.55], data, batch=True)
    assert np.array_equal(ref, got)
    Godambe.cache.clear()
    a = Godambe.FIM_uncert(func, [10], [1.6, 0.55], data, log=True, batch=False)
    Godambe.cache.clear()
    b = Godambe.FIM_uncert(func, [10], [1.6, 0.55], data, log=True, batch=True)
    assert np.array_equal(a, b)


def test_param_projection_round_trip():
    fixed = [None, 0.3, None, 2.0]
    p = np.array([1.5, 0.7])
    up = Inference._project_params_up(p, fixed)
    assert list(up) == [1.5, 0.3, 0.7, 2.0]
    assert list(Inference._project_params_down(up, fixed)) == [1.5, 0.7]


@pytest.mark.parametrize("params", [(np.log(5), 2.0, np.log(5), 2.0),
                                    (np.log(4), 1.8, np.log(7), 2.2)])
def test_cache2d_integrate_accounts_for_all_mass(params):
    """Cache2D.integrate (dadi/DFE/Cache2D_mod.py:209-310) with every cached
    spectrum equal to one returns the mass of the density it integrates: the
    trapezoid over the cached square, the four strips outside it (scipy quad
    of each marginal beyond the grid, assigned to the edge spectra) and three
    of the four outer corners (dblquad) -- the corner beyond the grid in both
    coordinates is left out, as in the reference.  The density is a product of
    two lognormals times 1e4 (integrate is linear in it; the scale keeps the
    reference's absolute quadrature tolerance of 1e-4 small against every
    weight), with 8-19% of each marginal's mass on either side of a grid on
    [0.5, 50]: every strip is >= 7% and every corner >= 0.9% of the total,
    so a missing term (or one integrated over the wrong range) exceeds the
    5e-3 tolerance.  Symmetric and asymmetric densities take different code
    paths."""
    from scipy.stats import lognorm
    from dadi_amd.DFE.Cache2D_mod import Cache2D
    mu1, s1, mu2, s2 = params
    d1 = lognorm(s=s1, scale=np.exp(mu1))
    d2 = lognorm(s=s2, scale=np.exp(mu2))
    K = 1e4

    def sel_dist(x, y, p):
        x, y = np.asarray(x, dtype=float), np.asarray(y, dtype=float)
        if x.ndim == 1 and y.ndim == 1:
            return K * d1.pdf(x)[:, None] * d2.pdf(y)[None, :]
        return K * d1.pdf(x) * d2.pdf(y)

    lo, hi, N = 0.5, 50., 50
    c = Cache2D.__new__(Cache2D)
    c.neg_gammas = -np.logspace(np.log10(hi), np.log10(lo), N)
    c.gammas = c.neg_gammas
    c.spectra = np.ones((N, N, 2, 2))
    for p in (d1.cdf(lo), d1.sf(hi), d2.cdf(lo), d2.sf(hi)):
        assert p > 0.08
    fs = c.integrate(params, None, sel_dist, 1.0, None)
    np.testing.assert_allclose(np.asarray(fs), K * (1 - d1.sf(hi) * d2.sf(hi)), rtol=5e-3)
    # without the exterior terms only the cached square's mass remains
    inner = c.integrate(params, None, sel_dist, 1.0, None, exterior_int=False)
    sq = (d1.cdf(hi) - d1.cdf(lo)) * (d2.cdf(hi) - d2.cdf(lo))
    np.testing.assert_allclose(np.asarray(inner), K * sq, rtol=5e-3)


@pytest.mark.parametrize("attr", ["folded", "folded_ancestral", "folded_major"])
def test_folded_data_with_unfolded_model_raises(attr):
    """The reference folds the model to match folded data
    (dadi/Inference.py:493-501, 676-683); without folding support the
    likelihoods and the theta scaling refuse instead of comparing mismatched
    spectra.  Folded model with folded data, and unfolded with unfolded, work."""
    import numpy as np
    import dadi_amd
    rng = np.random.default_rng(0)
    model = dadi_amd.Spectrum(rng.uniform(1, 5, (6, 6)))
    data = dadi_amd.Spectrum(rng.poisson(3.0, (6, 6)).astype(float))
    setattr(data, attr, True)
    for f in (dadi_amd.Inference.ll, dadi_amd.Inference.ll_multinom,
              dadi_amd.Inference.optimal_sfs_scaling, dadi_amd.Inference.ll_per_bin):
        with pytest.raises(NotImplementedError):
            f(model, data)
    setattr(model, attr, True)
    assert np.isfinite(dadi_amd.Inference.ll(model, data))
