"""
GPU tests of optimiser-side batching with a real model: BFGS in
log(params), a grid search and Godambe uncertainties evaluate their points in
batches (integrations coalesced into shared launches) and return exactly
what the one-evaluation-at-a-time path returns.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    import dadi_amd as d
    assert d.cuda_enabled(True), "HIP backend failed to initialise"
    func_ex = d.Numerics.make_extrap_log_func(d.Demographics2D.split_mig)
    pts = [12, 16, 20]
    ns = (8, 8)
    model = func_ex([1.5, 0.7, 0.3, 1.2], ns, pts)
    rng = np.random.default_rng(11)
    data = d.Spectrum(rng.poisson(2000 * np.maximum(model.data, 0)).astype(float))
    data.mask = model.mask
    return d, func_ex, pts, data


def test_optimize_log_batched_equals_serial(setup):
    d, func_ex, pts, data = setup
    kw = dict(lower_bound=[1e-2, 1e-2, 1e-3, 0], upper_bound=[20, 20, 5, 10], maxiter=3,
              full_output=True)
    p0 = [1.0, 1.0, 0.5, 1.0]
    ref = d.Inference.optimize_log(p0, data, func_ex, pts, batch=False, **kw)
    got = d.Inference.optimize_log(p0, data, func_ex, pts, batch=True, **kw)
    assert np.array_equal(ref[0], got[0]) and ref[1] == got[1]
    cache = d.Inference.last_model_cache
    assert cache.batched >= 5 and cache.serial == 0


def test_grid_and_godambe_batched_equal_serial(setup):
    d, func_ex, pts, data = setup
    grid = d.Inference.index_exp[1.0:2.0:3j, 0.5:1.0:3j]
    fixed = [None, None, 0.3, 1.2]
    ref = d.Inference.optimize_grid(data, func_ex, pts, grid, fixed_params=fixed,
                                    full_output=True, batch=False)
    got = d.Inference.optimize_grid(data, func_ex, pts, grid, fixed_params=fixed,
                                    full_output=True, batch=True)
    assert np.array_equal(ref[0], got[0]) and np.array_equal(ref[3], got[3])
    popt = [1.5, 0.7, 0.3, 1.2]
    d.Godambe.cache.clear()
    a = d.Godambe.FIM_uncert(func_ex, pts, popt, data, batch=False)
    d.Godambe.cache.clear()
    b = d.Godambe.FIM_uncert(func_ex, pts, popt, data, batch=True)
    assert np.array_equal(a, b) and np.all(np.isfinite(a))


def test_optimize_log_batched_agrees_with_oracle_model(setup, oracle):
    """The batched GPU optimisation against the same BFGS driven by an
    oracle-backed split_mig (the oracle's integration and sampling, evaluated
    one point at a time on the CPU): the spectra differ only by the sampling's
    summation order (~1e-13 relative), so the optimum agrees to 1e-6 relative
    and the log-likelihood to 1e-6 absolute (dadi/Inference.py:80-176 for the
    driver)."""
    d, func_ex, pts, data = setup

    def oracle_split_mig(params, ns, pts_):
        nu1, nu2, T, m = params
        xx = oracle.default_grid(pts_)
        phi = oracle.phi_1D_to_2D(xx, oracle.phi_1D(xx))
        phi = oracle.integrate_const(phi, xx, T, [nu1, nu2], [[m], [m]], [0, 0], [0.5, 0.5])
        fs = d.Spectrum(oracle.from_phi(phi, ns, xx))
        fs.extrap_x = xx[1]
        return fs

    oracle_ex = d.Numerics.make_extrap_log_func(oracle_split_mig)
    kw = dict(lower_bound=[1e-2, 1e-2, 1e-3, 0], upper_bound=[20, 20, 5, 10], maxiter=3,
              full_output=True)
    p0 = [1.0, 1.0, 0.5, 1.0]
    want = d.Inference.optimize_log(p0, data, oracle_ex, pts, batch=False, **kw)
    got = d.Inference.optimize_log(p0, data, func_ex, pts, batch=True, **kw)
    np.testing.assert_allclose(got[0], want[0], rtol=1e-6)
    assert abs(got[1] - want[1]) <= 1e-6
    # and the optimum moved away from the start, so the comparison has teeth
    assert np.max(np.abs(np.log(got[0]) - np.log(p0))) > 1e-2
