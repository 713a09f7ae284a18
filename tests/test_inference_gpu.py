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
