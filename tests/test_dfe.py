"""
DFE cache generation on the batched GPU engine, checked against the
reference's stored fitdadi spectra (tests/test_DFE.py:68-93 of the
reference: Cache1D of IM_sel_single_gamma, ns (8, 12), pts (60, 80, 100),
gamma_bounds (1e-2, 10), 100 gammas, lognormal DFE; np.allclose there).
"""
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _read_fs(path):
    lines = [l for l in open(path).read().split("\n") if l and not l.startswith("#")]
    shape = tuple(int(v) for v in lines[0].split()[:-1])
    data = np.array([float(v) for v in lines[1].split()]).reshape(shape)
    mask = np.array([int(v) for v in lines[2].split()]).reshape(shape).astype(bool)
    return np.ma.masked_array(data, mask)


@pytest.fixture(scope="module")
def cache1d():
    import dadi_amd
    assert dadi_amd.cuda_enabled(True)
    from dadi_amd.DFE import DemogSelModels, Cache1D
    return Cache1D([0.5, 2, 0.5, 0.1, 0, 0], [8, 12], DemogSelModels.IM_sel_single_gamma,
                   [60, 80, 100], gamma_bounds=(1e-2, 10), gamma_pts=100)


@pytest.mark.gpu
def test_cache1d_matches_fitdadi_golden(cache1d):
    from dadi_amd.DFE import PDFs
    fs = cache1d.integrate([-0.5, 0.5], None, PDFs.lognormal, 1e5, None, exterior_int=False)
    comp = _read_fs(os.path.join(HERE, "golden", "fitdadi.IM_no_ext_test.fs"))
    assert np.ma.allclose(fs, comp, rtol=1e-5)
    fs = cache1d.integrate([-0.5, 0.5], None, PDFs.lognormal, 1e5, None, exterior_int=True)
    comp = _read_fs(os.path.join(HERE, "golden", "fitdadi.IM_test.fs"))
    assert np.ma.allclose(fs, comp, rtol=1e-5)


@pytest.mark.gpu
def test_cache_entries_equal_sequential_evaluations(cache1d):
    """Batched (coalesced) cache entries are bit-identical to evaluating the
    same model one call at a time."""
    import dadi_amd
    from dadi_amd.DFE import DemogSelModels
    ex = dadi_amd.Numerics.make_extrap_func(DemogSelModels.IM_sel_single_gamma)
    for ii in (0, 37, 99):
        g = cache1d.gammas[ii]
        fs = ex((0.5, 2, 0.5, 0.1, 0, 0, g), (8, 12), (60, 80, 100))
        assert np.array_equal(np.asarray(fs.data), cache1d.spectra[ii])


@pytest.mark.gpu
def test_cache1d_point_pos_matches_fitdadi_golden(cache1d):
    """Reference tests/test_DFE.py:87-92: a positive point mass (ppos 0.1,
    gammapos 4.3) beside the lognormal DFE; gammapos is not cached, so it is
    evaluated with IM_sel_single_gamma and appended (times theta, as the
    reference stores it)."""
    import copy
    from dadi_amd.DFE import DemogSelModels, PDFs
    c = copy.deepcopy(cache1d)
    fs = c.integrate_point_pos([-0.5, 0.5, 0.1, 4.3], None, PDFs.lognormal, 1e5,
                               DemogSelModels.IM_sel_single_gamma)
    comp = _read_fs(os.path.join(HERE, "golden", "fitdadi.IM_point_pos_test.fs"))
    assert np.ma.allclose(fs, comp, rtol=1e-5)
    assert 4.3 in c.gammas
    # without a model function an uncached positive gamma is an error
    with pytest.raises(IndexError):
        copy.deepcopy(cache1d).integrate_point_pos([-0.5, 0.5, 0.1, 4.3], None,
                                                   PDFs.lognormal, 1e5)


@pytest.mark.gpu
def test_cache2d_point_pos_matches_fitdadi_golden():
    """Reference tests/test_DFE.py:178-193: Cache2D of IM_sel, ns (8, 12),
    pts (60, 80, 100), 5 negative gammas in (1e-2, 10) plus 4.3; symmetric
    positive point mass over a bivariate lognormal DFE."""
    import dadi_amd
    assert dadi_amd.cuda_enabled(True)
    from dadi_amd.DFE import DemogSelModels, PDFs, Cache2D
    s2 = Cache2D([0.5, 2, 0.5, 0.1, 0, 0], [8, 12], DemogSelModels.IM_sel, pts=[60, 80, 100],
                 gamma_pts=5, gamma_bounds=(1e-2, 10), additional_gammas=[4.3])
    fs = s2.integrate_symmetric_point_pos([-0.5, 0.5, 0.5, 0.1, 4.3], None, PDFs.biv_lognormal,
                                          1e5)
    comp = _read_fs(os.path.join(HERE, "golden", "fitdadi.IM_2D_point_pos_test.fs"))
    assert np.ma.allclose(fs, comp, rtol=1e-5)


@pytest.mark.gpu
def test_cache2d_point_pos_needs_cached_gammapos():
    """Reference tests/test_DFE.py:149-165: the point mass may only use a
    cached positive gamma."""
    import dadi_amd
    assert dadi_amd.cuda_enabled(True)
    from dadi_amd.DFE import DemogSelModels, PDFs, Cache2D
    s2 = Cache2D([0.5, 2, 0.5, 0.01, 0, 0], [3, 3], DemogSelModels.IM_sel, [20],
                 gamma_bounds=(1e-4, 2), gamma_pts=2, additional_gammas=[0.2])
    fs = s2.integrate([2, 1, 0.4], None, PDFs.biv_lognormal, 1, None)
    assert np.all(np.isfinite(np.asarray(fs)))
    s2.integrate_symmetric_point_pos([2, 1, 0.4, 0.1, 0.2], None, PDFs.biv_lognormal, 1, None)
    with pytest.raises(IndexError):
        s2.integrate_symmetric_point_pos([2, 1, 0.4, 0.1, 2], None, PDFs.biv_lognormal, 1, None)


def _hand_cache(spectra, gammas):
    """A Cache2D object with given entries (no generation), for host-side tests."""
    from dadi_amd.DFE import Cache2D
    c = Cache2D.__new__(Cache2D)
    c.gammas = np.asarray(gammas, dtype=float)
    c.neg_gammas = c.gammas
    c.spectra = spectra
    return c


def test_cache2d_merge_checks_conflicts_and_completeness():
    """Cache2D.merge (dadi/DFE/Cache2D_mod.py:469-505): disjoint split_jobs
    parts merge into the full cache; conflicting entries and an incomplete
    merge raise ValueError; the first part is not modified."""
    from dadi_amd.DFE import Cache2D
    g = [-2.0, -1.0]
    full = [[np.full((3, 3), 10 * i + j, dtype=float) for j in range(2)] for i in range(2)]
    parts = []
    for job in range(3):
        sp = [[None] * 2 for _ in range(2)]
        for k, (i, j) in enumerate([(i, j) for i in range(2) for j in range(2)]):
            if k % 3 == job:
                sp[i][j] = full[i][j]
        parts.append(_hand_cache(sp, g))
    merged = Cache2D.merge(parts)
    assert np.array_equal(merged.spectra, np.array(full))
    assert parts[0].spectra[0][1] is None
    with pytest.raises(ValueError):
        Cache2D.merge(parts[:2])
    bad = [[None] * 2 for _ in range(2)]
    bad[0][0] = full[0][0] + 1
    with pytest.raises(ValueError):
        Cache2D.merge([parts[0], _hand_cache(bad, g), parts[1], parts[2]])


def test_cache2d_point_pos_quadrant_weights():
    """integrate_point_pos on a hand-made cache whose entries are constant
    spectra: the four quadrant weights of dadi/DFE/Cache2D_mod.py:412-429 add
    up as the formula says; at rho = 0 the result is a plain mixture."""
    from dadi_amd.DFE import PDFs
    neg = -np.logspace(1, -1, 6)
    gam = np.concatenate([neg, [3.0]])
    sp = np.zeros((7, 7, 2, 2))
    sp[-1, -1] = 7.0       # both positive
    sp[-1, :6] = 5.0       # population 1 positive
    sp[:6, -1] = 3.0       # population 2 positive
    sp[:6, :6] = 1.0       # both negative
    c = _hand_cache(sp, gam)
    c.neg_gammas = neg
    dens = [0.2, 0.6, 0.0]     # mu, sigma, rho
    w = PDFs.biv_lognormal(-neg, -neg, dens)
    trap = lambda v, axis: np.trapezoid(v, neg, axis=axis) if hasattr(np, "trapezoid") \
        else np.trapz(v, neg, axis=axis)
    marg2 = trap(trap(w, 0)[:, None, None] * np.ones((6, 2, 2)), 0)
    marg1 = trap(trap(w, 1)[:, None, None] * np.ones((6, 2, 2)), 0)
    nn = np.asarray(c.integrate(dens, None, PDFs.biv_lognormal, 1))
    p1, p2 = 0.1, 0.3
    got = c.integrate_point_pos(dens + [p1, 3.0, p2, 3.0], None, PDFs.biv_lognormal, 2.0)
    want = 2.0 * (p1 * p2 * 7 + p1 * (1 - p2) * 5 * marg2 + (1 - p1) * p2 * 3 * marg1
                  + (1 - p1) * (1 - p2) * nn)
    keep = ~np.ma.getmaskarray(got)
    assert keep.sum() == 2      # the corners of the spectrum are masked
    np.testing.assert_allclose(np.asarray(got)[keep], want[keep], rtol=1e-12)
    with pytest.raises(IndexError):
        c.integrate_point_pos(dens + [p1, 4.0, p2, 3.0], None, PDFs.biv_lognormal, 2.0)


def test_cache2d_pairs_go_most_selective_first_and_come_back_in_order():
    """Cache2D hands its gamma pairs to the workers most selective first (the
    larger |gamma| of a pair sets its time step, so launch rounds then hold
    integrations of similar length) and stores every spectrum at its own
    (i, j), as the unsorted loop does (host only: the model is a recorder)."""
    from dadi_amd.DFE import Cache2D
    seen = []

    def model(p, ns, pts):
        seen.append(tuple(p[-2:]))
        return np.full(tuple(n + 1 for n in ns), 1000 * p[-2] + p[-1])

    c = Cache2D((1.0,), (2, 2), model, [10], gamma_bounds=(1e-2, 100), gamma_pts=4,
                workers=1)
    g = c.gammas
    key = [max(abs(a), abs(b)) for a, b in seen]
    assert key == sorted(key, reverse=True)
    assert sorted(seen) == sorted((a, b) for a in g for b in g)
    for i in range(len(g)):
        for j in range(len(g)):
            assert np.all(c.spectra[i][j] == 1000 * g[i] + g[j])
