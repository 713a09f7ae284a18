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
