"""
Two-population DFE cache over pairs of selection coefficients
(dadi/DFE/Cache2D_mod.py).  Generation batches every (gamma1, gamma2, pts)
evaluation on the GPU through engine.map_batched.
"""
import numpy as np
import scipy.integrate

from ..Spectrum_mod import Spectrum
from .Cache1D_mod import _evaluate_grid, _trapz


class Cache2D:
    """Spectra cached over the grid gammas x gammas (dadi/DFE/Cache2D_mod.py:13-92).
    ``split_jobs``/``this_job_id`` keep only every split_jobs-th pair, as the
    reference does for manual sharding; ``merge`` recombines them."""

    def __init__(self, params, ns, demo_sel_func, pts, gamma_bounds=(1e-4, 2000.),
                 gamma_pts=100, additional_gammas=[], cpus=None, gpus=0, verbose=False,
                 split_jobs=1, this_job_id=0, workers=512):
        self.ns, self.pts, self.params = ns, pts, params
        self.func_name = demo_sel_func.__name__
        self.gammas = -np.logspace(np.log10(gamma_bounds[1]), np.log10(gamma_bounds[0]),
                                   gamma_pts)
        self.neg_gammas = self.gammas
        self.gammas = np.concatenate((self.gammas, additional_gammas))
        ng = len(self.gammas)
        pairs = [(i, j) for i in range(ng) for j in range(ng)]
        mine = pairs[this_job_id::split_jobs]
        sets = [tuple(params) + (self.gammas[i], self.gammas[j]) for i, j in mine]
        specs = _evaluate_grid(demo_sel_func, sets, ns, pts, workers)
        self.spectra = [[None] * ng for _ in range(ng)]
        for (i, j), s in zip(mine, specs):
            self.spectra[i][j] = s.data if hasattr(s, 'data') else s
        if split_jobs == 1:
            self.spectra = np.array(self.spectra)
        if verbose:
            print('cached %d spectra' % len(mine))

    @staticmethod
    def merge(caches):
        """Combine caches generated with split_jobs (dadi/DFE/Cache2D_mod.py:469-505)."""
        out = caches[0]
        ng = len(out.gammas)
        for c in caches[1:]:
            for i in range(ng):
                for j in range(ng):
                    if out.spectra[i][j] is None:
                        out.spectra[i][j] = c.spectra[i][j]
        out.spectra = np.array(out.spectra)
        return out

    def integrate(self, params, ns, sel_dist, theta, pts=None, exterior_int=True):
        """Spectrum integrated over a bivariate density sel_dist(xx, yy, params)
        on the negative-gamma grid, plus (exterior_int) the mass outside the
        cached domain assigned to the edge spectra
        (dadi/DFE/Cache2D_mod.py:209-310)."""
        Nneg = len(self.neg_gammas)
        spectra = self.spectra[:Nneg, :Nneg]
        params = np.array(params)
        weights = sel_dist(-self.neg_gammas, -self.neg_gammas, params)
        weighted = weights[:, :, None, None] * spectra
        temp = _trapz(weighted, self.neg_gammas, axis=0)
        fs = _trapz(temp, self.neg_gammas, axis=0)
        if not exterior_int:
            return Spectrum(theta * fs)
        testx = np.logspace(-2, 2, 3)
        testout = sel_dist(testx, testx, params)
        symmetric = np.allclose(testout, testout.T, atol=0, rtol=1e-12)
        max_gamma = -self.neg_gammas[-1]
        min_gamma = -self.neg_gammas[0]
        g = -self.neg_gammas
        w1lo, w1hi, w2lo, w2hi = (np.zeros(Nneg) for _ in range(4))
        for ii, gamma in enumerate(g):
            a, _ = scipy.integrate.quad(sel_dist, min_gamma, np.inf, epsabs=1e-4, epsrel=1e-3,
                                        args=(gamma, params))
            b, _ = scipy.integrate.quad(sel_dist, 0, max_gamma, epsabs=1e-4, epsrel=1e-3,
                                        args=(gamma, params))
            if symmetric:
                c, d = a, b
            else:
                marg = lambda g2: sel_dist(gamma, g2, params)
                c, _ = scipy.integrate.quad(marg, min_gamma, np.inf, epsabs=1e-4, epsrel=1e-3)
                d, _ = scipy.integrate.quad(marg, 0, max_gamma, epsabs=1e-4, epsrel=1e-3)
            w1lo[ii], w1hi[ii], w2lo[ii], w2hi[ii] = a, b, c, d
        fs = fs + _trapz(spectra[:, 0] * w2lo[:, None, None], self.neg_gammas, axis=0)
        fs = fs + _trapz(spectra[:, -1] * w2hi[:, None, None], self.neg_gammas, axis=0)
        fs = fs + _trapz(spectra[0, :] * w1lo[:, None, None], self.neg_gammas, axis=0)
        fs = fs + _trapz(spectra[-1, :] * w1hi[:, None, None], self.neg_gammas, axis=0)
        w, _ = scipy.integrate.dblquad(sel_dist, 0, max_gamma, lambda _: 0,
                                       lambda _: max_gamma, epsrel=1e-3, epsabs=1e-4,
                                       args=[params])
        fs = fs + spectra[-1, -1] * w
        w, _ = scipy.integrate.dblquad(sel_dist, 0, max_gamma, lambda _: min_gamma,
                                       lambda _: np.inf, epsrel=1e-3, epsabs=1e-4, args=[params])
        fs = fs + spectra[0, -1] * w
        if not symmetric:
            w, _ = scipy.integrate.dblquad(sel_dist, min_gamma, np.inf, lambda _: 0,
                                           lambda _: max_gamma, epsrel=1e-3, epsabs=1e-4,
                                           args=[params])
        fs = fs + spectra[-1, 0] * w
        return Spectrum(theta * fs)
