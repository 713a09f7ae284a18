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
        # the more selective population of a pair sets its time step, so the
        # number of steps grows with max |gamma|: pairs go out most selective
        # first (launch rounds of similar integrations)
        cost = [max(abs(self.gammas[i]), abs(self.gammas[j])) for i, j in mine]
        specs = _evaluate_grid(demo_sel_func, sets, ns, pts, workers, cost)
        self.spectra = [[None] * ng for _ in range(ng)]
        for (i, j), s in zip(mine, specs):
            self.spectra[i][j] = s.data if hasattr(s, 'data') else s
        if split_jobs == 1:
            self.spectra = np.array(self.spectra)
        if verbose:
            print('cached %d spectra' % len(mine))

    @staticmethod
    def merge(caches):
        """Combine caches generated with split_jobs (dadi/DFE/Cache2D_mod.py:469-505):
        entries two caches both hold must agree, and the result must be complete
        (ValueError otherwise).  The first cache is copied, not modified."""
        import copy
        out = copy.deepcopy(caches[0])
        ng = len(out.gammas)
        for c in caches[1:]:
            for i in range(ng):
                for j in range(ng):
                    fs = c.spectra[i][j]
                    if fs is None:
                        continue
                    have = out.spectra[i][j]
                    if have is not None and not np.all(np.asarray(have) == np.asarray(fs)):
                        raise ValueError("Merged cached conflicts with current.")
                    out.spectra[i][j] = fs
        for i in range(ng):
            for j in range(ng):
                if out.spectra[i][j] is None:
                    raise ValueError("Cache is incomplete after merging. First missing entry "
                                     "is {0},{1}.".format(i, j))
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

    def integrate_point_pos(self, params, ns, biv_seldist, theta, rho=0, pts=None):
        """Spectrum of a DFE made of four quadrants (dadi/DFE/Cache2D_mod.py:312-430):
        both gammas negative (integrate() of the bivariate density biv_seldist,
        exterior terms included, at theta 1), both positive (the cached
        pair (gammapos1, gammapos2)), and one of each (the cached row or column of
        the positive gamma weighted by the other population's marginal density).
        params ends with (ppos1, gammapos1, ppos2, gammapos2); the quadrant
        weights are p1 p2 + rho (sqrt(p1 p2) - p1 p2), (1-rho) p1 (1-p2),
        (1-rho) (1-p1) p2 and (1-p1)(1-p2) + rho (1 - sqrt(p1 p2) - (1-p1)(1-p2)).
        No normalisation: mass outside the cached gammas is taken as unseen.
        Raises IndexError when a positive gamma is not among the cached ones."""
        biv_params = params[:-4]
        ppos1, gammapos1, ppos2, gammapos2 = params[-4:]
        Nneg = len(self.neg_gammas)
        weights = biv_seldist(-self.neg_gammas, -self.neg_gammas, biv_params)
        neg_neg = self.integrate(biv_params, ns, biv_seldist, 1, pts)
        i1 = np.flatnonzero(self.gammas == gammapos1)
        i2 = np.flatnonzero(self.gammas == gammapos2)
        if len(i1) == 0 or len(i2) == 0:
            raise IndexError('Failed to find requested gammapos1={0:.4f} and/or '
                             'gammapos2={1:0.4f} in cached spectra. Were they included in '
                             'additional_gammas during cache generation?'
                             .format(gammapos1, gammapos2))
        i1, i2 = int(i1[0]), int(i2[0])
        pos_pos = self.spectra[i1, i2]
        # population 1 positive: its cached row, over population 2's marginal
        marg2 = _trapz(weights, self.neg_gammas, axis=0)
        pos_neg = _trapz(marg2[:, None, None] * self.spectra[i1, :Nneg], self.neg_gammas,
                         axis=0)
        # population 2 positive: its cached column, over population 1's marginal
        marg1 = _trapz(weights, self.neg_gammas, axis=1)
        neg_pos = _trapz(marg1[:, None, None] * self.spectra[:Nneg, i2], self.neg_gammas,
                         axis=0)
        both = ppos1 * ppos2
        w_pp = both + rho * (np.sqrt(both) - both)
        w_pn = (1 - rho) * ppos1 * (1 - ppos2)
        w_np = (1 - rho) * (1 - ppos1) * ppos2
        w_nn = (1 - ppos1) * (1 - ppos2) + rho * (1 - np.sqrt(both) - (1 - ppos1) * (1 - ppos2))
        fs = w_pp * pos_pos + w_pn * pos_neg + w_np * neg_pos + w_nn * neg_neg
        return theta * fs

    def integrate_symmetric_point_pos(self, params, ns, biv_seldist, theta, pts=None):
        """integrate_point_pos with the same positive point mass (ppos, gammapos)
        in both populations; the last density parameter is the correlation rho
        (dadi/DFE/Cache2D_mod.py:432-467)."""
        dens = list(params[:-2])
        ppos, gammapos = params[-2:]
        return self.integrate_point_pos(dens + [ppos, gammapos, ppos, gammapos], ns,
                                        biv_seldist, theta, rho=dens[-1], pts=None)


def mixture(params, ns, s1, s2, sel_dist1, sel_dist2, theta, pts, exterior_int=True):
    """(1 - p2d) x the Cache1D spectrum (perfectly correlated gammas) + p2d x the
    Cache2D spectrum; params = shared density parameters, rho, p2d
    (dadi/DFE/Cache2D_mod.py:509-547)."""
    fs1 = s1.integrate(params[:-2], None, sel_dist1, theta, None, exterior_int)
    fs2 = s2.integrate(params[:-1], None, sel_dist2, theta, None, exterior_int)
    p2d = params[-1]
    return (1 - p2d) * fs1 + p2d * fs2


def mixture_symmetric_point_pos(params, ns, s1, s2, sel_dist1, sel_dist2, theta, pts=None):
    """mixture with a symmetric positive point mass; params = shared density
    parameters, rho, ppos, gammapos, p2d (dadi/DFE/Cache2D_mod.py:549-587).
    The Cache2D part receives (density parameters, rho, ppos, gammapos, ppos,
    gammapos), exactly the arguments the reference passes, so its results are
    the reference's."""
    dens = list(params[:-4])
    rho, ppos, gammapos, p2d = params[-4:]
    fs1 = s1.integrate_point_pos(dens + [ppos, gammapos], None, sel_dist1, theta, Npos=1,
                                 pts=None)
    fs2 = s2.integrate_symmetric_point_pos(dens + [rho, ppos, gammapos, ppos, gammapos], None,
                                           sel_dist2, theta, None)
    return (1 - p2d) * fs1 + p2d * fs2


def mixture_point_pos(params, ns, s1, s2, sel_dist1, sel_dist2, theta, pts=None):
    """mixture with positive point masses (ppos1, gammapos1) and (ppos2,
    gammapos2); params = shared density parameters, rho, ppos1, gammapos1,
    ppos2, gammapos2, p2d (dadi/DFE/Cache2D_mod.py:589-631)."""
    dens = list(params[:-6])
    rho, ppos1, gammapos1, ppos2, gammapos2, p2d = params[-6:]
    fs1 = s1.integrate_point_pos(dens + [ppos1, gammapos1], None, sel_dist1, theta, Npos=1,
                                 pts=None)
    fs2 = s2.integrate_point_pos(dens + [rho, ppos1, gammapos1, ppos2, gammapos2], None,
                                 sel_dist2, theta, None)
    return (1 - p2d) * fs1 + p2d * fs2
