"""
Cache of selected spectra over a grid of selection coefficients, for
inferring a distribution of fitness effects (dadi/DFE/Cache1D_mod.py).

Generation is the expensive part: one extrapolated model evaluation per gamma
(500 x 3 grid sizes by default).  Here every evaluation runs in its own worker
thread against the ordinary Integration API and all of their integrations are
coalesced into batched persistent launches on the GPU
(engine.map_batched), instead of dadi's process pool of serial evaluations
(dadi/DFE/Cache1D_mod.py:96-179).
"""
import numpy as np
import scipy.integrate

from .. import Numerics, engine
from ..Spectrum_mod import Spectrum

try:
    _trapz = np.trapezoid
except AttributeError:  # numpy < 2
    _trapz = np.trapz


def _evaluate_grid(demo_sel_func, param_sets, ns, pts, workers, cost=None):
    """Extrapolated spectra of demo_sel_func for every parameter set: each
    (set, grid size) is one worker task; their integrations share launches.
    cost: a number per set; sets are evaluated most costly first, so that the
    launch rounds (at most `workers` tasks each, each bounded by its longest
    integration) hold integrations of similar length.  Results come back in
    the order of param_sets either way."""
    order = list(range(len(param_sets)))
    if cost is not None:
        order.sort(key=lambda i: -cost[i])
    tasks = [(tuple(param_sets[i]), tuple(ns), int(pt)) for i in order for pt in pts]
    raw = engine.map_batched(lambda p, n, pt: demo_sel_func(p, n, pt), tasks, workers=workers)
    for r in raw:
        if isinstance(r, Exception):
            raise r
    k = len(pts)
    where = {i: r for r, i in enumerate(order)}
    out = []
    for i in range(len(param_sets)):
        res = raw[where[i] * k:(where[i] + 1) * k]
        if k == 1:
            out.append(res[0])
        else:
            out.append(Numerics.extrapolate(res, [r.extrap_x for r in res]))
    return out


class Cache1D:
    """Spectra cached over negative gammas log-spaced in gamma_bounds
    (dadi/DFE/Cache1D_mod.py:22-78).  ``cpus`` and ``gpus`` are accepted for
    compatibility; all generation runs batched on the one GPU."""

    def __init__(self, params, ns, demo_sel_func, pts, gamma_bounds=(1e-4, 2000),
                 gamma_pts=500, additional_gammas=[], cpus=None, gpus=0, verbose=False,
                 workers=512):
        self.params, self.ns, self.pts = tuple(params), tuple(ns), tuple(pts)
        self.gammas = -np.logspace(np.log10(gamma_bounds[1]), np.log10(gamma_bounds[0]),
                                   gamma_pts)
        self.neg_gammas = self.gammas
        self.gammas = np.concatenate((self.gammas, additional_gammas))
        self.func_name = demo_sel_func.__name__
        sets = [self.params + (g,) for g in self.gammas] + [self.params + (0,)]
        specs = _evaluate_grid(demo_sel_func, sets, self.ns, self.pts, workers)
        self.spectra = np.array([s.data if hasattr(s, 'data') else s for s in specs[:-1]])
        self.neu_spec = specs[-1]
        if verbose:
            print('cached %d spectra' % len(self.gammas))

    def integrate(self, params, ns, sel_dist, theta, pts=None, exterior_int=True):
        """Spectrum integrated over sel_dist on the negative gammas
        (dadi/DFE/Cache1D_mod.py:181-230)."""
        Nneg = len(self.neg_gammas)
        spectra = self.spectra[:Nneg]
        weights = sel_dist(-self.neg_gammas, params)
        weighted = 0 * spectra
        for ii, w in enumerate(weights):
            weighted[ii] = w * spectra[ii]
        fs = _trapz(weighted, self.neg_gammas, axis=0)
        if not exterior_int:
            return Spectrum(theta * fs)
        smallest_gamma = self.neg_gammas[-1]
        largest_gamma = self.neg_gammas[0]
        weight_neu, _ = scipy.integrate.quad(sel_dist, 0, -smallest_gamma, args=params)
        weight_del, _ = scipy.integrate.quad(sel_dist, -largest_gamma, np.inf, args=params)
        fs = fs + self.neu_spec.data * weight_neu
        fs = fs + spectra[0] * weight_del
        return Spectrum(theta * fs)

    def integrate_point_pos(self, params, ns, sel_dist, theta, demo_sel_func=None, Npos=1,
                            pts=None, exterior_int=True):
        """(1 - sum ppos) x integrate() plus Npos point masses of positive selection,
        each its cached spectrum times ppos (dadi/DFE/Cache1D_mod.py:232-285).  A
        positive gamma that is not cached is evaluated with demo_sel_func (on the
        GPU) and appended to the cache, multiplied by theta as the reference
        stores it; without demo_sel_func it raises IndexError."""
        pdf_params = params[:-2 * Npos]
        ppos_l, gammapos_l = params[-2 * Npos::2], params[-2 * Npos + 1::2]
        pdf_fs = self.integrate(pdf_params, None, sel_dist, theta, None,
                                exterior_int=exterior_int)
        result = (1 - np.sum(ppos_l)) * pdf_fs
        for ppos, gammapos in zip(ppos_l, gammapos_l):
            if gammapos not in self.gammas:
                if demo_sel_func is None:
                    raise IndexError('Failed to find requested gammapos={0:.4f} in Cache1D '
                                     'spectra.'.format(gammapos))
                # as the reference (dadi/DFE/Cache1D_mod.py:275-279), a spectrum
                # computed here is stored multiplied by theta
                ex = Numerics.make_extrap_func(demo_sel_func)
                pos_fs = theta * ex(tuple(self.params) + (gammapos,), self.ns, self.pts)
                self.gammas = np.append(self.gammas, gammapos)
                self.spectra = np.append(self.spectra, [pos_fs.data], axis=0)
            ii = list(self.gammas).index(gammapos)
            result += ppos * Spectrum(self.spectra[ii])
        return result
