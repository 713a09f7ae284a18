"""
Poisson and multinomial log-likelihoods of a data spectrum given a model
spectrum (dadi/Inference.py:466-575, 668-685).
"""
import logging

import numpy
from numpy import logical_and, logical_not
from scipy.special import gammaln

logger = logging.getLogger('Inference')


def _masked(a):
    if isinstance(a, numpy.ma.MaskedArray):
        return a
    return numpy.ma.masked_array(numpy.asarray(a, dtype=float))


_FOLDS = ('folded', 'folded_ancestral', 'folded_major')


def _refuse_folded(model, data):
    """The reference folds an unfolded model to match folded data before
    comparing them (dadi/Inference.py:493-501, 676-683).  Spectrum folding is not
    part of this package, so such a comparison raises instead of returning a
    likelihood of mismatched spectra."""
    for attr in _FOLDS:
        if getattr(data, attr, False) and not getattr(model, attr, False):
            raise NotImplementedError(
                'data spectrum is %s but the model is not; folding the model '
                '(Spectrum.fold*) is not implemented in dadi_amd' % attr.replace('_', ' '))


def ll_per_bin(model, data, missing_model_cutoff=1e-6):
    """Poisson log-likelihood of each entry: -model + data*log(model) - lnGamma(data+1),
    masked where either spectrum is masked or the model is not positive
    (dadi/Inference.py:482-548).  Folded data with an unfolded model raises
    NotImplementedError."""
    _refuse_folded(model, data)
    model = _masked(model)
    data = _masked(data)
    result = -model.data + data.data * numpy.ma.log(model) - gammaln(data + 1.)
    if numpy.all(result.mask == numpy.ma.getmaskarray(data)):
        return result
    not_data_mask = logical_not(numpy.ma.getmaskarray(data))
    data_sum = data.sum()
    missing = logical_and(model < 0, not_data_mask)
    if numpy.any(missing) and data[missing].sum() / data_sum > missing_model_cutoff:
        logger.warning('Model is < 0 where data is not masked.')
    missing = logical_and(model == 0, logical_and(data > 0, not_data_mask))
    if numpy.any(missing) and data[missing].sum() / data_sum > missing_model_cutoff:
        logger.warning('Model is 0 where data is neither masked nor 0.')
    return result


def ll(model, data):
    """Sum of ll_per_bin (dadi/Inference.py:466-480)."""
    return ll_per_bin(model, data).sum()


def optimal_sfs_scaling(model, data):
    """sum(data)/sum(model) over entries masked in neither (dadi/Inference.py:668-685).
    Folded data with an unfolded model raises NotImplementedError."""
    _refuse_folded(model, data)
    model = _masked(model)
    data = _masked(data)
    mask = numpy.ma.getmaskarray(model) | numpy.ma.getmaskarray(data)
    m = numpy.ma.masked_array(model.data, mask=mask)
    d = numpy.ma.masked_array(data.data, mask=mask)
    return d.sum() / m.sum()


def ll_multinom_per_bin(model, data):
    """Poisson ll per bin of the optimally scaled model (dadi/Inference.py:550-557)."""
    theta_opt = optimal_sfs_scaling(model, data)
    return ll_per_bin(theta_opt * model, data)


def ll_multinom(model, data):
    """Log-likelihood with the optimal theta scaling (dadi/Inference.py:559-575)."""
    return ll_multinom_per_bin(model, data).sum()


# ---------------------------------------------------------------------------
# Optimisation with batched model evaluations
# ---------------------------------------------------------------------------
# The optimisers below follow dadi's (dadi/Inference.py:22-178, 337-458,
# 840-1036, 1152-1285): the same objective, the same scipy drivers, the same
# arguments.  What differs is how model spectra are produced: every point an
# optimiser is about to request is evaluated together with the others --
# for BFGS the point and its finite-difference stencil, for a grid search the
# whole grid -- through engine.map_batched, so their integrations share
# persistent GPU launches.  Spectra are cached under the exact bits of their
# parameters, and the optimiser then sees the very numbers a one-at-a-time run
# would have produced: its trajectory is unchanged.

import os
import sys

import scipy.optimize

from . import engine

#: Stores thetas
_theta_store = {}
#: Counts calls to object_func
_counter = 0
#: Returned when object_func is passed out-of-bounds params or gets a NaN ll.
_out_of_bounds_val = -1e8


class ModelCache:
    """Model spectra under the exact bits of their parameter vectors.
    ``prefetch`` evaluates every parameter set not yet cached in one
    engine.map_batched call (worker threads whose integrations are coalesced
    into shared launches)."""

    def __init__(self, model_func, ns, pts, func_args=(), func_kwargs=None, workers=256):
        self.model_func = model_func
        self.ns = ns
        self.pts = pts
        self.func_args = list(func_args)
        self.func_kwargs = dict(func_kwargs or {})
        self.workers = workers
        self.store = {}
        self.batched = 0   # evaluations made in prefetch batches
        self.serial = 0    # evaluations made one at a time (not prefetched)

    @staticmethod
    def key(params):
        return numpy.asarray(params, dtype=float).tobytes()

    def _eval(self, params):
        kw = self.func_kwargs.copy()
        kw['pts'] = self.pts
        return self.model_func(*([params, self.ns] + self.func_args), **kw)

    def prefetch(self, params_list):
        todo, seen = [], set()
        for p in params_list:
            k = self.key(p)
            if k not in self.store and k not in seen:
                seen.add(k)
                todo.append(p)
        if len(todo) < 2:
            return
        res = engine.map_batched(self._eval, [(p,) for p in todo], workers=self.workers)
        self.batched += len(todo)
        for p, r in zip(todo, res):
            if not isinstance(r, Exception):   # failures recur (and raise) when requested
                self.store[self.key(p)] = r

    def __call__(self, params):
        k = self.key(params)
        hit = self.store.get(k)
        if hit is None:
            self.serial += 1
            hit = self._eval(params)
            self.store[k] = hit
        return hit


def _in_bounds(params_up, lower_bound, upper_bound):
    if lower_bound is not None:
        for pval, bound in zip(params_up, lower_bound):
            if bound is not None and pval < bound:
                return False
    if upper_bound is not None:
        for pval, bound in zip(params_up, upper_bound):
            if bound is not None and pval > bound:
                return False
    return True


def _fd_points(x, epsilon, lb=None, ub=None):
    """x and the points of scipy's 2-point finite-difference gradient with
    absolute step epsilon (scipy.optimize._numdiff.approx_derivative: x with
    x[i] + h[i], h = epsilon unless that step vanishes in floating point); with
    bounds, also the backward step of components the forward step leaves."""
    x = numpy.asarray(x, dtype=float)
    h = numpy.full(x.shape, float(epsilon))
    sign = (x >= 0).astype(float) * 2 - 1
    h = numpy.where(((x + h) - x) == 0,
                    numpy.finfo(float).eps ** 0.5 * sign * numpy.maximum(1.0, numpy.abs(x)), h)
    out = [x]
    for i in range(x.size):
        x1 = x.copy()
        x1[i] += h[i]
        out.append(x1)
        if ub is not None and ub[i] is not None and x1[i] > ub[i]:
            x2 = x.copy()
            x2[i] -= h[i]
            out.append(x2)
    return out


class _Prefetch:
    """Before the objective evaluates a new optimiser point x, evaluate x and
    its finite-difference stencil together."""

    def __init__(self, cache, log, fixed_params, lower_bound, upper_bound, epsilon, ub=None):
        self.cache = cache
        self.log = log
        self.fixed_params = fixed_params
        self.lower_bound = lower_bound
        self.upper_bound = upper_bound
        self.epsilon = epsilon
        self.ub = ub

    def params_of(self, x):
        p = numpy.exp(x) if self.log else x
        return _project_params_up(p, self.fixed_params)

    def __call__(self, x):
        if self.cache.key(self.params_of(x)) in self.cache.store:
            return
        ups = [self.params_of(p) for p in _fd_points(x, self.epsilon, ub=self.ub)]
        self.cache.prefetch([u for u in ups if _in_bounds(u, self.lower_bound,
                                                          self.upper_bound)])


def _object_func(params, data, model_func, pts, lower_bound=None, upper_bound=None,
                 verbose=0, multinom=True, flush_delay=0, func_args=[], func_kwargs={},
                 fixed_params=None, ll_scale=1, output_stream=sys.stdout, store_thetas=False,
                 model_cache=None):
    """Objective of the optimisers: -ll of the model at params (out of
    bounds or NaN: 1e8), dadi/Inference.py:22-72.  model_cache, when given,
    supplies the model spectra."""
    global _counter
    _counter += 1

    params_up = _project_params_up(params, fixed_params)
    if not _in_bounds(params_up, lower_bound, upper_bound):
        return -_out_of_bounds_val / ll_scale

    ns = data.sample_sizes
    if model_cache is not None:
        sfs = model_cache(params_up)
    else:
        all_args = [params_up, ns] + list(func_args)
        func_kwargs = func_kwargs.copy()
        func_kwargs['pts'] = pts
        sfs = model_func(*all_args, **func_kwargs)
    if multinom:
        result = ll_multinom(sfs, data)
    else:
        result = ll(sfs, data)

    if store_thetas:
        _theta_store[tuple(params)] = optimal_sfs_scaling(sfs, data)

    if numpy.isnan(result):
        result = _out_of_bounds_val

    if (verbose > 0) and (_counter % verbose == 0):
        param_str = 'array([%s])' % (', '.join(['%- 12g' % v for v in params_up]))
        output_stream.write('%-8i, %-12g, %s%s' % (_counter, result, param_str, os.linesep))
        output_stream.flush()

    return -result / ll_scale


def _object_func_log(log_params, *args, **kwargs):
    """Objective in log(params), dadi/Inference.py:74-78."""
    return _object_func(numpy.exp(log_params), *args, **kwargs)


def _batched_objective(base, prefetch):
    def f(x, *args):
        if prefetch is not None:
            prefetch(x)
        return base(x, *args)
    return f


def _project_params_down(pin, fixed_params):
    """Eliminate fixed parameters from pin (dadi/Inference.py:1152-1168)."""
    if fixed_params is None:
        return pin
    if len(pin) != len(fixed_params):
        raise ValueError('fixed_params list must have same length as input '
                         'parameter array.')
    pout = []
    for curr_val, fixed_val in zip(pin, fixed_params):
        if fixed_val is None:
            pout.append(curr_val)
    return numpy.array(pout)


def _project_params_up(pin, fixed_params):
    """Fold fixed parameters into pin (dadi/Inference.py:1170-1189)."""
    if fixed_params is None:
        return pin
    if numpy.isscalar(pin):
        pin = [pin]
    pout = numpy.zeros(len(fixed_params))
    orig_ii = 0
    for out_ii, val in enumerate(fixed_params):
        if val is None:
            pout[out_ii] = pin[orig_ii]
            orig_ii += 1
        else:
            pout[out_ii] = fixed_params[out_ii]
    return pout


#: the model cache of the last batched optimisation (evaluation counts)
last_model_cache = None


def _setup(data, model_func, pts, func_args, func_kwargs, batch, workers):
    global last_model_cache
    if not batch:
        return None
    last_model_cache = ModelCache(model_func, data.sample_sizes, pts, func_args, func_kwargs,
                                  workers)
    return last_model_cache


def _bfgs(log, p0, data, model_func, pts, lower_bound, upper_bound, verbose, epsilon, gtol,
          multinom, maxiter, full_output, func_args, func_kwargs, fixed_params, ll_scale,
          output_file, batch, workers):
    output_stream = open(output_file, 'w') if output_file else sys.stdout
    cache = _setup(data, model_func, pts, func_args, func_kwargs, batch, workers)
    pre = None if cache is None else _Prefetch(cache, log, fixed_params, lower_bound,
                                               upper_bound, epsilon)
    args = (data, model_func, pts, lower_bound, upper_bound, verbose, multinom, 0,
            func_args, func_kwargs, fixed_params, ll_scale, output_stream, False, cache)
    p0 = _project_params_down(p0, fixed_params)
    base = _object_func_log if log else _object_func
    x0 = numpy.log(p0) if log else p0
    outputs = scipy.optimize.fmin_bfgs(_batched_objective(base, pre), x0, epsilon=epsilon,
                                       args=args, gtol=gtol, full_output=True, disp=False,
                                       maxiter=maxiter)
    xopt, fopt, gopt, Bopt, func_calls, grad_calls, warnflag = outputs
    xopt = _project_params_up(numpy.exp(xopt) if log else xopt, fixed_params)
    if output_file:
        output_stream.close()
    if not full_output:
        return xopt
    return xopt, fopt, gopt, Bopt, func_calls, grad_calls, warnflag


def optimize_log(p0, data, model_func, pts, lower_bound=None, upper_bound=None, verbose=0,
                 flush_delay=0.5, epsilon=1e-3, gtol=1e-5, multinom=True, maxiter=None,
                 full_output=False, func_args=[], func_kwargs={}, fixed_params=None,
                 ll_scale=1, output_file=None, batch=True, workers=256):
    """BFGS in log(params) (dadi/Inference.py:80-176).  batch=True evaluates
    each new point together with its finite-difference stencil; the result is
    the same as batch=False, one model evaluation at a time."""
    return _bfgs(True, p0, data, model_func, pts, lower_bound, upper_bound, verbose, epsilon,
                 gtol, multinom, maxiter, full_output, func_args, func_kwargs, fixed_params,
                 ll_scale, output_file, batch, workers)


def optimize(p0, data, model_func, pts, lower_bound=None, upper_bound=None, verbose=0,
             flush_delay=0.5, epsilon=1e-3, gtol=1e-5, multinom=True, maxiter=None,
             full_output=False, func_args=[], func_kwargs={}, fixed_params=None, ll_scale=1,
             output_file=None, batch=True, workers=256):
    """BFGS in params (dadi/Inference.py:840-923); batching as optimize_log."""
    return _bfgs(False, p0, data, model_func, pts, lower_bound, upper_bound, verbose, epsilon,
                 gtol, multinom, maxiter, full_output, func_args, func_kwargs, fixed_params,
                 ll_scale, output_file, batch, workers)


def _lbfgsb(log, p0, data, model_func, pts, lower_bound, upper_bound, verbose, epsilon, pgtol,
            multinom, maxiter, full_output, func_args, func_kwargs, fixed_params, ll_scale,
            output_file, batch, workers):
    output_stream = open(output_file, 'w') if output_file else sys.stdout
    cache = _setup(data, model_func, pts, func_args, func_kwargs, batch, workers)
    args = (data, model_func, pts, None, None, verbose, multinom, 0, func_args, func_kwargs,
            fixed_params, ll_scale, output_stream, False, cache)

    def opt_bounds(b):
        if b is None:
            return [None] * len(p0)
        if log:
            b = numpy.log(b)
            b[numpy.isnan(b)] = None
        return _project_params_down(b, fixed_params)

    lb, ub = opt_bounds(lower_bound), opt_bounds(upper_bound)
    bounds = list(zip(lb, ub))
    pre = None if cache is None else _Prefetch(cache, log, fixed_params, None, None, epsilon,
                                               ub=list(ub))
    p0 = _project_params_down(p0, fixed_params)
    base = _object_func_log if log else _object_func
    x0 = numpy.log(p0) if log else p0
    xopt, fopt, info_dict = scipy.optimize.fmin_l_bfgs_b(
        _batched_objective(base, pre), x0, bounds=bounds, epsilon=epsilon, args=args,
        iprint=-1, pgtol=pgtol, maxfun=maxiter, approx_grad=True)
    xopt = _project_params_up(numpy.exp(xopt) if log else xopt, fixed_params)
    if output_file:
        output_stream.close()
    if not full_output:
        return xopt
    return xopt, fopt, info_dict


def optimize_log_lbfgsb(p0, data, model_func, pts, lower_bound=None, upper_bound=None,
                        verbose=0, flush_delay=0.5, epsilon=1e-3, pgtol=1e-5, multinom=True,
                        maxiter=1e5, full_output=False, func_args=[], func_kwargs={},
                        fixed_params=None, ll_scale=1, output_file=None, batch=True,
                        workers=256):
    """L-BFGS-B in log(params) (dadi/Inference.py:337-458); batching as
    optimize_log."""
    return _lbfgsb(True, p0, data, model_func, pts, lower_bound, upper_bound, verbose, epsilon,
                   pgtol, multinom, maxiter, full_output, func_args, func_kwargs, fixed_params,
                   ll_scale, output_file, batch, workers)


def optimize_lbfgsb(p0, data, model_func, pts, lower_bound=None, upper_bound=None, verbose=0,
                    flush_delay=0.5, epsilon=1e-3, pgtol=1e-5, multinom=True, maxiter=1e5,
                    full_output=False, func_args=[], func_kwargs={}, fixed_params=None,
                    ll_scale=1, output_file=None, batch=True, workers=256):
    """L-BFGS-B in params (dadi/Inference.py:925-1035); batching as
    optimize_log."""
    return _lbfgsb(False, p0, data, model_func, pts, lower_bound, upper_bound, verbose,
                   epsilon, pgtol, multinom, maxiter, full_output, func_args, func_kwargs,
                   fixed_params, ll_scale, output_file, batch, workers)


index_exp = numpy.index_exp


def optimize_grid(data, model_func, pts, grid, verbose=0, flush_delay=0.5, multinom=True,
                  full_output=False, func_args=[], func_kwargs={}, fixed_params=None,
                  output_file=None, batch=True, workers=256):
    """Brute-force search over a grid given as index_exp ranges
    (dadi/Inference.py:1191-1285).  batch=True evaluates every grid point in
    one batch before scipy.optimize.brute walks the grid."""
    global _theta_store
    output_stream = open(output_file, 'w') if output_file else sys.stdout
    cache = _setup(data, model_func, pts, func_args, func_kwargs, batch, workers)
    args = (data, model_func, pts, None, None, verbose, multinom, 0, func_args, func_kwargs,
            fixed_params, 1.0, output_stream, full_output, cache)
    if full_output:
        _theta_store = {}
    if cache is not None:
        # dry run: the grid points brute will visit, in its own arithmetic
        seen = []
        scipy.optimize.brute(lambda x, *a: seen.append(
            _project_params_up(numpy.array(x, dtype=float, copy=True), fixed_params)) or 0.0,
            ranges=grid, args=args, full_output=False, finish=None)
        cache.prefetch(seen)
    outputs = scipy.optimize.brute(_object_func, ranges=grid, args=args,
                                   full_output=full_output, finish=False)
    if full_output:
        xopt, fopt, grid, fout = outputs
        thetas = numpy.zeros(fout.shape)
        for indices, _ in numpy.ndenumerate(fout):
            grid_indices = tuple([slice(None, None, None)] + list(indices))
            thetas[indices] = _theta_store[tuple(grid[grid_indices])]
    else:
        xopt = outputs
    xopt = _project_params_up(xopt, fixed_params)
    if output_file:
        output_stream.close()
    if not full_output:
        return xopt
    return xopt, fopt, grid, fout, thetas
