"""
Grid and Richardson extrapolation (dadi/Numerics.py:215-233, 303-501, 644-669).
"""
import functools
import logging

import numpy

logger = logging.getLogger('Numerics')


def exponential_grid(pts, crwd=8.):
    """Logistic-spaced grid on [0, 1], crowded towards both ends
    (dadi/Numerics.py:215-231)."""
    unif = numpy.linspace(-1, 1, pts)
    grid = 1. / (1. + numpy.exp(-crwd * unif))
    grid = (grid - grid[0]) / (grid[-1] - grid[0])
    return grid


#: dadi/Numerics.py:233
default_grid = exponential_grid


def trapz(yy, xx=None, dx=None, axis=-1):
    """Composite trapezoid rule along ``axis`` with a 1D grid
    (dadi/Numerics.py:303-332)."""
    if (xx is None) == (dx is None):
        raise ValueError('One and only one of xx or dx must be specified.')
    if xx is not None:
        dx = numpy.diff(xx)
    yy = numpy.asanyarray(yy)
    nd = yy.ndim
    if yy.shape[axis] != len(dx) + 1:
        raise ValueError('Length of xx must be equal to length of yy along specified axis.')
    s1 = [slice(None)] * nd
    s2 = [slice(None)] * nd
    sx = [numpy.newaxis] * nd
    s1[axis] = slice(1, None)
    s2[axis] = slice(None, -1)
    sx[axis] = slice(None)
    return numpy.sum(dx[tuple(sx)] * (yy[tuple(s1)] + yy[tuple(s2)]) / 2.0, axis=axis)


def linear_extrap(ys, xs):
    """Linear extrapolation of two (x, y) pairs to x = 0 (dadi/Numerics.py:644-655)."""
    y1, y2 = ys
    x1, x2 = xs
    return (x2 * y1 - x1 * y2) / (x2 - x1)


def quadratic_extrap(ys, xs):
    """Lagrange quadratic through three (x, y) pairs evaluated at x = 0
    (dadi/Numerics.py:657-669)."""
    y1, y2, y3 = ys
    x1, x2, x3 = xs
    return x2 * x3 / ((x1 - x2) * (x1 - x3)) * y1 + x1 * x3 / ((x2 - x1) * (x2 - x3)) * y2 \
        + x1 * x2 / ((x3 - x1) * (x3 - x2)) * y3


def extrapolate(result_l, x_l, extrap_log=False, fail_mag=10):
    """Extrapolate a list of results computed at grid spacings x_l to zero
    spacing, with the fail_mag safeguard of dadi/Numerics.py:417-469."""
    if extrap_log:
        result_l = [numpy.log(r) for r in result_l]
    if len(result_l) == 1:
        ex_result = result_l[0]
    elif len(result_l) == 2:
        ex_result = linear_extrap(result_l, x_l)
    elif len(result_l) == 3:
        ex_result = quadratic_extrap(result_l, x_l)
    else:
        raise ValueError('Number of calculations to use for extrapolation must be between 1 '
                         'and 3 (the reference names cubic..quintic forms it does not define)')
    if extrap_log:
        ex_result = numpy.exp(ex_result)
    if len(result_l) > 1:
        best_result = result_l[int(numpy.argmin(x_l))]
        if extrap_log:
            best_result = numpy.exp(best_result)
        extrap_failed = abs(numpy.log10(ex_result / best_result)) > fail_mag
        if numpy.any(extrap_failed):
            logger.warning('Extrapolation may have failed. Check resulting '
                           'frequency spectrum for unexpected results.')
        ex_result[extrap_failed] = best_result[extrap_failed]
    try:
        ex_result.pop_ids = result_l[0].pop_ids
    except AttributeError:
        pass
    return ex_result


# Evaluate the grids of an extrapolated function concurrently when called
# from a batching worker (engine.map_batched); False: one grid after another.
parallel_grids = True


def make_extrap_func(func, extrap_x_l=None, extrap_log=False, fail_mag=10):
    """A version of ``func`` whose last argument is a list of grid sizes and
    which returns the result extrapolated to infinitely many grid points
    (dadi/Numerics.py:334-475)."""
    x_l_from_results = extrap_x_l is None

    def extrap_func(*args, **kwargs):
        if 'pts' not in kwargs:
            other_args, pts_l = args[:-1], args[-1]
        else:
            other_args = args
            pts_l = kwargs.pop('pts')
        no_extrap = kwargs.pop('no_extrap', False)
        if numpy.isscalar(pts_l):
            pts_l = [pts_l]
        partial_func = functools.partial(func, *other_args, **kwargs)
        if parallel_grids:
            # inside engine.map_batched the grids are evaluated side by side,
            # their integrations in the same launches (same results)
            from . import engine
            result_l = engine.map_in_worker(partial_func, list(pts_l))
        else:
            result_l = list(map(partial_func, pts_l))
        if no_extrap:
            return result_l
        if x_l_from_results:
            try:
                x_l = [r.extrap_x for r in result_l]
            except AttributeError:
                raise ValueError("Extrapolation function error: No explicit extrapolation x_l "
                                 "provided, and results do not have 'extrap_x' attributes.")
        else:
            x_l = extrap_x_l
        return extrapolate(result_l, x_l, extrap_log, fail_mag)

    extrap_func.__name__ = func.__name__
    extrap_func.__doc__ = func.__doc__
    return extrap_func


def make_extrap_log_func(func, extrap_x_l=None):
    """make_extrap_func on log values (dadi/Numerics.py:477-501)."""
    return make_extrap_func(func, extrap_x_l=extrap_x_l, extrap_log=True)
