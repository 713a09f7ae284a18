"""Small helpers shared by the drivers."""
import numpy


def ensure_1arg_func(var):
    """Turn a constant into a function of time; check that callables take one
    argument.  Values are returned as numpy.float64, as in
    dadi/Misc.py:45-68."""
    if numpy.isscalar(var):
        var_f_tmp = lambda t: var
    else:
        var_f_tmp = var
    var_f = lambda t: numpy.float64(var_f_tmp(t))
    if not callable(var_f_tmp):
        raise ValueError('Argument is not a constant or a function.')
    try:
        var_f(0.0)
    except TypeError:
        raise ValueError('Argument is not a constant or a one-argument function.')
    return var_f
