"""
Parameter uncertainties and likelihood-ratio-test adjustments from the
Godambe (and Fisher) information (dadi/Godambe.py).

The finite-difference Hessian and gradients need the model at a fixed set of
parameter points that does not depend on the model's values.  Here those
points are first collected by running the very same difference formulas
with a recording function, the model spectra at all of them are evaluated in
one batch (engine.map_batched: their integrations share persistent GPU
launches), and the formulas then run again over the cached spectra -- the
same numbers, in the same order, as one evaluation at a time.
"""
import numpy

from . import Inference, engine
from .Spectrum_mod import Spectrum

two_pt_deriv_test = False


def hessian_elem(func, f0, p0, ii, jj, eps, args=(), one_sided=None):
    """Element [ii][jj] of the finite-difference Hessian
    (dadi/Godambe.py:11-86)."""
    if one_sided is None:
        one_sided = [False] * len(p0)

    pwork = numpy.array(p0, copy=True, dtype=float)
    if ii == jj:
        if pwork[ii] != 0 and not one_sided[ii]:
            pwork[ii] = p0[ii] + eps[ii]
            fp = func(pwork, *args)
            pwork[ii] = p0[ii] - eps[ii]
            fm = func(pwork, *args)
            element = (fp - 2 * f0 + fm) / eps[ii] ** 2
        else:
            pwork[ii] = p0[ii] + 2 * eps[ii]
            fpp = func(pwork, *args)
            pwork[ii] = p0[ii] + eps[ii]
            fp = func(pwork, *args)
            element = (fpp - 2 * fp + f0) / eps[ii] ** 2
    else:
        if pwork[ii] != 0 and pwork[jj] != 0 and not one_sided[ii] and not one_sided[jj]:
            pwork[ii] = p0[ii] + eps[ii]
            pwork[jj] = p0[jj] + eps[jj]
            fpp = func(pwork, *args)
            pwork[ii] = p0[ii] + eps[ii]
            pwork[jj] = p0[jj] - eps[jj]
            fpm = func(pwork, *args)
            pwork[ii] = p0[ii] - eps[ii]
            pwork[jj] = p0[jj] + eps[jj]
            fmp = func(pwork, *args)
            pwork[ii] = p0[ii] - eps[ii]
            pwork[jj] = p0[jj] - eps[jj]
            fmm = func(pwork, *args)
            element = (fpp - fpm - fmp + fmm) / (4 * eps[ii] * eps[jj])
        else:
            pwork[ii] = p0[ii] + eps[ii]
            pwork[jj] = p0[jj] + eps[jj]
            fpp = func(pwork, *args)
            pwork[ii] = p0[ii] + eps[ii]
            pwork[jj] = p0[jj]
            fpm = func(pwork, *args)
            pwork[ii] = p0[ii]
            pwork[jj] = p0[jj] + eps[jj]
            fmp = func(pwork, *args)
            element = (fpp - fpm - fmp + f0) / (eps[ii] * eps[jj])
    return element


def _steps(p0, eps_in):
    """Absolute finite-difference steps and one-sidedness per parameter
    (dadi/Godambe.py:106-119, 139-152)."""
    eps = numpy.empty([len(p0)])
    one_sided = [False] * len(p0)
    for i, pval in enumerate(p0):
        if pval != 0:
            if pval * eps_in < 1e-6:
                eps[i] = eps_in
                one_sided[i] = True
            else:
                eps[i] = eps_in * pval
        else:
            eps[i] = eps_in
    return eps, one_sided


def get_hess(func, p0, eps, args=()):
    """Finite-difference Hessian of func at p0 with fractional step eps
    (dadi/Godambe.py:88-126)."""
    eps, one_sided = _steps(p0, eps)
    f0 = func(p0, *args)
    hess = numpy.empty((len(p0), len(p0)))
    for ii in range(len(p0)):
        for jj in range(ii, len(p0)):
            hess[ii][jj] = hessian_elem(func, f0, p0, ii, jj, eps, args=args,
                                        one_sided=one_sided)
            hess[jj][ii] = hess[ii][jj]
    return hess


def get_grad(func, p0, eps, args=()):
    """Finite-difference gradient of func at p0 with fractional step eps
    (dadi/Godambe.py:128-177)."""
    eps, one_sided = _steps(p0, eps)
    grad = numpy.empty([len(p0), 1])
    for ii in range(len(p0)):
        pwork = numpy.array(p0, copy=True, dtype=float)
        if p0[ii] != 0 and not one_sided[ii]:
            pwork[ii] = p0[ii] + eps[ii]
            fp = func(pwork, *args)
            pwork[ii] = p0[ii] - eps[ii]
            fm = func(pwork, *args)
            grad[ii] = (fp - fm) / (2 * eps[ii])
        else:
            pwork[ii] = p0[ii] + eps[ii]
            fp = func(pwork, *args)
            pwork[ii] = p0[ii]
            fm = func(pwork, *args)
            grad[ii] = (fp - fm) / eps[ii]
            if two_pt_deriv_test:
                pwork[ii] = p0[ii] + 2 * eps[ii]
                fpp = func(pwork, *args)
                grad[ii] = (4 * fp - fpp - 3 * fm) / (2 * eps[ii])
    return grad


#: model spectra of get_godambe, keyed by (function, parameters, ns, grids)
cache = {}


def get_godambe(func_ex, grid_pts, all_boot, p0, data, eps, log=False, just_hess=False,
                boot_theta_adjusts=[], batch=True, workers=256):
    """Godambe information, Hessian, J and mean score (dadi/Godambe.py:179-253).
    batch=True evaluates every finite-difference point of the Hessian and
    gradients in one batch first."""
    ns = data.sample_sizes
    if not boot_theta_adjusts:
        boot_theta_adjusts = numpy.ones(len(all_boot))

    def key(params):
        return (func_ex.__hash__(), tuple(params), tuple(ns), tuple(grid_pts))

    def func(params, data, theta_adjust=1):
        k = key(params)
        if k not in cache:
            cache[k] = func_ex(params, ns, grid_pts)
        fs = theta_adjust * cache[k]
        return Inference.ll(fs, data)

    def log_func(logparams, data, theta_adjust=1):
        return func(numpy.exp(logparams), data, theta_adjust)

    if batch:
        # record the parameter points the formulas below will visit
        points = []

        def rec(params, data, theta_adjust=1):
            points.append(numpy.array(params, dtype=float, copy=True))
            return 0.0

        def rec_log(logparams, data, theta_adjust=1):
            return rec(numpy.exp(logparams), data, theta_adjust)

        f, x0 = (rec_log, numpy.log(p0)) if log else (rec, p0)
        get_hess(f, x0, eps, args=[data])
        if not just_hess and len(all_boot):
            get_grad(f, x0, eps, args=[Spectrum(all_boot[0]), 1])
        todo, seen = [], set()
        for p in points:
            k = key(p)
            if k not in cache and k not in seen:
                seen.add(k)
                todo.append(p)
        res = engine.map_batched(lambda p: func_ex(p, ns, grid_pts), [(p,) for p in todo],
                                 workers=workers)
        for p, r in zip(todo, res):
            if not isinstance(r, Exception):
                cache[key(p)] = r

    if not log:
        hess = -get_hess(func, p0, eps, args=[data])
    else:
        hess = -get_hess(log_func, numpy.log(p0), eps, args=[data])
    if just_hess:
        return hess

    J = numpy.zeros((len(p0), len(p0)))
    cU = numpy.zeros((len(p0), 1))
    for boot, theta_adjust in zip(all_boot, boot_theta_adjusts):
        boot = Spectrum(boot)
        if not log:
            grad_temp = get_grad(func, p0, eps, args=[boot, theta_adjust])
        else:
            grad_temp = get_grad(log_func, numpy.log(p0), eps, args=[boot, theta_adjust])
        J = J + numpy.outer(grad_temp, grad_temp)
        cU = cU + grad_temp
    J = J / len(all_boot)
    cU = cU / len(all_boot)

    J_inv = numpy.linalg.inv(J)
    godambe = numpy.dot(numpy.dot(hess, J_inv), hess)
    return godambe, hess, J, cU


def _with_theta(func_ex, grid_pts, p0, data):
    """Append the optimal theta to p0 and scale the model by it
    (dadi/Godambe.py:288-294)."""
    func_multi = func_ex
    model = func_multi(p0, data.sample_sizes, grid_pts)
    theta_opt = Inference.optimal_sfs_scaling(model, data)
    p0 = list(p0) + [theta_opt]
    return (lambda p, ns, pts: p[-1] * func_multi(p[:-1], ns, pts)), p0


def GIM_uncert(func_ex, grid_pts, all_boot, p0, data, log=False, multinom=True, eps=0.01,
               return_GIM=False, boot_theta_adjusts=None, batch=True):
    """Standard deviations of the parameters from the Godambe information
    (dadi/Godambe.py:255-302)."""
    if multinom:
        if boot_theta_adjusts:
            raise ValueError('boot_thetas option can only be used with multinom=False')
        func_ex, p0 = _with_theta(func_ex, grid_pts, p0, data)
    GIM, H, J, cU = get_godambe(func_ex, grid_pts, all_boot, p0, data, eps, log,
                                boot_theta_adjusts=boot_theta_adjusts or [], batch=batch)
    uncerts = numpy.sqrt(numpy.diag(numpy.linalg.inv(GIM)))
    if not return_GIM:
        return uncerts
    return uncerts, GIM, H


def FIM_uncert(func_ex, grid_pts, p0, data, log=False, multinom=True, eps=0.01,
               return_FIM=False, batch=True):
    """Standard deviations of the parameters from the Fisher information
    (dadi/Godambe.py:304-341)."""
    if multinom:
        func_ex, p0 = _with_theta(func_ex, grid_pts, p0, data)
    H = get_godambe(func_ex, grid_pts, [], p0, data, eps, log, just_hess=True, batch=batch)
    uncerts = numpy.sqrt(numpy.diag(numpy.linalg.inv(H)))
    if not return_FIM:
        return uncerts
    return uncerts, H


def LRT_adjust(func_ex, grid_pts, all_boot, p0, data, nested_indices, multinom=True,
               eps=0.01, boot_theta_adjusts=None, batch=True):
    """First-order moment-matching adjustment of the likelihood-ratio test
    statistic (dadi/Godambe.py:343-392)."""
    if multinom and boot_theta_adjusts:
        raise ValueError('Cannot use boot_theta_adjusts with multinom=True.')
    if multinom:
        func_ex, p0 = _with_theta(func_ex, grid_pts, p0, data)

    def diff_func(diff_params, ns, grid_pts):
        full_params = numpy.array(p0, copy=True, dtype=float)
        full_params[nested_indices] = diff_params
        return func_ex(full_params, ns, grid_pts)

    p_nested = numpy.asarray(p0)[nested_indices]
    GIM, H, J, cU = get_godambe(diff_func, grid_pts, all_boot, p_nested, data, eps, log=False,
                                boot_theta_adjusts=boot_theta_adjusts or [], batch=batch)
    return len(nested_indices) / numpy.trace(numpy.dot(J, numpy.linalg.inv(H)))
