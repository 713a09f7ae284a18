"""
Integration of the allele-frequency density phi forward in time, 1-5
populations, on the GPU.

The public functions keep dadi's signatures and semantics
(dadi/Integration.py:178-937): parameters may be numbers or one-argument
functions of time, 1-3 population calls return a new array, 4-5 population
calls update the caller's array in place, invalid parameters raise
ValueError, and one Demes event is logged per call.

Each call computes the time schedule on the host (the step never depends on
phi), ships it to the device once and runs every step there.  The arithmetic
reproduces the reference's CPU code path for the same call bit for bit:

* numbers only, 1 population  -> dadi/Integration.py:986-1035 order (PRE1)
* numbers only, 2-3 populations -> dadi/Integration.py:1037-1242 order (PRE)
* any function of time, or 4-5 populations -> dadi/integration{1..5}D.c order (TV)

phi may be a numpy array (copied to the device and back) or a float64 torch
tensor already on the device (kept there; no host round trip).
"""
import logging

import numpy
from numpy import newaxis as nuax  # noqa: F401  (kept for parity with dadi's module)

from . import Demes, Misc
from . import _lib
from . import engine
from . import schedule as _sched

logger = logging.getLogger('Integration')

#: Chang and Cooper's delj trick; only False (delj = 0.5) is supported on the
#: device, as in the reference GPU path (dadi/cuda/Integration.py:171-172).
use_delj_trick = False
#: Time step scale, dadi/Integration.py:39.
timescale_factor = 1e-3
#: Old time-step rule, dadi/Integration.py:41-44.
use_old_timestep = False
old_timescale_factor = 0.1
#: Set by dadi_amd.cuda_enabled(); True means the HIP backend is in use.
cuda_enabled = True


def _backend_ok():
    if not cuda_enabled:
        raise _lib.BackendUnavailable(
            "the HIP backend is disabled (dadi_amd.cuda_enabled(False)); this package has no "
            "CPU integrator -- use dadi's own CPU path")


def _is_dev(phi):
    return type(phi).__module__.startswith("torch")


def _prep_phi(phi, D, copy):
    if _is_dev(phi):
        if phi.dim() != D:
            raise ValueError("phi must be %d-dimensional" % D)
        import torch
        t = phi.to(dtype=torch.float64)
        if copy:
            t = t.clone(memory_format=torch.contiguous_format)
        elif not t.is_contiguous():
            t = t.contiguous()
        return t
    a = numpy.array(phi, dtype=numpy.float64, copy=True, order='C')
    if a.ndim != D:
        raise ValueError("phi must be %d-dimensional" % D)
    return a


def _scalar_all(vs):
    return numpy.all([numpy.isscalar(v) for v in vs])


def _run(D, phi, xx, T, nus, migs, gammas, hs, theta0, initial_t, frozen, nomut, beta,
         deme_ids, const_ok):
    """Shared driver of one_pop .. five_pops; phi is already a private copy
    (numpy) or the tensor to update (device)."""
    xx = numpy.asarray(xx, dtype=numpy.float64)
    if len(xx) != phi.shape[0]:
        raise ValueError("phi and xx sizes differ")
    flat_m = [m for ms in migs for m in ms]
    params = list(nus) + flat_m + list(gammas) + list(hs) + [theta0] + ([beta] if D == 1 else [])
    fz, inj = _sched.masks(D, frozen, nomut)
    if const_ok and _scalar_all(params):
        if numpy.any(numpy.less([T] + list(nus) + flat_m + [theta0], 0)):
            raise ValueError('A time, population size, migration rate, or theta0 '
                             'is < 0. Has the model been mis-specified?')
        if numpy.any(numpy.equal(list(nus), 0)):
            raise ValueError('A population size is 0. Has the model been '
                             'mis-specified?')
        Demes.cache.append(Demes.IntegrationConst(duration=T - initial_t,
                                                  start_sizes=list(nus), mig=flat_m,
                                                  deme_ids=deme_ids))
        dx = numpy.diff(xx)
        dt = min(_sched.compute_dt(nus[k], migs[k] if D > 1 else [0], gammas[k], hs[k],
                                   timescale_factor, use_old_timestep, old_timescale_factor, dx)
                 for k in range(D))
        seg = _sched.const_segments(D, xx, initial_t, T, dt, nus, migs, gammas, hs, theta0,
                                    beta)
        order = _lib.ORDER_PRE1 if D == 1 else _lib.ORDER_PRE
    else:
        f = Misc.ensure_1arg_func
        hist = []
        seg = _sched.tv_segments(D, xx, initial_t, T, [f(v) for v in nus],
                                 [[f(m) for m in ms] for ms in migs], [f(v) for v in gammas],
                                 [f(v) for v in hs], f(theta0), f(beta) if D == 1 else None,
                                 timescale_factor, use_old_timestep, old_timescale_factor,
                                 history=hist)
        Demes.cache.append(Demes.IntegrationNonConst(history=hist, deme_ids=deme_ids))
        order = _lib.ORDER_TV
    return engine.run_single(phi, xx, seg, order, fz, inj, delj=bool(use_delj_trick))


def _check_T(T, initial_t):
    if T - initial_t < 0:
        raise ValueError('Final integration time T (%f) is less than '
                         'intial_time (%f). Integration cannot be run '
                         'backwards.' % (T, initial_t))


def _mig_nonzero(ms):
    return any((callable(m) or m != 0) for m in ms)


def one_pop(phi, xx, T, nu=1, gamma=0, h=0.5, theta0=1.0, initial_t=0,
            frozen=False, beta=1, deme_ids=None):
    """Integrate a 1-population phi forward (dadi/Integration.py:178-258).

    nu, gamma, h, theta0 and beta may be functions of time.
    """
    phi = _prep_phi(phi, 1, copy=True)
    if frozen:
        return phi
    if T - initial_t == 0:
        return phi
    _check_T(T, initial_t)
    _backend_ok()
    return _run(1, phi, xx, T, [nu], [[]], [gamma], [h], theta0, initial_t, [False], [False],
                beta, deme_ids, const_ok=True)


def one_pop_X(phi, xx, T, nu=1, gamma=0, h=0.5, beta=1, alpha=1, theta0=1.0, initial_t=0,
              frozen=False):
    """Integrate a 1-population phi of X-chromosome alleles forward
    (dadi/Integration.py:1256-1352): V = 1/nu x(1-x) (2 beta+4)(beta+1)/(9 beta),
    M = gamma 4/3 (0.5+h+x(1-2h)) x(1-x), injection scaled by
    2/(1+2 beta) (1/(alpha+1) + beta).  beta = Nf/Nm, alpha = mu_m/mu_f.
    Constant parameters only, as in the reference."""
    phi = _prep_phi(phi, 1, copy=True)
    if frozen:
        return phi
    if T - initial_t == 0:
        return phi
    _check_T(T, initial_t)
    if not _scalar_all((nu, gamma, h, theta0, beta, alpha)):
        raise NotImplementedError('X chromosome integration currently only '
                                  'implemented for constant parameters.')
    if numpy.any(numpy.less([T, nu, theta0, beta, alpha], 0)):
        raise ValueError('A time, population size, migration rate, theta0, '
                         'beta, or alpha is < 0. Has the model been '
                         'mis-specified?')
    if numpy.any(numpy.equal([nu, beta], 0)):
        raise ValueError('A population size or beta is 0. Has the model been '
                         'mis-specified?')
    _backend_ok()
    xx = numpy.asarray(xx, dtype=numpy.float64)
    if len(xx) != phi.shape[0]:
        raise ValueError("phi and xx sizes differ")
    dx = numpy.diff(xx)
    dt = _sched.compute_dt(nu, [0], gamma, h, timescale_factor, use_old_timestep,
                           old_timescale_factor, dx)
    seg = _sched.xchrom_segments(xx, initial_t, T, dt, nu, gamma, h, beta, alpha, theta0)
    return engine.run_single(phi, xx, seg, _lib.ORDER_PRE1, 0, 1, delj=bool(use_delj_trick))


def two_pops(phi, xx, T, nu1=1, nu2=1, m12=0, m21=0, gamma1=0, gamma2=0,
             h1=0.5, h2=0.5, theta0=1, initial_t=0, frozen1=False,
             frozen2=False, nomut1=False, nomut2=False, enable_cuda_cached=False,
             deme_ids=None):
    """Integrate a 2-population phi forward (dadi/Integration.py:260-387).

    m12 is the migration rate into population 1 from population 2.
    ``enable_cuda_cached`` is accepted for compatibility and has no effect.
    """
    phi = _prep_phi(phi, 2, copy=True)
    if T - initial_t == 0:
        return phi
    _check_T(T, initial_t)
    if (frozen1 or frozen2) and _mig_nonzero([m12, m21]):
        raise ValueError('Population cannot be frozen and have non-zero '
                         'migration to or from it.')
    _backend_ok()
    return _run(2, phi, xx, T, [nu1, nu2], [[m12], [m21]], [gamma1, gamma2], [h1, h2], theta0,
                initial_t, [frozen1, frozen2], [nomut1, nomut2], 1, deme_ids, const_ok=True)


def three_pops(phi, xx, T, nu1=1, nu2=1, nu3=1,
               m12=0, m13=0, m21=0, m23=0, m31=0, m32=0,
               gamma1=0, gamma2=0, gamma3=0, h1=0.5, h2=0.5, h3=0.5,
               theta0=1, initial_t=0, frozen1=False, frozen2=False,
               frozen3=False, enable_cuda_cached=False, deme_ids=None):
    """Integrate a 3-population phi forward (dadi/Integration.py:389-554)."""
    phi = _prep_phi(phi, 3, copy=True)
    if T - initial_t == 0:
        return phi
    _check_T(T, initial_t)
    if (frozen1 and _mig_nonzero([m12, m21, m13, m31])) \
            or (frozen2 and _mig_nonzero([m12, m21, m23, m32])) \
            or (frozen3 and _mig_nonzero([m13, m31, m23, m32])):
        raise ValueError('Population cannot be frozen and have non-zero '
                         'migration to or from it.')
    _backend_ok()
    return _run(3, phi, xx, T, [nu1, nu2, nu3], [[m12, m13], [m21, m23], [m31, m32]],
                [gamma1, gamma2, gamma3], [h1, h2, h3], theta0, initial_t,
                [frozen1, frozen2, frozen3], [False] * 3, 1, deme_ids, const_ok=True)


def _inplace_result(orig, res):
    """4-5 population calls update the caller's array (dadi/Integration.py:635-937)."""
    if _is_dev(orig):
        if res is not orig:
            orig.copy_(res)
        return orig
    if isinstance(orig, numpy.ndarray) and orig.dtype == numpy.float64 and \
            orig.flags.c_contiguous and orig.flags.writeable:
        orig[...] = res
        return orig
    return res


def four_pops(phi, xx, T, nu1=1, nu2=1, nu3=1, nu4=1,
              m12=0, m13=0, m14=0, m21=0, m23=0, m24=0,
              m31=0, m32=0, m34=0, m41=0, m42=0, m43=0,
              gamma1=0, gamma2=0, gamma3=0, gamma4=0,
              h1=0.5, h2=0.5, h3=0.5, h4=0.5,
              theta0=1, initial_t=0, frozen1=False, frozen2=False,
              frozen3=False, frozen4=False, deme_ids=None):
    """Integrate a 4-population phi forward (dadi/Integration.py:556-727)."""
    if T - initial_t == 0:
        return phi
    _check_T(T, initial_t)
    if (frozen1 and _mig_nonzero([m12, m21, m13, m31, m41, m14])) \
            or (frozen2 and _mig_nonzero([m12, m21, m23, m32, m24, m42])) \
            or (frozen3 and _mig_nonzero([m13, m31, m23, m32, m34, m43])) \
            or (frozen4 and _mig_nonzero([m14, m41, m24, m42, m34, m43])):
        raise ValueError('Population cannot be frozen and have non-zero '
                         'migration to or from it.')
    _backend_ok()
    work = _prep_phi(phi, 4, copy=False)
    res = _run(4, work, xx, T, [nu1, nu2, nu3, nu4],
               [[m12, m13, m14], [m21, m23, m24], [m31, m32, m34], [m41, m42, m43]],
               [gamma1, gamma2, gamma3, gamma4], [h1, h2, h3, h4], theta0, initial_t,
               [frozen1, frozen2, frozen3, frozen4], [False] * 4, 1, deme_ids, const_ok=False)
    return _inplace_result(phi, res)


def five_pops(phi, xx, T, nu1=1, nu2=1, nu3=1, nu4=1, nu5=1,
              m12=0, m13=0, m14=0, m15=0, m21=0, m23=0, m24=0, m25=0,
              m31=0, m32=0, m34=0, m35=0, m41=0, m42=0, m43=0, m45=0,
              m51=0, m52=0, m53=0, m54=0,
              gamma1=0, gamma2=0, gamma3=0, gamma4=0, gamma5=0,
              h1=0.5, h2=0.5, h3=0.5, h4=0.5, h5=0.5,
              theta0=1, initial_t=0, frozen1=False, frozen2=False,
              frozen3=False, frozen4=False, frozen5=False, deme_ids=None):
    """Integrate a 5-population phi forward (dadi/Integration.py:729-937)."""
    if T - initial_t == 0:
        return phi
    _check_T(T, initial_t)
    if (frozen1 and _mig_nonzero([m12, m21, m13, m31, m41, m14, m15, m51])) \
            or (frozen2 and _mig_nonzero([m12, m21, m23, m32, m24, m42, m25, m52])) \
            or (frozen3 and _mig_nonzero([m13, m31, m23, m32, m34, m43, m35, m53])) \
            or (frozen4 and _mig_nonzero([m14, m41, m24, m42, m34, m43, m45, m54])) \
            or (frozen5 and _mig_nonzero([m15, m51, m25, m52, m35, m53, m45, m54])):
        raise ValueError('Population cannot be frozen and have non-zero '
                         'migration to or from it.')
    _backend_ok()
    work = _prep_phi(phi, 5, copy=False)
    res = _run(5, work, xx, T, [nu1, nu2, nu3, nu4, nu5],
               [[m12, m13, m14, m15], [m21, m23, m24, m25], [m31, m32, m34, m35],
                [m41, m42, m43, m45], [m51, m52, m53, m54]],
               [gamma1, gamma2, gamma3, gamma4, gamma5], [h1, h2, h3, h4, h5], theta0,
               initial_t, [frozen1, frozen2, frozen3, frozen4, frozen5], [False] * 5, 1,
               deme_ids, const_ok=False)
    return _inplace_result(phi, res)


def _Vfunc(x, nu, beta=1):
    return 1. / nu * x * (1 - x) * (beta + 1.) ** 2 / (4. * beta)


def _Mfunc1D(x, gamma, h):
    return gamma * 2 * (h + (1 - 2 * h) * x) * x * (1 - x)


def _compute_dt(dx, nu, ms, gamma, h):
    """dadi/Integration.py:147-176 with this module's settings."""
    return _sched.compute_dt(nu, ms, gamma, h, timescale_factor, use_old_timestep,
                             old_timescale_factor, dx)
