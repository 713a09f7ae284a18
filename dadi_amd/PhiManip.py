"""
Initial densities and population splits that feed the integrator
(dadi/PhiManip.py:12-495, 1036-1072).  These run once per model evaluation on
the host; the time stepping that follows runs on the device.
"""
import numpy
from numpy import newaxis as nuax
import scipy.integrate

from . import Demes, Numerics


def phi_1D(xx, nu=1.0, theta0=1.0, gamma=0, h=0.5, theta=None, beta=1, deme_ids=None):
    """Equilibrium density of one population with selection
    (dadi/PhiManip.py:12-108)."""
    Demes.cache = [Demes.Initiation(nu, deme_ids=deme_ids)]
    if theta is not None:
        raise ValueError('The parameter theta has been deprecated in favor of '
                         'parameters nu and theta0.')
    if h == 0.5:
        return phi_1D_genic(xx, nu, theta0, gamma, beta=beta)
    # Williamson et al. (2004) eqn 1 with beta rescaling of gamma; the
    # normalising integrals are done numerically.
    gamma = gamma * 4. * beta / (beta + 1.) ** 2
    Qadjust = 0
    if gamma < 0 and numpy.isinf(numpy.exp(-2 * gamma)):
        Qadjust = -2 * gamma
    integrand = lambda xi: numpy.exp(-4 * gamma * h * xi - 2 * gamma * (1 - 2 * h) * xi ** 2
                                     - Qadjust)
    int0, eps = scipy.integrate.quad(integrand, 0, 1, epsabs=0,
                                     points=numpy.linspace(0, 1, 41))
    ints = numpy.empty(len(xx))
    if gamma < 0:
        for ii, q in enumerate(xx):
            val, eps = scipy.integrate.quad(integrand, q, 1, epsabs=0,
                                            points=numpy.linspace(q, 1, 41))
            ints[ii] = val
        phi = numpy.exp(4 * gamma * h * xx + 2 * gamma * (1 - 2 * h) * xx ** 2) * ints / int0
    else:
        integrand = lambda xi, q: numpy.exp(-4 * gamma * h * (xi - q)
                                            - 2 * gamma * (1 - 2 * h) * (xi ** 2 - q ** 2))
        for ii, q in enumerate(xx):
            val, eps = scipy.integrate.quad(integrand, q, 1, args=(q,))
            ints[ii] = val
        phi = ints / int0
    phi[1:-1] *= 1. / (xx[1:-1] * (1 - xx[1:-1]))
    phi[0] = phi[1]
    if Qadjust == 0:
        phi[-1] = 1. / int0
    else:
        phi[-1] = min(phi[-1], phi[-2])
    return phi * nu * theta0 * 4. * beta / (beta + 1.) ** 2


def phi_1D_genic(xx, nu=1.0, theta0=1.0, gamma=0, theta=None, beta=1):
    """Closed-form equilibrium density with genic selection
    (dadi/PhiManip.py:110-164)."""
    if theta is not None:
        raise ValueError('The parameter theta has been deprecated.')
    if gamma == 0:
        return phi_1D_snm(xx, nu, theta0, beta=beta)
    gamma = gamma * 4. * beta / (beta + 1.) ** 2
    exp = numpy.exp
    if xx[0] == 0 and xx[-1] == 1:
        phi = 0 * xx
        if gamma > -300:
            phi[1:-1] = 1. / (xx[1:-1] * (1 - xx[1:-1])) \
                * (1 - exp(-2 * gamma * (1 - xx[1:-1]))) / (1 - exp(-2 * gamma))
        else:
            phi[1:-1] = 1. / (xx[1:-1] * (1 - xx[1:-1])) * exp(2 * gamma * xx[1:-1])
    else:
        if gamma > -300:
            phi = 1. / (xx * (1 - xx)) * (1 - exp(-2 * gamma * (1 - xx))) / (1 - exp(-2 * gamma))
        else:
            phi = 1. / (xx * (1 - xx)) * exp(2 * gamma * xx)
    if xx[0] == 0:
        phi[0] = phi[1]
    if xx[-1] == 1:
        if gamma < 300:
            limit = 2 * gamma * exp(2 * gamma) / (exp(2 * gamma) - 1)
        else:
            limit = 2 * gamma
        phi[-1] = limit
    return phi * nu * theta0 * 4. * beta / (beta + 1.) ** 2


def phi_1D_snm(xx, nu=1.0, theta0=1.0, theta=None, beta=1):
    """Standard neutral density nu*theta0/x (dadi/PhiManip.py:166-195)."""
    if theta is not None:
        raise ValueError('The parameter theta has been deprecated.')
    if xx[0] == 0:
        phi = 0 * xx
        phi[1:] = nu * theta0 / xx[1:]
        phi[0] = phi[1]
    else:
        phi = nu * theta0 / xx
    return phi * 4. * beta / (beta + 1.) ** 2


def check_xx(xx):
    """dadi/PhiManip.py:197-212."""
    if not xx[0] == 0 and xx[1] == 1:
        raise ValueError('Input xx argument does not run from 0 to 1.')
    if not numpy.all(numpy.diff(xx) >= 0):
        raise ValueError('Input xx argument is not monotonically increasing.')


def phi_1D_to_2D(xx, phi_1D, deme_ids=None):
    """One-to-two population split: phi on the diagonal, normalised by the
    trapezoid width (dadi/PhiManip.py:214-234)."""
    check_xx(xx)
    Demes.cache.append(Demes.Split(proportions=[1], deme_ids=deme_ids))
    pts = len(xx)
    phi_2D = numpy.zeros((pts, pts))
    ii = numpy.arange(1, pts - 1)
    phi_2D[ii, ii] = phi_1D[ii] * 2 / (xx[ii + 1] - xx[ii - 1])
    return phi_2D


def _admixture_intermediates(phi, ad_z, zz):
    """Linear-interpolation weights of admixed frequencies onto the new axis
    (dadi/PhiManip.py:268-300)."""
    upper = numpy.searchsorted(zz, ad_z)
    upper = numpy.minimum(upper, len(zz) - 1)
    upper = numpy.maximum(upper, 1)
    lower = upper - 1
    upper_z = zz[upper]
    lower_z = zz[lower]
    delz0 = numpy.where(lower == 0, 0, zz[lower] - zz[lower - 1])
    delz1 = numpy.where(upper == 0, 0, zz[upper] - zz[lower])
    delz2 = numpy.where(upper == len(zz) - 1, 0, zz[(upper + 1) % len(zz)] - zz[upper])
    frac_lower = (upper_z - ad_z) / (upper_z - lower_z)
    frac_upper = (ad_z - lower_z) / (upper_z - lower_z)
    norm = 2 * phi / (frac_lower * delz0 + delz1 + frac_upper * delz2)
    return lower, upper, frac_lower, frac_upper, norm


def _new_axis(phi, ad, zz):
    lower, upper, fl, fu, norm = _admixture_intermediates(phi, ad, zz)
    idx = numpy.indices(phi.shape, sparse=True)
    out = numpy.zeros(phi.shape + (len(zz),))
    out[tuple(idx) + (lower,)] = fl * norm
    out[tuple(idx) + (upper,)] += fu * norm
    return out


def phi_2D_to_3D_admix(phi, f1, xx, yy, zz, deme_ids=None):
    """Population 3 formed from fractions f1 of pop 1 and 1-f1 of pop 2
    (dadi/PhiManip.py:302-313, 372-409)."""
    Demes.cache.append(Demes.Split(proportions=[f1, 1 - f1], deme_ids=deme_ids))
    ad_z = f1 * xx[:, nuax] + (1 - f1) * yy[nuax, :]
    return _new_axis(phi, ad_z, zz)


phi_2D_to_3D = phi_2D_to_3D_admix


def phi_2D_to_3D_split_2(xx, phi_2D, deme_ids=None):
    """Split population 2 into populations 2 and 3 (dadi/PhiManip.py:236-250)."""
    check_xx(xx)
    return phi_2D_to_3D_admix(phi_2D, 0, xx, xx, xx, deme_ids)


def phi_2D_to_3D_split_1(xx, phi_2D, deme_ids=None):
    """Split population 1 into populations 1 and 3 (dadi/PhiManip.py:252-266)."""
    check_xx(xx)
    return phi_2D_to_3D_admix(phi_2D, 1, xx, xx, xx, deme_ids)


def phi_3D_to_4D(phi, f1, f2, xx, yy, zz, aa, deme_ids=None):
    """Population 4 from pops 1-3 in proportions f1, f2, 1-f1-f2
    (dadi/PhiManip.py:315-331, 412-451)."""
    Demes.cache.append(Demes.Split(proportions=[f1, f2, 1 - f1 - f2], deme_ids=deme_ids))
    if f1 + f2 > 1:
        raise ValueError('Admixture proportions (f1=%f, f2 = %f) are non-sensible.' % (f1, f2))
    ad_w = f1 * xx[:, nuax, nuax] + f2 * yy[nuax, :, nuax] + (1 - f1 - f2) * zz[nuax, nuax, :]
    return _new_axis(phi, ad_w, aa)


def phi_4D_to_5D(phi, f1, f2, f3, xx, yy, zz, aa, bb, deme_ids=None):
    """Population 5 from pops 1-4 (dadi/PhiManip.py:333-350, 453-495)."""
    Demes.cache.append(Demes.Split(proportions=[f1, f2, f3, 1 - f1 - f2 - f3],
                                   deme_ids=deme_ids))
    if f1 + f2 + f3 > 1:
        raise ValueError('Admixture proportions are non-sensible.')
    ad_w = f1 * xx[:, nuax, nuax, nuax] + f2 * yy[nuax, :, nuax, nuax] \
        + f3 * zz[nuax, nuax, :, nuax] + (1 - f1 - f2 - f3) * aa[nuax, nuax, nuax, :]
    return _new_axis(phi, ad_w, bb)


def remove_pop(phi, xx, popnum):
    """Integrate one population out of phi (dadi/PhiManip.py:1036-1050)."""
    Demes.cache.append(Demes.Remove(removed=popnum))
    return Numerics.trapz(phi, xx, axis=popnum - 1)


def filter_pops(phi, xx, tokeep):
    """Keep only the listed populations, numbered from 1 (dadi/PhiManip.py:1052-1072)."""
    toremove = list(range(1, phi.ndim + 1))
    for pop_ii in tokeep:
        toremove.remove(pop_ii)
    for pop_ii in sorted(toremove)[::-1]:
        phi = remove_pop(phi, xx, pop_ii)
    return phi


# ---------------------------------------------------------------------------
# Device-resident splits: the same operations on a float64 torch tensor that
# stays on the GPU between epochs (used by the serving loop and by model
# code that passes device tensors).  Each elementwise step is the IEEE
# operation the numpy version performs, so results are bit-identical.
# ---------------------------------------------------------------------------
def _dev_grid(xx, like):
    import torch
    return torch.as_tensor(numpy.asarray(xx, dtype=numpy.float64), device=like.device)


def phi_1D_to_2D_device(xx, phi_1D):
    """phi_1D_to_2D for a device tensor (dadi/PhiManip.py:214-234)."""
    import torch
    x = _dev_grid(xx, phi_1D)
    pts = len(xx)
    out = torch.zeros((pts, pts), dtype=torch.float64, device=phi_1D.device)
    ii = torch.arange(1, pts - 1, device=phi_1D.device)
    out[ii, ii] = phi_1D[1:-1] * 2 / (x[2:] - x[:-2])
    return out


def _new_axis_device(phi, ad, zz):
    """Device twin of _new_axis / _admixture_intermediates."""
    import torch
    z = _dev_grid(zz, phi)
    nz = len(zz)
    upper = torch.searchsorted(z, ad.contiguous())
    upper = torch.clamp(upper, max=nz - 1)
    upper = torch.clamp(upper, min=1)
    lower = upper - 1
    upper_z = z[upper]
    lower_z = z[lower]
    zero = torch.zeros((), dtype=torch.float64, device=phi.device)
    delz0 = torch.where(lower == 0, zero, z[lower] - z[(lower - 1) % nz])
    delz1 = torch.where(upper == 0, zero, z[upper] - z[lower])
    delz2 = torch.where(upper == nz - 1, zero, z[(upper + 1) % nz] - z[upper])
    frac_lower = (upper_z - ad) / (upper_z - lower_z)
    frac_upper = (ad - lower_z) / (upper_z - lower_z)
    norm = 2 * phi / (frac_lower * delz0 + delz1 + frac_upper * delz2)
    out = torch.zeros(tuple(phi.shape) + (nz,), dtype=torch.float64, device=phi.device)
    flat = out.view(-1, nz)
    rows = torch.arange(flat.shape[0], device=phi.device)
    flat[rows, lower.reshape(-1)] = (frac_lower * norm).reshape(-1)
    flat[rows, upper.reshape(-1)] += (frac_upper * norm).reshape(-1)
    return out


def phi_2D_to_3D_admix_device(phi, f1, xx, yy, zz):
    """phi_2D_to_3D_admix for a device tensor."""
    x = _dev_grid(xx, phi)
    y = _dev_grid(yy, phi)
    ad_z = f1 * x[:, None] + (1 - f1) * y[None, :]
    return _new_axis_device(phi, ad_z, zz)


def phi_2D_to_3D_split_2_device(xx, phi_2D):
    return phi_2D_to_3D_admix_device(phi_2D, 0, xx, xx, xx)


def phi_3D_to_4D_device(phi, f1, f2, xx, yy, zz, aa):
    """phi_3D_to_4D for a device tensor."""
    if f1 + f2 > 1:
        raise ValueError('Admixture proportions (f1=%f, f2 = %f) are non-sensible.' % (f1, f2))
    x = _dev_grid(xx, phi)
    y = _dev_grid(yy, phi)
    z = _dev_grid(zz, phi)
    ad_w = f1 * x[:, None, None] + f2 * y[None, :, None] + (1 - f1 - f2) * z[None, None, :]
    return _new_axis_device(phi, ad_w, aa)


def phi_2D_to_3D_split_1_device(xx, phi_2D):
    """phi_2D_to_3D_split_1 for a device tensor."""
    return phi_2D_to_3D_admix_device(phi_2D, 1, xx, xx, xx)


def phi_4D_to_5D_device(phi, f1, f2, f3, xx, yy, zz, aa, bb):
    """phi_4D_to_5D for a device tensor."""
    if f1 + f2 + f3 > 1:
        raise ValueError('Admixture proportions are non-sensible.')
    x = _dev_grid(xx, phi)
    y = _dev_grid(yy, phi)
    z = _dev_grid(zz, phi)
    a = _dev_grid(aa, phi)
    ad_w = f1 * x[:, None, None, None] + f2 * y[None, :, None, None] \
        + f3 * z[None, None, :, None] + (1 - f1 - f2 - f3) * a[None, None, None, :]
    return _new_axis_device(phi, ad_w, bb)


def _pairwise_last(a):
    """Sum over the last axis in the order numpy's pairwise summation uses for
    contiguous doubles: fewer than 8 terms one after another from 0; up to 128
    terms in 8 interleaved partial sums combined as ((r0+r1)+(r2+r3)) +
    ((r4+r5)+(r6+r7)), then the remainder; longer runs split in two at a
    multiple of 8 below the half."""
    n = a.shape[-1]
    if n < 8:
        import torch
        res = torch.zeros_like(a[..., 0])
        for i in range(n):
            res = res + a[..., i]
        return res
    if n <= 128:
        r = [a[..., j] for j in range(8)]
        i = 8
        while i < n - n % 8:
            r = [r[j] + a[..., i + j] for j in range(8)]
            i += 8
        res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]))
        while i < n:
            res = res + a[..., i]
            i += 1
        return res
    n2 = n // 2
    n2 -= n2 % 8
    return _pairwise_last(a[..., :n2]) + _pairwise_last(a[..., n2:])


def _trapz_device(phi, xx, axis):
    """Numerics.trapz(phi, xx, axis) on a device tensor with numpy's operations
    and summation order: the last axis is summed pairwise, any other axis
    term after term, so results are bit-identical to the host version."""
    x = _dev_grid(xx, phi)
    dx = x[1:] - x[:-1]
    n = phi.shape[axis]
    shape = [1] * phi.dim()
    shape[axis] = n - 1
    terms = dx.reshape(shape) * (phi.narrow(axis, 1, n - 1) + phi.narrow(axis, 0, n - 1)) / 2.0
    if axis == phi.dim() - 1:
        return _pairwise_last(terms)
    terms = terms.movedim(axis, 0)
    acc = terms[0]
    for k in range(1, n - 1):
        acc = acc + terms[k]
    return acc.contiguous()


def remove_pop_device(phi, xx, popnum):
    """remove_pop for a device tensor (dadi/PhiManip.py:1036-1050); no
    Demes.cache event."""
    return _trapz_device(phi, xx, popnum - 1)


def filter_pops_device(phi, xx, tokeep):
    """filter_pops for a device tensor (dadi/PhiManip.py:1052-1072)."""
    toremove = list(range(1, phi.dim() + 1))
    for pop_ii in tokeep:
        toremove.remove(pop_ii)
    for pop_ii in sorted(toremove)[::-1]:
        phi = remove_pop_device(phi, xx, pop_ii)
    return phi
