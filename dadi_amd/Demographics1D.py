"""
One-population models used by the configurations (dadi/Demographics1D.py).
"""
import numpy

from . import Numerics, PhiManip, Integration
from .Spectrum_mod import Spectrum


def snm(notused, ns, pts):
    """Standard neutral model (dadi/Demographics1D.py:9-27)."""
    xx = Numerics.default_grid(pts)
    phi = PhiManip.phi_1D(xx)
    return Spectrum.from_phi(phi, ns, (xx,))


def two_epoch(params, ns, pts):
    """Instantaneous size change to nu, T ago (dadi/Demographics1D.py:29-58)."""
    nu, T = params
    xx = Numerics.default_grid(pts)
    phi = PhiManip.phi_1D(xx)
    phi = Integration.one_pop(phi, xx, T, nu)
    return Spectrum.from_phi(phi, ns, (xx,))


two_epoch.__param_names__ = ['nu', 'T']


def growth(params, ns, pts):
    """Exponential growth to nu over T (dadi/Demographics1D.py:60-89)."""
    nu, T = params
    xx = Numerics.default_grid(pts)
    phi = PhiManip.phi_1D(xx)
    nu_func = lambda t: numpy.exp(numpy.log(nu) * t / T)
    phi = Integration.one_pop(phi, xx, T, nu_func)
    return Spectrum.from_phi(phi, ns, (xx,))


growth.__param_names__ = ['nu', 'T']


def bottlegrowth(params, ns, pts):
    """Instantaneous change to nuB then exponential growth to nuF
    (dadi/Demographics1D.py:91-126)."""
    nuB, nuF, T = params
    xx = Numerics.default_grid(pts)
    phi = PhiManip.phi_1D(xx)
    nu_func = lambda t: nuB * numpy.exp(numpy.log(nuF / nuB) * t / T)
    phi = Integration.one_pop(phi, xx, T, nu_func)
    return Spectrum.from_phi(phi, ns, (xx,))


bottlegrowth.__param_names__ = ['nuB', 'nuF', 'T']


def three_epoch(params, ns, pts):
    """Bottleneck nuB for TB, then nuF for TF (dadi/Demographics1D.py:128-161)."""
    nuB, nuF, TB, TF = params
    xx = Numerics.default_grid(pts)
    phi = PhiManip.phi_1D(xx)
    phi = Integration.one_pop(phi, xx, TB, nuB)
    phi = Integration.one_pop(phi, xx, TF, nuF)
    return Spectrum.from_phi(phi, ns, (xx,))


three_epoch.__param_names__ = ['nuB', 'nuF', 'TB', 'TF']
