"""
Two-population models used by the configurations (dadi/Demographics2D.py).
"""
from . import Numerics, PhiManip, Integration
from .Spectrum_mod import Spectrum


def snm(notused, ns, pts):
    """Standard neutral model, two populations (dadi/Demographics2D.py:10-27)."""
    xx = Numerics.default_grid(pts)
    phi = PhiManip.phi_1D(xx)
    phi = PhiManip.phi_1D_to_2D(xx, phi)
    return Spectrum.from_phi(phi, ns, (xx, xx))


def split_mig(params, ns, pts):
    """Split into sizes nu1, nu2 T ago with symmetric migration m
    (dadi/Demographics2D.py:133-167)."""
    nu1, nu2, T, m = params
    xx = Numerics.default_grid(pts)
    phi = PhiManip.phi_1D(xx)
    phi = PhiManip.phi_1D_to_2D(xx, phi)
    phi = Integration.two_pops(phi, xx, T, nu1, nu2, m12=m, m21=m)
    return Spectrum.from_phi(phi, ns, (xx, xx))


split_mig.__param_names__ = ['nu1', 'nu2', 'T', 'm']


def split_asym_mig(params, ns, pts):
    """Split with asymmetric migration (dadi/Demographics2D.py:190-226)."""
    nu1, nu2, T, m12, m21 = params
    xx = Numerics.default_grid(pts)
    phi = PhiManip.phi_1D(xx)
    phi = PhiManip.phi_1D_to_2D(xx, phi)
    phi = Integration.two_pops(phi, xx, T, nu1, nu2, m12=m12, m21=m21)
    return Spectrum.from_phi(phi, ns, (xx, xx))


split_asym_mig.__param_names__ = ['nu1', 'nu2', 'T', 'm12', 'm21']


def IM(params, ns, pts):
    """Isolation with migration and exponential growth
    (dadi/Demographics2D.py:268-309)."""
    s, nu1, nu2, T, m12, m21 = params
    xx = Numerics.default_grid(pts)
    phi = PhiManip.phi_1D(xx)
    phi = PhiManip.phi_1D_to_2D(xx, phi)
    nu1_func = lambda t: s * (nu1 / s) ** (t / T)
    nu2_func = lambda t: (1 - s) * (nu2 / (1 - s)) ** (t / T)
    phi = Integration.two_pops(phi, xx, T, nu1_func, nu2_func, m12=m12, m21=m21)
    return Spectrum.from_phi(phi, ns, (xx, xx))


IM.__param_names__ = ['s', 'nu1', 'nu2', 'T', 'm12', 'm21']
