"""
Demographic models with selection used to build DFE caches
(dadi/DFE/DemogSelModels.py).
"""
from .. import Numerics, PhiManip, Integration
from ..Spectrum_mod import Spectrum


def equil(params, ns, pts):
    """Equilibrium with selection gamma (dadi/DFE/DemogSelModels.py:11-24)."""
    gamma = params[0]
    xx = Numerics.default_grid(pts)
    phi = PhiManip.phi_1D(xx, gamma=gamma)
    return Spectrum.from_phi(phi, ns, (xx,))


def two_epoch_sel(params, ns, pts):
    """Size change to nu T ago with selection gamma (dadi/DFE/DemogSelModels.py:33-61)."""
    nu, T, gamma = params
    xx = Numerics.default_grid(pts)
    phi = PhiManip.phi_1D(xx, gamma=gamma)
    phi = Integration.one_pop(phi, xx, T, nu, gamma=gamma)
    return Spectrum.from_phi(phi, ns, (xx,))


two_epoch_sel.__param_names__ = ['nu', 'T', 'gamma']


def split_mig_sel(params, ns, pts):
    """Split with symmetric migration and selection gamma1 (pop 1 and
    ancestral), gamma2 (pop 2) (dadi/DFE/DemogSelModels.py:235-276)."""
    nu1, nu2, T, m, gamma1, gamma2 = params
    xx = Numerics.default_grid(pts)
    phi = PhiManip.phi_1D(xx, gamma=gamma1)
    phi = PhiManip.phi_1D_to_2D(xx, phi)
    phi = Integration.two_pops(phi, xx, T, nu1, nu2, m12=m, m21=m, gamma1=gamma1,
                               gamma2=gamma2)
    return Spectrum.from_phi(phi, ns, (xx, xx))


split_mig_sel.__param_names__ = ['nu1', 'nu2', 'T', 'm', 'gamma1', 'gamma2']


def split_mig_sel_single_gamma(params, ns, pts):
    nu1, nu2, T, m, gamma = params
    return split_mig_sel((nu1, nu2, T, m, gamma, gamma), ns, pts)


def IM_pre_sel(params, ns, pts):
    """Size change, split, exponential growth, migration, selection
    (dadi/DFE/DemogSelModels.py:63-126)."""
    nuPre, TPre, s, nu1, nu2, T, m12, m21, gamma1, gamma2 = params
    xx = Numerics.default_grid(pts)
    phi = PhiManip.phi_1D(xx, gamma=gamma1)
    phi = Integration.one_pop(phi, xx, TPre, nu=nuPre, gamma=gamma1)
    phi = PhiManip.phi_1D_to_2D(xx, phi)
    nu1_0 = nuPre * s
    nu2_0 = nuPre * (1 - s)
    nu1_func = lambda t: nu1_0 * (nu1 / nu1_0) ** (t / T)
    nu2_func = lambda t: nu2_0 * (nu2 / nu2_0) ** (t / T)
    phi = Integration.two_pops(phi, xx, T, nu1_func, nu2_func, m12=m12, m21=m21,
                               gamma1=gamma1, gamma2=gamma2)
    return Spectrum.from_phi(phi, ns, (xx, xx))


def IM_sel(params, ns, pts):
    """dadi/DFE/DemogSelModels.py:162-197."""
    s, nu1, nu2, T, m12, m21, gamma1, gamma2 = params
    return IM_pre_sel((1, 0, s, nu1, nu2, T, m12, m21, gamma1, gamma2), ns, pts)


def IM_sel_single_gamma(params, ns, pts):
    s, nu1, nu2, T, m12, m21, gamma = params
    return IM_sel((s, nu1, nu2, T, m12, m21, gamma, gamma), ns, pts)
