"""
Three- and four-population chain models: one ancestral population, a split
into two, a split into three (and four), each epoch with constant sizes and
symmetric migration.  They follow the pattern of
dadi/Demographics3D.py:10-74 (out_of_africa) and exercise every integration
dimensionality up to the one they end in.
"""
from . import Numerics, PhiManip, Integration
from .Spectrum_mod import Spectrum


def three_pop_chain_phi(params, pts):
    """phi of (nuA, TA, nu1, nu2, T2, m2, nu1b, nu2b, nu3, T3, m3)."""
    nuA, TA, nu1, nu2, T2, m2, nu1b, nu2b, nu3, T3, m3 = params
    xx = Numerics.default_grid(pts)
    phi = PhiManip.phi_1D(xx)
    phi = Integration.one_pop(phi, xx, TA, nu=nuA)
    phi = PhiManip.phi_1D_to_2D(xx, phi)
    phi = Integration.two_pops(phi, xx, T2, nu1, nu2, m12=m2, m21=m2)
    phi = PhiManip.phi_2D_to_3D_split_2(xx, phi)
    phi = Integration.three_pops(phi, xx, T3, nu1b, nu2b, nu3, m12=m3, m13=m3, m21=m3,
                                 m23=m3, m31=m3, m32=m3)
    return phi, xx


def three_pop_chain(params, ns, pts):
    phi, xx = three_pop_chain_phi(params, pts)
    return Spectrum.from_phi(phi, ns, (xx, xx, xx))


three_pop_chain.__param_names__ = ['nuA', 'TA', 'nu1', 'nu2', 'T2', 'm2', 'nu1b', 'nu2b',
                                   'nu3', 'T3', 'm3']


def four_pop_chain(params, ns, pts):
    """three_pop_chain followed by a split of population 3 into 3 and 4
    (phi_3D_to_4D with f1 = f2 = 0) and a 4-population epoch of
    (nu1c, nu2c, nu3c, nu4, T4, m4)."""
    nu1c, nu2c, nu3c, nu4, T4, m4 = params[11:]
    phi, xx = three_pop_chain_phi(params[:11], pts)
    phi = PhiManip.phi_3D_to_4D(phi, 0, 0, xx, xx, xx, xx)
    m = m4
    phi = Integration.four_pops(phi, xx, T4, nu1c, nu2c, nu3c, nu4,
                                m12=m, m13=m, m14=m, m21=m, m23=m, m24=m,
                                m31=m, m32=m, m34=m, m41=m, m42=m, m43=m)
    return Spectrum.from_phi(phi, ns, (xx,) * 4)
