"""
Probability densities for distributions of fitness effects
(dadi/DFE/PDFs.py).  Host-side weights; they only scale cached spectra.
"""
import numpy as np
import scipy.stats.distributions as ssd


def exponential(xx, params):
    """params: [scale]"""
    return ssd.expon.pdf(xx, scale=params[0])


def lognormal(xx, params):
    """params: [mu, sigma] (log scale)"""
    mu, sigma = params
    return ssd.lognorm.pdf(xx, sigma, scale=np.exp(mu))


def gamma(xx, params):
    """params: [alpha, beta] = [shape, scale]"""
    alpha, beta = params
    return ssd.gamma.pdf(xx, alpha, scale=beta)


def beta(xx, params):
    """params: [alpha, beta]"""
    alpha, beta = params
    return ssd.beta.pdf(xx, alpha, beta)


def normal(xx, mu, sigma):
    return ssd.norm.pdf(xx, loc=mu, scale=sigma)


def biv_lognormal(xx, yy, params):
    """Bivariate lognormal density on the grid xx (rows) x yy (columns);
    params (mu, sigma, rho) or (mu1, mu2, sigma1, sigma2, rho)
    (dadi/DFE/PDFs.py biv_lognormal_py)."""
    xx = np.atleast_1d(xx)
    yy = np.atleast_1d(yy)
    if len(params) == 3:
        mu, sigma, rho = params
        mu1 = mu2 = mu
        sigma1 = sigma2 = sigma
    elif len(params) == 5:
        mu1, mu2, sigma1, sigma2, rho = params
    else:
        raise ValueError('Parameter array for bivariate lognormal must have length 3 or 5.')
    delx = (np.log(xx[:, np.newaxis]) - mu1) / sigma1
    dely = (np.log(yy[np.newaxis, :]) - mu2) / sigma2
    norm = 2 * np.pi * sigma1 * sigma2 * np.sqrt(1. - rho ** 2) * np.outer(xx, yy)
    q = (delx ** 2 - 2. * rho * delx * dely + dely ** 2) / (1. - rho ** 2)
    return np.squeeze(np.exp(-q / 2.) / norm)


def biv_ind_gamma(xx, yy, params):
    """Product of independent gamma densities; params (alpha, beta) or
    (alpha1, alpha2, beta1, beta2); a trailing odd parameter is ignored."""
    xx = np.atleast_1d(xx)
    yy = np.atleast_1d(yy)
    if len(params) in (2, 3):
        alpha, beta = params[:2]
        a1 = a2 = alpha
        b1 = b2 = beta
    elif len(params) in (4, 5):
        a1, a2, b1, b2 = params[:4]
    else:
        raise ValueError('Parameter array for bivariate gamma must have length 2-5.')
    px = ssd.gamma.pdf(xx, a1, scale=b1)
    py = ssd.gamma.pdf(yy, a2, scale=b2)
    return np.squeeze(np.outer(px, py))
