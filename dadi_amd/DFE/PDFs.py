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


def _biv_lognormal_params(params):
    """(mu1, mu2, sigma1, sigma2, rho) from (mu, sigma, rho) or the five values."""
    if len(params) == 3:
        return params[0], params[0], params[1], params[1], params[2]
    if len(params) == 5:
        return tuple(params)
    raise ValueError('Parameter array for bivariate lognormal must have length 3 or 5.')


def biv_lognormal(xx, yy, params):
    """Density of (X, Y) = (exp U, exp V) with (U, V) bivariate normal, on the
    grid xx (rows) x yy (columns); params (mu, sigma, rho) or (mu1, mu2,
    sigma1, sigma2, rho), as dadi/DFE/PDFs.py:41-43 takes them.

    Written through the standardised coordinates s = (ln x - mu1)/sigma1 and
    w = ((ln y - mu2)/sigma2 - rho s)/sqrt(1 - rho^2), which are independent
    standard normals: f(x, y) = phi(s) phi(w) / (sigma1 sigma2 sqrt(1-rho^2) x y),
    with phi the standard normal density (the Jacobian of (x, y) -> (s, w)).
    Scalar arguments give a 0-d result."""
    mu1, mu2, sigma1, sigma2, rho = _biv_lognormal_params(params)
    x = np.atleast_1d(np.asarray(xx, dtype=float))
    y = np.atleast_1d(np.asarray(yy, dtype=float))
    with np.errstate(divide='ignore', invalid='ignore'):
        s = (np.log(x) - mu1) / sigma1
        t = (np.log(y) - mu2) / sigma2
        root = np.sqrt(1. - rho * rho)
        w = (t[None, :] - rho * s[:, None]) / root
        jac = (sigma1 * x)[:, None] * (sigma2 * y)[None, :] * root
        dens = ssd.norm.pdf(s)[:, None] * ssd.norm.pdf(w) / jac
    return np.squeeze(dens)


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
