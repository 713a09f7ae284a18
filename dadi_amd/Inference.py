"""
Poisson and multinomial log-likelihoods of a data spectrum given a model
spectrum (dadi/Inference.py:466-575, 668-685).
"""
import logging

import numpy
from numpy import logical_and, logical_not
from scipy.special import gammaln

logger = logging.getLogger('Inference')


def _masked(a):
    if isinstance(a, numpy.ma.MaskedArray):
        return a
    return numpy.ma.masked_array(numpy.asarray(a, dtype=float))


def ll_per_bin(model, data, missing_model_cutoff=1e-6):
    """Poisson log-likelihood of each entry: -model + data*log(model) - lnGamma(data+1),
    masked where either spectrum is masked or the model is not positive
    (dadi/Inference.py:482-548)."""
    model = _masked(model)
    data = _masked(data)
    result = -model.data + data.data * numpy.ma.log(model) - gammaln(data + 1.)
    if numpy.all(result.mask == numpy.ma.getmaskarray(data)):
        return result
    not_data_mask = logical_not(numpy.ma.getmaskarray(data))
    data_sum = data.sum()
    missing = logical_and(model < 0, not_data_mask)
    if numpy.any(missing) and data[missing].sum() / data_sum > missing_model_cutoff:
        logger.warning('Model is < 0 where data is not masked.')
    missing = logical_and(model == 0, logical_and(data > 0, not_data_mask))
    if numpy.any(missing) and data[missing].sum() / data_sum > missing_model_cutoff:
        logger.warning('Model is 0 where data is neither masked nor 0.')
    return result


def ll(model, data):
    """Sum of ll_per_bin (dadi/Inference.py:466-480)."""
    return ll_per_bin(model, data).sum()


def optimal_sfs_scaling(model, data):
    """sum(data)/sum(model) over entries masked in neither (dadi/Inference.py:668-685)."""
    model = _masked(model)
    data = _masked(data)
    mask = numpy.ma.getmaskarray(model) | numpy.ma.getmaskarray(data)
    m = numpy.ma.masked_array(model.data, mask=mask)
    d = numpy.ma.masked_array(data.data, mask=mask)
    return d.sum() / m.sum()


def ll_multinom_per_bin(model, data):
    """Poisson ll per bin of the optimally scaled model (dadi/Inference.py:550-557)."""
    theta_opt = optimal_sfs_scaling(model, data)
    return ll_per_bin(theta_opt * model, data)


def ll_multinom(model, data):
    """Log-likelihood with the optimal theta scaling (dadi/Inference.py:559-575)."""
    return ll_multinom_per_bin(model, data).sum()
