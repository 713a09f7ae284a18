"""
Frequency spectra and the phi -> spectrum sampling integral.

``Spectrum.from_phi`` follows dadi/Spectrum_mod.py:1938-2062 with the
analytic samplers (1307-1348, 1509-1548, 1774-1871): phi is taken as
piecewise linear between grid points and integrated exactly against the
binomial sampling probabilities, whose interval integrals are differences of
regularised incomplete beta functions.  The beta tables are built on the host
with scipy (as dadi does, cached per sample size and grid); the contractions,
one axis at a time, run on the device (``dh_sample_axis``).

With ``het_ascertained`` or ``force_direct`` the direct samplers of
dadi/Spectrum_mod.py:1263-1279, 1350-1380, 1550-1592, 1594-1643 apply: the
trapezoid rule against the binomial sampling factors (times x(1-x) along a
heterozygote-ascertained axis), again one device contraction per axis
(``dh_contract_axis``) against host-built tables of factor x trapezoid weight.
"""
import logging

import numpy
from scipy.special import betainc, comb

from . import _lib
from . import engine

logger = logging.getLogger('Spectrum_mod')

_dbeta_cache = {}


def cached_dbeta(nx, xx):
    """(nx+1, len(xx)-1) tables of betainc(i+1, nx-i+1, x) and
    betainc(i+2, nx-i+1, x) differences over the grid intervals
    (dadi/Spectrum_mod.py:25-41)."""
    key = nx, tuple(xx)
    if key not in _dbeta_cache:
        x = numpy.minimum(numpy.maximum(xx, 0), 1.0)
        ii = numpy.arange(nx + 1)[:, None]
        b1 = betainc(ii + 1, nx - ii + 1, x[None, :])
        b2 = betainc(ii + 2, nx - ii + 1, x[None, :])
        if len(_dbeta_cache) > 512:
            _dbeta_cache.clear()
        _dbeta_cache[key] = (numpy.ascontiguousarray(b1[:, 1:] - b1[:, :-1]),
                             numpy.ascontiguousarray(b2[:, 1:] - b2[:, :-1]))
    return _dbeta_cache[key]


_dev_dbeta = {}


def _device_dbeta(ctx, m, xx):
    key = (m, xx.tobytes())
    hit = _dev_dbeta.get(key)
    if hit is None:
        d1, d2 = cached_dbeta(m, xx)
        hit = (ctx.to_dev(d1), ctx.to_dev(d2))
        if len(_dev_dbeta) > 256:
            _dev_dbeta.clear()
        _dev_dbeta[key] = hit
    return hit


_direct_tables = {}


def direct_table(m, xx, het=False):
    """(m+1, len(xx)) table: comb(m, k) x^k (1-x)^(m-k) [x (1-x)] (the factor of
    dadi/Spectrum_mod.py:1272-1274) times the trapezoid weight of each grid
    point, so that contracting phi with row k is the trapezoid integral of
    factor_k * phi."""
    xx = numpy.asarray(xx, dtype=numpy.float64)
    key = (m, xx.tobytes(), bool(het))
    hit = _direct_tables.get(key)
    if hit is None:
        k = numpy.arange(m + 1)[:, None]
        F = comb(m, k) * xx[None, :] ** k * (1 - xx[None, :]) ** (m - k)
        if het:
            F = F * (xx * (1 - xx))[None, :]
        d = numpy.diff(xx)
        w = numpy.zeros_like(xx)
        w[:-1] += d / 2.0
        w[1:] += d / 2.0
        hit = numpy.ascontiguousarray(F * w[None, :])
        if len(_direct_tables) > 256:
            _direct_tables.clear()
        _direct_tables[key] = hit
    return hit


def sample_device_direct(phi, ns, xxs, het_ascertained=None, ctx=None):
    """Direct (trapezoid) sample spectrum data of a 1-4D phi on the device;
    every axis may have its own grid.  Returns a device tensor of shape
    (n0+1, ..., n_{D-1}+1)."""
    import torch
    ctx = ctx or engine.context()
    if isinstance(phi, numpy.ndarray):
        cur = ctx.to_dev(numpy.ascontiguousarray(phi, dtype=numpy.float64))
    else:
        cur = phi.to(dtype=torch.float64).contiguous()
    D = cur.dim()
    if D > 4:
        raise NotImplementedError('direct sampling is provided for 1-4 populations, as in '
                                  'the reference')
    names = ('xx', 'yy', 'zz', 'aa')
    stream = ctx.stream()
    keep = []
    for k in range(D - 1, -1, -1):
        xx = numpy.asarray(xxs[k], dtype=numpy.float64)
        n = int(cur.shape[-1])
        if n != len(xx):
            raise ValueError('phi and grid sizes differ')
        m = int(ns[k])
        A = ctx.to_dev(direct_table(m, xx, het_ascertained == names[k]))
        P = cur.numel() // n
        out = torch.empty((m + 1,) + tuple(cur.shape[:-1]), dtype=torch.float64,
                          device=ctx.device)
        _lib.check(ctx.lib.dh_contract_axis(cur.data_ptr(), P, n, A.data_ptr(), m + 1,
                                            out.data_ptr(), stream), "dh_contract_axis")
        keep.append((cur, A))
        cur = out
    torch.cuda.current_stream(ctx.device).synchronize()
    del keep
    return cur


def sample_device(phi, ns, xx, ctx=None, batched=False):
    """Expected sample spectrum data of phi on the device; returns a device
    tensor of shape (n0+1, ..., n_{D-1}+1).  phi may be numpy or a device
    tensor; every axis uses grid xx.  batched: phi has a leading batch axis
    (B, n, ..., n) and the result is (B, n0+1, ..., n_{D-1}+1); each entry is
    computed with the same operations as for its phi alone."""
    import torch
    ctx = ctx or engine.context()
    if isinstance(phi, numpy.ndarray):
        cur = ctx.to_dev(numpy.ascontiguousarray(phi, dtype=numpy.float64))
    else:
        cur = phi.to(dtype=torch.float64).contiguous()
    D = cur.dim() - (1 if batched else 0)
    if len(ns) != D:
        raise ValueError('Dimensionality of phi and lengths of ns do not agree.')
    xx = numpy.asarray(xx, dtype=numpy.float64)
    # 1D: dadi clamps the grid before forming slopes (Spectrum_mod.py:1329)
    xs = numpy.minimum(numpy.maximum(xx, 0), 1.0) if D == 1 else xx
    xsd = ctx.grid(xs)
    stream = ctx.stream()
    keep = []
    for k in range(D - 1, -1, -1):
        n = int(cur.shape[-1])
        if n != len(xx):
            raise ValueError('phi and grid sizes differ')
        m = int(ns[k])
        P = cur.numel() // n
        d1, d2 = _device_dbeta(ctx, m, xx)
        work = torch.empty(2 * P * (n - 1), dtype=torch.float64, device=ctx.device)
        out = torch.empty((m + 1,) + tuple(cur.shape[:-1]), dtype=torch.float64,
                          device=ctx.device)
        _lib.check(ctx.lib.dh_sample_axis(cur.data_ptr(), P, n, xsd.data_ptr(), m,
                                          d1.data_ptr(), d2.data_ptr(), work.data_ptr(),
                                          out.data_ptr(), stream), "dh_sample_axis")
        keep.append((cur, work))
        cur = out
    if batched:   # the batch axis ends last: (n0+1, ..., n_{D-1}+1, B)
        cur = cur.permute((D,) + tuple(range(D))).contiguous()
    torch.cuda.current_stream(ctx.device).synchronize()
    del keep
    return cur


def sample_jobs(jobs, ctx=None):
    """Runs waiting engine.SampleJob samplings: the phis of one shape,
    sample sizes and grid go to the device in one copy and are sampled in
    one pass per axis (at most ~2^27 grid points per pass, which bounds the
    slope workspace to ~2 GB); results back to the host in one copy."""
    ctx = ctx or engine.context()
    groups = {}
    for j in jobs:
        groups.setdefault((j.phi.shape, j.ns, j.xx.tobytes()), []).append(j)
    for (shape, ns, _), grp in groups.items():
        per = max(1, (1 << 27) // int(numpy.prod(shape)))
        for c0 in range(0, len(grp), per):
            part = grp[c0:c0 + per]
            stack = ctx.to_dev(numpy.stack([j.phi for j in part]))
            data = sample_device(stack, ns, part[0].xx, ctx, batched=True).cpu().numpy()
            for i, j in enumerate(part):
                j.result = numpy.array(data[i])
            del stack


class Spectrum(numpy.ma.masked_array):
    """A frequency spectrum: a masked array whose corner entries ('absent in
    all' and 'fixed in all') are masked by default (dadi/Spectrum_mod.py:74-182)."""

    def __new__(subtype, data, mask=numpy.ma.nomask, mask_corners=True, data_folded=None,
                check_folding=True, dtype=float, copy=True, fill_value=numpy.nan,
                keep_mask=True, shrink=True, pop_ids=None, extrap_x=None):
        data = numpy.asanyarray(data)
        if mask is numpy.ma.nomask:
            mask = numpy.ma.make_mask_none(data.shape)
        subarr = numpy.ma.masked_array(data, mask=mask, dtype=dtype, copy=copy,
                                       fill_value=fill_value, keep_mask=True, shrink=False)
        subarr = subarr.view(subtype)
        if hasattr(data, 'folded'):
            subarr.folded = data.folded
        else:
            subarr.folded = bool(data_folded)
        if mask_corners:
            subarr.mask_corners()
        subarr.pop_ids = pop_ids
        subarr.extrap_x = extrap_x
        return subarr

    def __array_finalize__(self, obj):
        if obj is None:
            return
        numpy.ma.masked_array.__array_finalize__(self, obj)
        self.folded = getattr(obj, 'folded', False)
        self.pop_ids = getattr(obj, 'pop_ids', None)
        self.extrap_x = getattr(obj, 'extrap_x', None)

    def mask_corners(self):
        """Mask the 'seen in 0 samples' and 'seen in all samples' entries."""
        self.mask[(0,) * self.ndim] = True
        self.mask[tuple(numpy.array(self.shape) - 1)] = True

    @property
    def sample_sizes(self):
        return numpy.asarray(self.shape) - 1

    def log(self):
        """numpy.ma.log, keeping the Spectrum type."""
        return numpy.ma.log(self)

    @staticmethod
    def from_phi(phi, ns, xxs, mask_corners=True, pop_ids=None, admix_props=None,
                 het_ascertained=None, force_direct=False):
        """Expected spectrum of sample sizes ``ns`` from the density ``phi`` on
        grids ``xxs`` (dadi/Spectrum_mod.py:1938-2062): the analytic
        (piecewise-linear phi) samplers, which need one grid for every axis,
        or with het_ascertained / force_direct the direct trapezoid samplers
        (1-4 populations).  admix_props is not provided."""
        if admix_props:
            raise NotImplementedError('the admix_props samplers are not provided by the GPU '
                                      'backend; implement admixture with PhiManip')
        ndim = phi.ndim if hasattr(phi, 'ndim') else phi.dim()
        if not ndim == len(ns) == len(xxs):
            raise ValueError('Dimensionality of phi and lengths of ns and xxs do not all agree.')
        if het_ascertained and het_ascertained not in ['xx', 'yy', 'zz']:
            raise ValueError("If used, het_ascertained must be 'xx', 'yy', or 'zz'.")
        if ndim > 5:
            raise ValueError('Only implemented for dimensions 1-5.')
        if het_ascertained or force_direct:
            data = sample_device_direct(phi, ns, xxs, het_ascertained).cpu().numpy()
            fs = Spectrum(data, mask_corners=mask_corners, pop_ids=pop_ids)
            fs.extrap_x = xxs[0][1]
            return fs
        xx = numpy.asarray(xxs[0], dtype=numpy.float64)
        for other in xxs[1:]:
            if not (len(other) == len(xx) and numpy.allclose(other, xx)):
                raise ValueError('Must have xx==yy to use linear algebra calculation of FS '
                                 'from phi.')
        co = engine.current_coalescer()
        if co is not None and isinstance(phi, numpy.ndarray):
            # a batching worker: sampled with every other waiting phi
            job = engine.SampleJob(numpy.ascontiguousarray(phi, dtype=numpy.float64), ns, xx)
            co.submit(job)
            data = job.result
        else:
            data = sample_device(phi, ns, xx).cpu().numpy()
        fs = Spectrum(data, mask_corners=mask_corners, pop_ids=pop_ids)
        fs.extrap_x = xxs[0][1]
        for x in xxs[1:]:
            if not x[1] == fs.extrap_x:
                logger.warning('Spectrum calculated from phi different grids for '
                               'different dimensions. Extrapolation may fail.')
        return fs
