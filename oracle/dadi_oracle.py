"""
CPU oracle for the dadi ADI hot path.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()``
and by bench.py's CPU-baseline leg, as the checker.  The dadi_amd package never
imports it.

Contents, each a restatement of the reference function it cites:

* ``default_grid``            dadi/Numerics.py:215-233
* ``phi_1D_snm``/``phi_1D_genic``/``phi_1D``   dadi/PhiManip.py:166-195, 110-164, 12-108
* ``phi_1D_to_2D``            dadi/PhiManip.py:214-234
* ``phi_2D_to_3D_admix``/``_split_1``/``_split_2``/``phi_3D_to_4D``/``phi_4D_to_5D``
                              dadi/PhiManip.py:236-300, 302-370, 372-495
* ``remove_pop``/``filter_pops``  dadi/PhiManip.py:1036-1072
* ``compute_dt``              dadi/Integration.py:147-176
* ``inject_value``            dadi/Integration.py:70-145
* ``integrate_const``         dadi/Integration.py:986-1242 (constant parameters)
* ``integrate_tv``            dadi/Integration.py:178-937 (time-varying loops)
* ``one_pop_X``               dadi/Integration.py:1244-1352 (X chromosome, constant)
* ``from_phi``                dadi/Spectrum_mod.py:25-41, 1307-1348, 1509-1548,
                              1774-1871, 1938-2062
* ``from_phi_exact``          the same formula with compensated (once-rounded) sums
* ``from_phi_direct``         dadi/Spectrum_mod.py:1263-1279, 1350-1380, 1550-1592,
                              1594-1643 (trapezoid samplers, heterozygote ascertainment)
* ``quadratic_extrap``/``linear_extrap``/``extrapolate``
                              dadi/Numerics.py:334-475, 644-669
* ``ll``/``ll_multinom``      dadi/Inference.py:466-575, 668-685
* ``marginalize``             the entry sums of Spectrum.marginalize

The sweeps run either through this repository's C restatement
(``oracle/libdadi_oracle.so``, backend ``"oracle"``) or through the reference C
compiled from source (``oracle/_ref/libdadi_ref.so``, backend ``"ref"``), so the
two can be compared bit for bit.
"""
import ctypes
import os

import numpy as np
from numpy import newaxis as nuax
from scipy.special import betainc, comb, gammaln

try:
    _trapezoid = np.trapezoid
except AttributeError:  # numpy < 2
    _trapezoid = np.trapz

HERE = os.path.dirname(os.path.abspath(__file__))
ORC_TV, ORC_PRE, ORC_PRE1 = 0, 1, 2

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def tridiag(a, b, c, r):
    """Thomas solve through the C restatement of dadi/tridiag.c:14-40."""
    n = len(r)
    a, b, c, r = (np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, c, r))
    u = np.empty(n)
    gam = np.empty(n)
    lib().orc_tridiag(_ptr(a), _ptr(b), _ptr(c), _ptr(r), _ptr(u), n, _ptr(gam))
    return u


_lib = None
_ref = None


def lib():
    """The C restatement (built by ``make -C oracle``)."""
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "libdadi_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/libdadi_oracle.so is not built (make -C oracle)")
        L = ctypes.CDLL(path)
        L.orc_sweep.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                _dp, ctypes.c_double, _dp, ctypes.c_double, ctypes.c_double,
                                ctypes.c_double, ctypes.c_double]
        L.orc_sweep.restype = None
        L.orc_tridiag.argtypes = [_dp, _dp, _dp, _dp, _dp, ctypes.c_int, _dp]
        L.orc_tridiag.restype = None
        L.orc_run.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_long,
                              _dp, _dp, _dp, _ip, _ip, ctypes.c_double]
        L.orc_run.restype = None
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(os.path.join(HERE, "_ref", "libdadi_ref.so"))


def ref():
    """The reference C kernels compiled from source (``make -C oracle ref``)."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libdadi_ref.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/_ref/libdadi_ref.so is not built")
        _ref = ctypes.CDLL(path)
    return _ref


# ---------------------------------------------------------------------------
# Grid and initial densities
# ---------------------------------------------------------------------------
def default_grid(pts, crwd=8.0):
    """dadi/Numerics.py:215-233 (exponential_grid)."""
    unif = np.linspace(-1, 1, pts)
    grid = 1.0 / (1.0 + np.exp(-crwd * unif))
    grid = (grid - grid[0]) / (grid[-1] - grid[0])
    return grid


def phi_1D_snm(xx, nu=1.0, theta0=1.0, beta=1):
    """dadi/PhiManip.py:166-195."""
    if xx[0] == 0:
        phi = 0 * xx
        phi[1:] = nu * theta0 / xx[1:]
        phi[0] = phi[1]
    else:
        phi = nu * theta0 / xx
    return phi * 4.0 * beta / (beta + 1.0) ** 2


def phi_1D_genic(xx, nu=1.0, theta0=1.0, gamma=0, beta=1):
    """dadi/PhiManip.py:110-164."""
    if gamma == 0:
        return phi_1D_snm(xx, nu, theta0, beta=beta)
    gamma = gamma * 4.0 * beta / (beta + 1.0) ** 2
    exp = np.exp
    if xx[0] == 0 and xx[-1] == 1:
        phi = 0 * xx
        if gamma > -300:
            phi[1:-1] = 1.0 / (xx[1:-1] * (1 - xx[1:-1])) \
                * (1 - exp(-2 * gamma * (1 - xx[1:-1]))) / (1 - exp(-2 * gamma))
        else:
            phi[1:-1] = 1.0 / (xx[1:-1] * (1 - xx[1:-1])) * exp(2 * gamma * xx[1:-1])
    else:
        if gamma > -300:
            phi = 1.0 / (xx * (1 - xx)) * (1 - exp(-2 * gamma * (1 - xx))) / (1 - exp(-2 * gamma))
        else:
            phi = 1.0 / (xx * (1 - xx)) * exp(2 * gamma * xx)
    if xx[0] == 0:
        phi[0] = phi[1]
    if xx[-1] == 1:
        if gamma < 300:
            limit = 2 * gamma * exp(2 * gamma) / (exp(2 * gamma) - 1)
        else:
            limit = 2 * gamma
        phi[-1] = limit
    return phi * nu * theta0 * 4.0 * beta / (beta + 1.0) ** 2


def phi_1D(xx, nu=1.0, theta0=1.0, gamma=0, h=0.5, beta=1):
    """dadi/PhiManip.py:12-108 (h == 0.5 branch; the quad branch is not restated)."""
    if h != 0.5:
        raise NotImplementedError("oracle restates only the h == 0.5 branch of phi_1D")
    return phi_1D_genic(xx, nu, theta0, gamma, beta=beta)


def phi_1D_to_2D(xx, phi_1D):
    """dadi/PhiManip.py:214-234."""
    pts = len(xx)
    phi_2D = np.zeros((pts, pts))
    for ii in range(1, pts - 1):
        phi_2D[ii, ii] = phi_1D[ii] * 2 / (xx[ii + 1] - xx[ii - 1])
    return phi_2D


def _admixture_intermediates(phi, ad_z, zz):
    """dadi/PhiManip.py:268-300."""
    upper_z_index = np.searchsorted(zz, ad_z)
    upper_z_index = np.minimum(upper_z_index, len(zz) - 1)
    upper_z_index = np.maximum(upper_z_index, 1)
    lower_z_index = upper_z_index - 1
    upper_z = zz[upper_z_index]
    lower_z = zz[lower_z_index]
    delz0 = zz[lower_z_index] - zz[lower_z_index - 1]
    delz0 = np.where(lower_z_index == 0, 0, delz0)
    delz1 = zz[upper_z_index] - zz[lower_z_index]
    delz1 = np.where(upper_z_index == 0, 0, delz1)
    delz2 = zz[(upper_z_index + 1) % len(zz)] - zz[upper_z_index]
    delz2 = np.where(upper_z_index == len(zz) - 1, 0, delz2)
    frac_lower = (upper_z - ad_z) / (upper_z - lower_z)
    frac_upper = (ad_z - lower_z) / (upper_z - lower_z)
    norm = 2 * phi / (frac_lower * delz0 + delz1 + frac_upper * delz2)
    return lower_z_index, upper_z_index, frac_lower, frac_upper, norm


def _scatter_new_axis(phi, ad, zz):
    lo, up, fl, fu, norm = _admixture_intermediates(phi, ad, zz)
    idx = np.indices(phi.shape, sparse=True)
    out = np.zeros(phi.shape + (len(zz),))
    out[tuple(idx) + (lo,)] = fl * norm
    out[tuple(idx) + (up,)] += fu * norm
    return out


def phi_2D_to_3D_admix(phi, f1, xx, yy, zz):
    """dadi/PhiManip.py:302-313, 372-409."""
    ad_z = f1 * xx[:, nuax] + (1 - f1) * yy[nuax, :]
    return _scatter_new_axis(phi, ad_z, zz)


def phi_2D_to_3D_split_2(xx, phi_2D):
    """dadi/PhiManip.py:236-250."""
    return phi_2D_to_3D_admix(phi_2D, 0, xx, xx, xx)


def phi_2D_to_3D_split_1(xx, phi_2D):
    """dadi/PhiManip.py:252-266."""
    return phi_2D_to_3D_admix(phi_2D, 1, xx, xx, xx)


def _trapz_axis(yy, xx, axis):
    """dadi/Numerics.py:303-332 with a grid: sum of dx (y_i + y_{i+1}) / 2."""
    yy = np.asarray(yy, dtype=np.float64)
    dx = np.diff(xx)
    shape = [1] * yy.ndim
    shape[axis] = len(dx)
    lo = np.take(yy, np.arange(len(xx) - 1), axis=axis)
    hi = np.take(yy, np.arange(1, len(xx)), axis=axis)
    return np.sum(dx.reshape(shape) * (hi + lo) / 2.0, axis=axis)


def remove_pop(phi, xx, popnum):
    """dadi/PhiManip.py:1036-1050: integrate population popnum (from 1) out."""
    return _trapz_axis(phi, xx, popnum - 1)


def filter_pops(phi, xx, tokeep):
    """dadi/PhiManip.py:1052-1072: remove every population not in tokeep,
    highest number first."""
    for pop in sorted(set(range(1, np.ndim(phi) + 1)) - set(tokeep))[::-1]:
        phi = remove_pop(phi, xx, pop)
    return phi


def phi_3D_to_4D(phi, f1, f2, xx, yy, zz, aa):
    """dadi/PhiManip.py:315-331, 412-451."""
    ad_w = f1 * xx[:, nuax, nuax] + f2 * yy[nuax, :, nuax] + (1 - f1 - f2) * zz[nuax, nuax, :]
    return _scatter_new_axis(phi, ad_w, aa)


def phi_4D_to_5D(phi, f1, f2, f3, xx, yy, zz, aa, bb):
    """dadi/PhiManip.py:333-350, 453-495."""
    ad_w = f1 * xx[:, nuax, nuax, nuax] + f2 * yy[nuax, :, nuax, nuax] \
        + f3 * zz[nuax, nuax, :, nuax] + (1 - f1 - f2 - f3) * aa[nuax, nuax, nuax, :]
    return _scatter_new_axis(phi, ad_w, bb)


# ---------------------------------------------------------------------------
# Time step and injection
# ---------------------------------------------------------------------------
timescale_factor = 1e-3


def compute_dt(nu, ms, gamma, h):
    """dadi/Integration.py:147-176 (use_old_timestep False)."""
    maxVM = max(0.25 / nu, sum(ms),
                abs(gamma) * 2 * max(np.abs(h + (1 - 2 * h) * 0.5) * 0.5 * (1 - 0.5),
                                     np.abs(h + (1 - 2 * h) * 0.25) * 0.25 * (1 - 0.25)))
    if maxVM > 0:
        dt = timescale_factor / maxVM
    else:
        dt = np.inf
    if dt == 0:
        raise ValueError("Timestep is zero.")
    return dt


def inject_value(D, dt, xx, theta0):
    """Injected mass at index 1 of one axis, dadi/Integration.py:70-145:
    dt/x1 * theta0/2 * 2**D / ((x2 - x0) * x1**(D-1)), evaluated left to right."""
    den = (xx[2] - xx[0])
    for _ in range(D - 1):
        den = den * xx[1]
    return dt / xx[1] * theta0 / 2 * (2 ** D) / den


def _inject(phi, D, dt, xx, theta0, on):
    val = inject_value(D, dt, xx, theta0)
    for k in range(D):
        if on[k]:
            idx = tuple(1 if q == k else 0 for q in range(D))
            phi[idx] += val


# ---------------------------------------------------------------------------
# Sweeps: generic C restatement or the reference's own functions
# ---------------------------------------------------------------------------
def sweep(order, phi, axis, xx, nu, mig, gamma, h, beta, dt):
    """One implicit half-step through the C restatement, in place."""
    assert phi.flags.c_contiguous and phi.dtype == np.float64
    mig = np.ascontiguousarray(mig, dtype=np.float64).reshape(-1)
    if mig.size == 0:
        mig = np.zeros(1)
    lib().orc_sweep(order, _ptr(phi), phi.ndim, phi.shape[0], axis, _ptr(xx), float(nu),
                    _ptr(mig), float(gamma), float(h), float(beta), float(dt))


_c_int, _c_d = ctypes.c_int, ctypes.c_double


def ref_sweep_tv(phi, axis, xx, nu, mig, gamma, h, beta, dt, delj=0):
    """The reference's implicit_<D>D<axis> (dadi/integration{1..5}D.c) via the
    argument order of dadi/integration_c.pyx:68-196; delj: use_delj_trick."""
    R = ref()
    D, n = phi.ndim, phi.shape[0]
    P = _ptr(phi)
    X = _ptr(xx)
    nu, gamma, h, dt = _c_d(nu), _c_d(gamma), _c_d(h), _c_d(dt)
    ms = [_c_d(m) for m in mig]
    N = _c_int(n)
    z = _c_int(0)
    dj = _c_int(1 if delj else 0)
    if D == 1:
        R.implicit_1Dx(P, X, nu, gamma, h, _c_d(beta), dt, N, dj)
    elif D == 2:
        f = [R.implicit_2Dx, R.implicit_2Dy][axis]
        f(P, X, X, nu, ms[0], gamma, h, dt, N, N, dj, z, N)
    elif D == 3:
        f = [R.implicit_3Dx, R.implicit_3Dy, R.implicit_3Dz][axis]
        f(P, X, X, X, nu, ms[0], ms[1], gamma, h, dt, N, N, N, dj, z, N)
    elif D == 4:
        f = [R.implicit_4Dx, R.implicit_4Dy, R.implicit_4Dz, R.implicit_4Da][axis]
        f(P, X, X, X, X, nu, ms[0], ms[1], ms[2], gamma, h, dt, N, N, N, N, dj)
    elif D == 5:
        f = [R.implicit_5Dx, R.implicit_5Dy, R.implicit_5Dz, R.implicit_5Da, R.implicit_5Db][axis]
        f(P, X, X, X, X, X, nu, ms[0], ms[1], ms[2], ms[3], gamma, h, dt, N, N, N, N, N, dj)
    else:
        raise ValueError(D)


def _Vfunc(x, nu, beta=1):
    return 1. / nu * x * (1 - x) * (beta + 1.) ** 2 / (4. * beta)


def _Mfunc(x, others, ms, gamma, h):
    """dadi/Integration.py:943-950 generalised: sum m*(y-x) left to right, then selection."""
    sel = gamma * 2 * (h + (1 - 2 * h) * x) * x * (1 - x)
    if not others:
        return sel
    acc = ms[0] * (others[0] - x)
    for y, m in zip(others[1:], ms[1:]):
        acc = acc + m * (y - x)
    return acc + sel


def compute_delj(dx, MInt, VInt):
    """Chang and Cooper's delta of the numpy paths, dadi/Integration.py:960-984
    (dx and VInt already broadcast against MInt): NaN and inf become 1/2."""
    wj = 2 * MInt * dx
    with np.errstate(all='ignore'):
        epsj = np.exp(wj / VInt)
        delj = (-epsj * wj + epsj * VInt - VInt) / (wj - epsj * wj)
    delj = np.where(np.isnan(delj), 0.5, delj)
    delj = np.where(np.isinf(delj), 0.5, delj)
    return delj


def precalc_arrays(D, xx, axis, nu, mig, gamma, h, beta=1, delj=False):
    """numpy a, b (without 1/dt) and c arrays of the constant-parameter paths,
    dadi/Integration.py:997-1015 (1D), 1056-1095 (2D), 1133-1221 (3D); delj:
    use_delj_trick."""
    n = len(xx)
    shp = [1] * D

    def along(v, k):
        s = [1] * D
        s[k] = len(v)
        return v.reshape(s)
    others = [along(xx, k) for k in range(D) if k != axis]
    xI = along((xx[:-1] + xx[1:]) / 2, axis)
    xN = along(xx, axis)
    V = _Vfunc(xx, nu, beta)
    M = _Mfunc(xN, others, mig, gamma, h)
    MInt = _Mfunc(xI, others, mig, gamma, h)
    dx = np.diff(xx)
    dfact = np.zeros(len(dx) + 1)
    dfact[1:-1] = 2 / (dx[:-1] + dx[1:])
    dfact[0] = 2 / dx[0]
    dfact[-1] = 2 / dx[-1]
    if delj:
        VInt = _Vfunc((xx[:-1] + xx[1:]) / 2, nu, beta)
        delj = compute_delj(along(dx, axis), MInt, along(VInt, axis))
    else:
        delj = 0.5
    shape = (n,) * D
    a, b, c = [np.zeros(shape) for _ in range(3)]
    sl = lambda s: tuple(s if k == axis else slice(None) for k in range(D))
    lo, hi = sl(slice(None, -1)), sl(slice(1, None))
    a[hi] += along(dfact[1:], axis) * (-MInt * delj - along(V[:-1], axis) / (2 * along(dx, axis)))
    c[lo] += along(dfact[:-1], axis) * (MInt * (1 - delj) - along(V[1:], axis) / (2 * along(dx, axis)))
    b[lo] += along(dfact[:-1], axis) * (MInt * delj + along(V[:-1], axis) / (2 * along(dx, axis)))
    b[hi] += along(dfact[1:], axis) * (-MInt * (1 - delj) + along(V[1:], axis) / (2 * along(dx, axis)))
    M = np.broadcast_to(M, shape)
    first = (0,) * D
    last = (n - 1,) * D
    if M[first] <= 0:
        b[first] += (0.5 / nu - M[first]) * 2 / dx[0]
    if M[last] >= 0:
        b[last] += -(-0.5 / nu - M[last]) * 2 / dx[-1]
    return a, b, c


def ref_sweep_pre(phi, axis, arrays, dt):
    """The reference's implicit_precalc_{2,3}D* (dadi/integration2D.c:162-217,
    dadi/integration3D.c:241-345) or, for 1D, tridiag(a, b+1/dt, c, phi/dt)
    (dadi/Integration.py:1027-1034)."""
    R = ref()
    a, b, c = arrays
    D, n = phi.ndim, phi.shape[0]
    N = _c_int(n)
    z = _c_int(0)
    if D == 1:
        bb = b + 1 / dt
        r = phi / dt
        u = np.empty(n)
        R.tridiag(_ptr(a), _ptr(bb), _ptr(c), _ptr(r), _ptr(u), N)
        phi[:] = u
    elif D == 2:
        f = [R.implicit_precalc_2Dx, R.implicit_precalc_2Dy][axis]
        f(_ptr(phi), _ptr(a), _ptr(b), _ptr(c), _c_d(dt), N, N, z, N)
    elif D == 3:
        f = [R.implicit_precalc_3Dx, R.implicit_precalc_3Dy, R.implicit_precalc_3Dz][axis]
        f(_ptr(phi), _ptr(a), _ptr(b), _ptr(c), _c_d(dt), N, N, N, z, N)
    else:
        raise ValueError("no constant-parameter path for D=%d in the reference" % D)


# ---------------------------------------------------------------------------
# Drivers
# ---------------------------------------------------------------------------
def _check(T, initial_t):
    if T - initial_t < 0:
        raise ValueError("Final integration time T is less than initial_time.")


def _Vfunc_X(x, nu, beta):
    """dadi/Integration.py:1244-1245."""
    return 1. / nu * x * (1 - x) * (2 * beta + 4.) * (beta + 1.) / (9. * beta)


def _Mfunc1D_X(x, gamma, h, beta):
    """dadi/Integration.py:1246-1247."""
    return gamma * 4. / 3. * (0.5 + h + x * (1 - 2 * h)) * x * (1 - x)


def one_pop_X(phi, xx, T, nu=1, gamma=0, h=0.5, beta=1, alpha=1, theta0=1.0, initial_t=0,
              dt=None, delj=False):
    """Constant-parameter X-chromosome integration, dadi/Integration.py:1256-1352
    (numpy coefficient arrays as the reference forms them, the C Thomas
    solver).  dt: the time step (default: the reference's _compute_dt)."""
    phi = np.array(phi, dtype=np.float64, copy=True)
    if T - initial_t == 0:
        return phi
    xx = np.asarray(xx, dtype=np.float64)
    M = _Mfunc1D_X(xx, gamma, h, beta)
    MInt = _Mfunc1D_X((xx[:-1] + xx[1:]) / 2, gamma, h, beta)
    V = _Vfunc_X(xx, nu, beta)
    dx = np.diff(xx)
    dfactor = np.zeros(len(dx) + 1)
    dfactor[1:-1] = 2 / (dx[:-1] + dx[1:])
    dfactor[0] = 2 / dx[0]
    dfactor[-1] = 2 / dx[-1]
    if delj:
        delj = compute_delj(dx, MInt, _Vfunc_X((xx[:-1] + xx[1:]) / 2, nu, beta))
    else:
        delj = 0.5
    a = np.zeros(phi.shape)
    a[1:] += dfactor[1:] * (-MInt * delj - V[:-1] / (2 * dx))
    c = np.zeros(phi.shape)
    c[:-1] += -dfactor[:-1] * (-MInt * (1 - delj) + V[1:] / (2 * dx))
    b = np.zeros(phi.shape)
    b[:-1] += -dfactor[:-1] * (-MInt * delj - V[:-1] / (2 * dx))
    b[1:] += dfactor[1:] * (-MInt * (1 - delj) + V[1:] / (2 * dx))
    if M[0] <= 0:
        b[0] += (0.5 / nu - M[0]) * 2 / dx[0]
    if M[-1] >= 0:
        b[-1] += -(-0.5 / nu - M[-1]) * 2 / dx[-1]
    if dt is None:
        dt = compute_dt(nu, [0], gamma, h)
    factor = 2. / (1. + 2. * beta) * (1. / (alpha + 1.) + beta)
    current_t = initial_t
    while current_t < T:
        this_dt = min(dt, T - current_t)
        phi[1] += this_dt / xx[1] * theta0 / 2. * factor * 2. / (xx[2] - xx[0])
        r = phi / this_dt
        phi = tridiag(a, b + 1. / this_dt, c, r)
        current_t += this_dt
    return phi


def integrate_const(phi, xx, T, nus, migs, gammas, hs, theta0=1.0, initial_t=0.0,
                    frozen=None, nomut=None, beta=1, backend="oracle", delj=False):
    """Constant-parameter drivers, dadi/Integration.py:986-1242.

    nus/gammas/hs: per population; migs[k]: rates into population k from the
    others in increasing order.  D = phi.ndim in {1, 2, 3}.  delj
    (use_delj_trick) needs backend "ref" (numpy arrays + reference C).
    """
    if delj and backend != "ref":
        raise ValueError("the delta trick is checked through the reference backend")
    phi = np.array(phi, dtype=np.float64, copy=True, order="C")
    D = phi.ndim
    frozen = [False] * D if frozen is None else list(frozen)
    nomut = [False] * D if nomut is None else list(nomut)
    if T - initial_t == 0:
        return phi
    _check(T, initial_t)
    if np.any(np.less([T] + list(nus) + [m for ms in migs for m in ms] + [theta0], 0)):
        raise ValueError("A time, population size, migration rate, or theta0 is < 0.")
    if np.any(np.equal(nus, 0)):
        raise ValueError("A population size is 0.")
    order = ORC_PRE1 if D == 1 else ORC_PRE
    arrays = None
    if backend == "ref":
        arrays = [precalc_arrays(D, xx, k, nus[k], migs[k], gammas[k], hs[k], beta, delj)
                  for k in range(D)]
    dt = min(compute_dt(nus[k], migs[k] if D > 1 else [0], gammas[k], hs[k]) for k in range(D))
    on = [not frozen[k] and not nomut[k] for k in range(D)]
    current_t = initial_t
    while current_t < T:
        this_dt = min(dt, T - current_t)
        _inject(phi, D, this_dt, xx, theta0, on)
        for k in range(D):
            if frozen[k]:
                continue
            if backend == "ref":
                ref_sweep_pre(phi, k, arrays[k], this_dt)
            else:
                sweep(order, phi, k, xx, nus[k], migs[k], gammas[k], hs[k], beta, this_dt)
        current_t += this_dt
    return phi


def _f(v):
    return v if callable(v) else (lambda t, v=v: v)


def integrate_tv(phi, xx, T, nus, migs, gammas, hs, theta0=1.0, initial_t=0.0,
                 frozen=None, nomut=None, beta=1, backend="oracle", delj=False):
    """Time-varying drivers, dadi/Integration.py:219-258, 329-387, 471-554,
    611-727, 786-937.  Parameters may be floats or 1-argument callables of t.
    The first population's beta is used in 1D only.  delj (use_delj_trick)
    needs backend "ref"."""
    if delj and backend != "ref":
        raise ValueError("the delta trick is checked through the reference backend")
    phi = np.array(phi, dtype=np.float64, copy=True, order="C")
    D = phi.ndim
    frozen = [False] * D if frozen is None else list(frozen)
    nomut = [False] * D if nomut is None else list(nomut)
    if T - initial_t == 0:
        return phi
    _check(T, initial_t)
    nu_f = [_f(v) for v in nus]
    g_f = [_f(v) for v in gammas]
    h_f = [_f(v) for v in hs]
    m_f = [[_f(v) for v in ms] for ms in migs]
    th_f = _f(theta0)
    be_f = _f(beta)
    t = initial_t
    nu = [f(t) for f in nu_f]
    ga = [f(t) for f in g_f]
    hh = [f(t) for f in h_f]
    mm = [[f(t) for f in fs] for fs in m_f]
    on = [not frozen[k] and not nomut[k] for k in range(D)]
    while t < T:
        dt = min(compute_dt(nu[k], mm[k] if D > 1 else [0], ga[k], hh[k]) for k in range(D))
        this_dt = min(dt, T - t)
        next_t = t + this_dt
        nu = [f(next_t) for f in nu_f]
        ga = [f(next_t) for f in g_f]
        hh = [f(next_t) for f in h_f]
        mm = [[f(next_t) for f in fs] for fs in m_f]
        th = th_f(next_t)
        bt = be_f(next_t)
        if np.any(np.less([T] + nu + [m for ms in mm for m in ms] + [th], 0)):
            raise ValueError("A time, population size, migration rate, or theta0 is < 0.")
        if np.any(np.equal(nu, 0)):
            raise ValueError("A population size is 0.")
        _inject(phi, D, this_dt, xx, th, on)
        for k in range(D):
            if frozen[k]:
                continue
            if backend == "ref":
                ref_sweep_tv(phi, k, xx, nu[k], mm[k], ga[k], hh[k], bt, this_dt, delj)
            else:
                sweep(ORC_TV, phi, k, xx, nu[k], mm[k], ga[k], hh[k], bt, this_dt)
        t = next_t
    return phi


# ---------------------------------------------------------------------------
# Sampling, extrapolation, likelihood
# ---------------------------------------------------------------------------
def cached_dbeta(nx, xx):
    """dadi/Spectrum_mod.py:25-41 (without the cache)."""
    xx = np.minimum(np.maximum(xx, 0), 1.0)
    dbeta1 = np.empty((nx + 1, len(xx) - 1))
    dbeta2 = np.empty((nx + 1, len(xx) - 1))
    for ii in range(0, nx + 1):
        b = betainc(ii + 1, nx - ii + 1, xx)
        dbeta1[ii] = b[1:] - b[:-1]
        b = betainc(ii + 2, nx - ii + 1, xx)
        dbeta2[ii] = b[1:] - b[:-1]
    return dbeta1, dbeta2


def _from_phi_1D(n, xx, phi):
    """dadi/Spectrum_mod.py:1307-1348."""
    data = np.zeros(n + 1)
    xx = np.minimum(np.maximum(xx, 0), 1.0)
    s = (phi[1:] - phi[:-1]) / (xx[1:] - xx[:-1])
    c1 = (phi[:-1] - s * xx[:-1]) / (n + 1)
    for d in range(0, n + 1):
        c2 = s * (d + 1) / ((n + 1) * (n + 2))
        beta1 = betainc(d + 1, n - d + 1, xx)
        beta2 = betainc(d + 2, n - d + 1, xx)
        entries = c1 * (beta1[1:] - beta1[:-1]) + c2 * (beta2[1:] - beta2[:-1])
        data[d] = np.sum(entries)
    return data


def _contract_last(phi, n, xx):
    """One axis of the 2-5D linalg samplers, dadi/Spectrum_mod.py:1527-1532 and
    1789-1803: integrate the piecewise-linear phi along its last axis against
    the binomial sampling weights.  Returns shape phi.shape[:-1] + (n+1,)."""
    db1, db2 = cached_dbeta(n, xx)
    s = (phi[..., 1:] - phi[..., :-1]) / (xx[1:] - xx[:-1])
    c1 = (phi[..., :-1] - s * xx[:-1]) / (n + 1)
    term1 = np.dot(c1, db1.T)
    term2 = np.dot(s, db2.T)
    term2 *= np.arange(1, n + 2) / ((n + 1) * (n + 2))
    return term1 + term2


def _two_sum(a, b):
    """Knuth's TwoSum: a + b = s + e exactly."""
    s = a + b
    bb = s - a
    return s, (a - (s - bb)) + (b - bb)


def _split(a):
    """Veltkamp's split of a double into two 26-bit halves."""
    c = 134217729.0 * a
    hi = c - (c - a)
    return hi, a - hi


def _two_prod(a, b):
    """Dekker's TwoProduct: a * b = p + e exactly (no fused multiply-add
    needed; exact while the product stays clear of under- and overflow)."""
    p = a * b
    ah, al = _split(a)
    bh, bl = _split(b)
    return p, al * bl - (((p - ah * bh) - al * bh) - ah * bl)


def _dot2(A, B):
    """Compensated dot products over the last axis, A (..., L) against the rows
    of B (K, L), as (s, c) with s + c the result (Ogita, Rump and Oishi, Dot2)."""
    shp = A.shape[:-1] + (B.shape[0],)
    s = np.zeros(shp)
    c = np.zeros(shp)
    for i in range(A.shape[-1]):
        q, eq = _two_prod(A[..., i:i + 1], B[:, i])
        s, es = _two_sum(s, q)
        c = c + (es + eq)
    return s, c


def _contract_last_exact(phi, n, xx):
    """_contract_last with compensated sums: term1 + term2 of
    dadi/Spectrum_mod.py:1527-1532 rounded once (up to eps^2-level error),
    instead of numpy's BLAS summation order."""
    db1, db2 = cached_dbeta(n, xx)
    s = (phi[..., 1:] - phi[..., :-1]) / (xx[1:] - xx[:-1])
    c1 = (phi[..., :-1] - s * xx[:-1]) / (n + 1)
    s1, k1 = _dot2(c1, db1)
    s2, k2 = _dot2(s, db2)
    fac = np.arange(1, n + 2) / ((n + 1) * (n + 2))
    q, eq = _two_prod(s2, fac)
    t, et = _two_sum(s1, q)
    return t + (et + ((k1 + eq) + k2 * fac))


def from_phi_exact(phi, ns, xx):
    """from_phi with every sampling sum rounded once (compensated dot
    products) rather than in numpy's BLAS order: the value the reference's
    formula defines for these inputs, against which any summation order is an
    approximation.  Its closeness to exact rational sums is checked in the
    tests."""
    phi = np.asarray(phi, dtype=np.float64)
    out = phi
    for k in range(phi.ndim - 1, -1, -1):
        if phi.ndim == 1:
            out = _contract_last_exact(out[None, :], ns[0], xx)[0]
            break
        out = _contract_last_exact(out, ns[k], xx)
        out = np.moveaxis(out, -1, 0)
    return out


def contract_fraction(phi_rows, n, xx, ks):
    """Exact rational value of the sampling contraction of each row of
    phi_rows (P, pts) at sample entries ks, with the reference's double
    inputs (s, c1, the betainc tables, fac) taken as exact; for checking the
    compensated sums on small cases."""
    from fractions import Fraction as F
    db1, db2 = cached_dbeta(n, xx)
    phi_rows = np.atleast_2d(phi_rows)
    s = (phi_rows[:, 1:] - phi_rows[:, :-1]) / (xx[1:] - xx[:-1])
    c1 = (phi_rows[:, :-1] - s * xx[:-1]) / (n + 1)
    fac = np.arange(1, n + 2) / ((n + 1) * (n + 2))
    out = np.empty((len(phi_rows), len(ks)))
    for p in range(len(phi_rows)):
        for j, k in enumerate(ks):
            t1 = sum(F(float(a)) * F(float(b)) for a, b in zip(c1[p], db1[k]))
            t2 = sum(F(float(a)) * F(float(b)) for a, b in zip(s[p], db2[k]))
            out[p, j] = float(t1 + t2 * F(float(fac[k])))
    return out


def from_phi(phi, ns, xx):
    """Sampled spectrum data (unmasked array) of a 1-5D phi on grid xx for every
    axis, following dadi/Spectrum_mod.py:1938-2062 (analytic samplers)."""
    phi = np.asarray(phi, dtype=np.float64)
    D = phi.ndim
    if D == 1:
        return _from_phi_1D(ns[0], xx, phi)
    out = phi
    # contract the last axis first, then move the new sample axis to the front
    for k in range(D - 1, -1, -1):
        out = _contract_last(out, ns[k], xx)
        out = np.moveaxis(out, -1, 0)
    # after D contractions the sample axes are in order (n0, n1, ..., n_{D-1})
    return out


def _binom_factor(n, ii, xx, het):
    """comb(n, ii) x^ii (1-x)^(n-ii) [* x (1-x)], dadi/Spectrum_mod.py:1272-1274."""
    f = comb(n, ii) * xx ** ii * (1 - xx) ** (n - ii)
    if het:
        f *= xx * (1 - xx)
    return f


def from_phi_direct(phi, ns, xxs, het_ascertained=None):
    """Sampled spectrum data of a 1-4D phi by the trapezoid rule against the
    binomial sampling factors, following the loops of the reference's direct
    samplers (dadi/Spectrum_mod.py:1263-1279, 1350-1380, 1550-1592,
    1594-1643); het_ascertained 'xx'/'yy'/'zz' weights that axis by x(1-x)."""
    phi = np.asarray(phi, dtype=np.float64)
    D = phi.ndim
    xxs = [np.asarray(x, dtype=np.float64) for x in xxs]
    het = [het_ascertained == name for name in ('xx', 'yy', 'zz', 'aa')[:D]]
    if D == 1:
        n, xx = ns[0], xxs[0]
        data = np.zeros(n + 1)
        for ii in range(n + 1):
            data[ii] = _trapezoid(_binom_factor(n, ii, xx, het[0]) * phi, xx)
        return data
    if D == 2:
        (nx, ny), (xx, yy) = ns, xxs
        data = np.zeros((nx + 1, ny + 1))
        fx = {ii: _binom_factor(nx, ii, xx, het[0]) for ii in range(nx + 1)}
        dx, dy = np.diff(xx), np.diff(yy)
        for jj in range(ny + 1):
            fy = _binom_factor(ny, jj, yy, het[1])
            over_y = _trapezoid(fy[nuax, :] * phi, dx=dy)
            for ii in range(nx + 1):
                data[ii, jj] = _trapezoid(fx[ii] * over_y, dx=dx)
        return data
    if D == 3:
        (nx, ny, nz), (xx, yy, zz) = ns, xxs
        data = np.zeros((nx + 1, ny + 1, nz + 1))
        dy, dz = np.diff(yy), np.diff(zz)
        half_dx = np.diff(xx) / 2.0
        fx = {ii: _binom_factor(nx, ii, xx, het[0]) for ii in range(nx + 1)}
        fy = {jj: _binom_factor(ny, jj, yy, het[1])[nuax, :] for jj in range(ny + 1)}
        for kk in range(nz + 1):
            fz = _binom_factor(nz, kk, zz, het[2])
            over_z = _trapezoid(fz[nuax, nuax, :] * phi, dx=dz)
            for jj in range(ny + 1):
                over_y = _trapezoid(fy[jj] * over_z, dx=dy)
                for ii in range(nx + 1):
                    integrand = fx[ii] * over_y
                    data[ii, jj, kk] = np.sum(half_dx * (integrand[1:] + integrand[:-1]))
        return data
    if D == 4:
        (nx, ny, nz, na), (xx, yy, zz, aa) = ns, xxs
        data = np.zeros((nx + 1, ny + 1, nz + 1, na + 1))
        dy, dz, da = np.diff(yy), np.diff(zz), np.diff(aa)
        half_dx = np.diff(xx) / 2.0
        fx = {ii: _binom_factor(nx, ii, xx, het[0]) for ii in range(nx + 1)}
        fy = {jj: _binom_factor(ny, jj, yy, het[1])[nuax, :] for jj in range(ny + 1)}
        fz = {kk: _binom_factor(nz, kk, zz, het[2])[nuax, nuax, :] for kk in range(nz + 1)}
        for ll in range(na + 1):
            fa = _binom_factor(na, ll, aa, het[3])
            over_a = _trapezoid(fa[nuax, nuax, nuax, :] * phi, dx=da)
            for kk in range(nz + 1):
                over_z = _trapezoid(fz[kk] * over_a, dx=dz)
                for jj in range(ny + 1):
                    over_y = _trapezoid(fy[jj] * over_z, dx=dy)
                    for ii in range(nx + 1):
                        integrand = fx[ii] * over_y
                        data[ii, jj, kk, ll] = np.sum(half_dx * (integrand[1:] + integrand[:-1]))
        return data
    raise ValueError("direct sampling is defined for 1-4 populations")


def marginalize(fs, over):
    """Spectrum.marginalize's data (dadi/Spectrum_mod.py): the entries summed
    over the listed axes."""
    return np.sum(np.asarray(fs), axis=tuple(over))


def corner_mask(shape):
    m = np.zeros(shape, dtype=bool)
    m[(0,) * len(shape)] = True
    m[tuple(s - 1 for s in shape)] = True
    return m


def linear_extrap(ys, xs):
    """dadi/Numerics.py:644-655."""
    y1, y2 = ys
    x1, x2 = xs
    return (x2 * y1 - x1 * y2) / (x2 - x1)


def quadratic_extrap(ys, xs):
    """dadi/Numerics.py:657-669."""
    y1, y2, y3 = ys
    x1, x2, x3 = xs
    return x2 * x3 / ((x1 - x2) * (x1 - x3)) * y1 + x1 * x3 / ((x2 - x1) * (x2 - x3)) * y2 \
        + x1 * x2 / ((x3 - x1) * (x3 - x2)) * y3


def extrapolate(results, xs, log=False, fail_mag=10):
    """Extrapolation step of make_extrap_func, dadi/Numerics.py:417-469.

    The reference applies it to Spectrum objects, i.e. numpy masked arrays,
    whose log10 masks non-positive arguments: an entry whose extrapolation is
    zero or of the opposite sign to the finest grid's value never counts as
    failed (it keeps the extrapolated value), while an infinite ratio does.
    The same rule is applied here on plain arrays."""
    res = [np.asarray(r, dtype=np.float64) for r in results]
    with np.errstate(all="ignore"):
        if log:
            res = [np.log(r) for r in res]
        if len(res) == 1:
            ex = res[0].copy()
        elif len(res) == 2:
            ex = linear_extrap(res, xs)
        elif len(res) == 3:
            ex = quadratic_extrap(res, xs)
        else:
            raise ValueError("1-3 grids supported")
        if log:
            ex = np.exp(ex)
        if len(res) > 1:
            best = res[int(np.argmin(xs))]
            if log:
                best = np.exp(best)
            ratio = ex / best
            failed = (ratio > 0) & (np.abs(np.log10(np.where(ratio > 0, ratio, 1.0))) > fail_mag)
            ex = np.where(failed, best, ex)
    return ex


def ll(model, data, mask=None):
    """Poisson log-likelihood over unmasked entries, dadi/Inference.py:466-509."""
    model = np.asarray(model, dtype=np.float64)
    data = np.asarray(data, dtype=np.float64)
    if mask is None:
        mask = corner_mask(model.shape)
    keep = ~mask
    with np.errstate(all="ignore"):
        terms = -model + data * np.log(model) - gammaln(data + 1.0)
    return float(np.sum(terms[keep]))


def ll_multinom(model, data, mask=None):
    """dadi/Inference.py:550-575 with optimal_sfs_scaling (668-685)."""
    model = np.asarray(model, dtype=np.float64)
    data = np.asarray(data, dtype=np.float64)
    if mask is None:
        mask = corner_mask(model.shape)
    keep = ~mask
    theta = data[keep].sum() / model[keep].sum()
    return ll(theta * model, data, mask)
