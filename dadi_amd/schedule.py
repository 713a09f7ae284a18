"""
Host-side time schedule of an integration.

dadi's time step never depends on phi (dadi/Integration.py:147-176 uses only
the demographic parameters), so the whole sequence of (dt, parameters,
injected mass) can be computed on the host before anything is launched and
handed to the device as a short list of segments (``dh_segment``: a run of
identical steps).  The arithmetic below is the reference's, operation for
operation, so the step sequence and every injected value are bit-identical to
what dadi's Python loops produce.
"""
import numpy as np

from ._lib import SEG_DTYPE


def compute_dt(nu, ms, gamma, h, timescale_factor=1e-3, use_old_timestep=False,
               old_timescale_factor=0.1, dx=None):
    """Time step from the largest V or M on the domain, dadi/Integration.py:147-176."""
    if use_old_timestep:
        return old_timescale_factor * dx[0]
    maxVM = max(0.25 / nu, sum(ms),
                abs(gamma) * 2 * max(np.abs(h + (1 - 2 * h) * 0.5) * 0.5 * (1 - 0.5),
                                     np.abs(h + (1 - 2 * h) * 0.25) * 0.25 * (1 - 0.25)))
    if maxVM > 0:
        dt = timescale_factor / maxVM
    else:
        dt = np.inf
    if dt == 0:
        raise ValueError('Timestep is zero. Values passed in are nu=%f, ms=%s,'
                         'gamma=%f, h=%f.' % (nu, str(ms), gamma, h))
    return dt


def inject_value(D, dt, xx, theta0):
    """Mass added at index 1 of one axis per step, dadi/Integration.py:70-145:
    dt/x1 * theta0/2 * 2**D / ((x2 - x0) * x1 * ... * x1), left to right.
    All axes share the grid, so every population gets the same value."""
    den = (xx[2] - xx[0])
    for _ in range(D - 1):
        den = den * xx[1]
    return dt / xx[1] * theta0 / 2 * (2 ** D) / den


def const_steps(initial_t, T, dt):
    """Step sizes of ``while t < T: this_dt = min(dt, T - t); t += this_dt``
    (dadi/Integration.py:1027-1034, 1108-1116, 1232-1242) as
    (number of full steps, [tail step sizes]).  Full steps are accumulated with
    a sequential cumulative sum, which performs the same additions in the same
    order as the Python loop."""
    t0 = float(initial_t)
    T = float(T)
    if not (t0 < T):
        return 0, []
    nfull = 0
    t = t0
    if np.isfinite(dt) and dt > 0:
        chunk = int(min(max((T - t0) / dt + 4, 16), 1 << 22))
        while True:
            arr = np.empty(chunk + 1)
            arr[0] = t
            arr[1:] = dt
            ts = np.cumsum(arr)
            ok = (ts < T) & (dt <= T - ts)
            # ok is a prefix of True values: full steps continue while it holds
            bad = np.flatnonzero(~ok)
            if bad.size:
                k = int(bad[0])
                nfull += k
                t = float(ts[k])
                break
            nfull += chunk
            t = float(ts[chunk])
    tails = []
    while t < T:
        this_dt = min(dt, T - t)
        tails.append(this_dt)
        t += this_dt
    return nfull, tails


def _blank(count):
    seg = np.zeros(count, dtype=SEG_DTYPE)
    seg["beta_p"] = 1.0
    seg["beta_r"] = 1.0
    seg["beta_q"] = 1.0
    seg["nu"] = 1.0
    return seg


def _fill_params(seg, D, nus, migs, gammas, hs, beta):
    for k in range(D):
        seg["nu"][:, k] = nus[k]
        seg["gamma"][:, k] = gammas[k]
        seg["h"][:, k] = hs[k]
        for j, m in enumerate(migs[k]):
            seg["mig"][:, k, j] = m
    if D == 1:
        seg["beta_p"] = (beta + 1.) ** 2
        seg["beta_q"] = 4 * beta


def const_segments(D, xx, initial_t, T, dt, nus, migs, gammas, hs, theta0, beta=1):
    """Segments of a constant-parameter integration: one run of full steps plus
    the (usually one) tail step."""
    nfull, tails = const_steps(initial_t, T, dt)
    count = (1 if nfull else 0) + len(tails)
    seg = _blank(count)
    _fill_params(seg, D, nus, migs, gammas, hs, beta)
    i = 0
    if nfull:
        seg["nsteps"][0] = nfull
        seg["dt"][0] = dt
        seg["inject"][0] = inject_value(D, dt, xx, theta0)
        i = 1
    for td in tails:
        seg["nsteps"][i] = 1
        seg["dt"][i] = td
        seg["inject"][i] = inject_value(D, td, xx, theta0)
        i += 1
    return seg


def inject_value_X(dt, xx, theta0, beta, alpha):
    """Mass added at index 1 per step of an X-chromosome integration,
    dadi/Integration.py:1248-1254:
    dt/x1 * theta0/2 * factor * 2/(x2 - x0), factor = 2/(1+2 beta) (1/(alpha+1) + beta),
    left to right."""
    factor = 2. / (1. + 2. * beta) * (1. / (alpha + 1.) + beta)
    return dt / xx[1] * theta0 / 2. * factor * 2. / (xx[2] - xx[0])


def xchrom_segments(xx, initial_t, T, dt, nu, gamma, h, beta, alpha, theta0):
    """Segments of a constant-parameter X-chromosome integration
    (dadi/Integration.py:1301-1352): V = 1/nu x(1-x) (2 beta+4)(beta+1)/(9 beta)
    and the X selection term (the kernel's chrom = 1), the X injection."""
    nfull, tails = const_steps(initial_t, T, dt)
    count = (1 if nfull else 0) + len(tails)
    seg = _blank(count)
    _fill_params(seg, 1, [nu], [[]], [gamma], [h], 1)
    seg["beta_p"] = 2 * beta + 4.
    seg["beta_r"] = beta + 1.
    seg["beta_q"] = 9. * beta
    seg["chrom"] = 1
    i = 0
    if nfull:
        seg["nsteps"][0] = nfull
        seg["dt"][0] = dt
        seg["inject"][0] = inject_value_X(dt, xx, theta0, beta, alpha)
        i = 1
    for td in tails:
        seg["nsteps"][i] = 1
        seg["dt"][i] = td
        seg["inject"][i] = inject_value_X(td, xx, theta0, beta, alpha)
        i += 1
    return seg


def total_steps(seg):
    return int(np.sum(seg["nsteps"]))


def compress(seg):
    """Merge consecutive segments whose step records are identical."""
    if len(seg) <= 1:
        return seg
    body = seg.copy()
    body["nsteps"] = 0
    raw = body.view(np.uint8).reshape(len(seg), -1)
    same = np.all(raw[1:] == raw[:-1], axis=1)
    starts = np.concatenate([[0], np.flatnonzero(~same) + 1])
    out = seg[starts].copy()
    counts = np.add.reduceat(seg["nsteps"], starts)
    out["nsteps"] = counts
    return out


def tv_segments(D, xx, initial_t, T, nu_f, mig_f, gamma_f, h_f, theta0_f, beta_f,
                timescale_factor=1e-3, use_old_timestep=False, old_timescale_factor=0.1,
                history=None):
    """Segments of a time-varying integration, following the loops of
    dadi/Integration.py:219-258 (1D), 347-386 (2D), 504-553 (3D), 671-726 (4D),
    874-936 (5D): dt from the parameters at the current time, the step's
    parameters and theta0 at next_t, the same < 0 and == 0 checks.

    nu_f[k], gamma_f[k], h_f[k]: callables per population; mig_f[k]: list of
    callables for the rates into population k (other populations in
    increasing order).  ``history`` (a list) receives the per-step demography
    records dadi keeps for its Demes cache."""
    dx = np.diff(xx)
    t = initial_t
    nu = [f(t) for f in nu_f]
    ga = [f(t) for f in gamma_f]
    hh = [f(t) for f in h_f]
    mm = [[f(t) for f in fs] for fs in mig_f]
    if history is not None:
        history.append([0, list(nu), [m for ms in mm for m in ms]])
    recs = []
    while t < T:
        dt = min(compute_dt(nu[k], mm[k] if D > 1 else [0], ga[k], hh[k], timescale_factor,
                            use_old_timestep, old_timescale_factor, dx) for k in range(D))
        this_dt = min(dt, T - t)
        next_t = t + this_dt
        nu = [f(next_t) for f in nu_f]
        ga = [f(next_t) for f in gamma_f]
        hh = [f(next_t) for f in h_f]
        mm = [[f(next_t) for f in fs] for fs in mig_f]
        th = theta0_f(next_t)
        bt = beta_f(next_t) if beta_f is not None else 1
        if history is not None:
            history.append([next_t, list(nu), [m for ms in mm for m in ms]])
        if np.any(np.less([T] + nu + [m for ms in mm for m in ms] + [th], 0)):
            raise ValueError('A time, population size, migration rate, or '
                             'theta0 is < 0. Has the model been mis-specified?')
        if np.any(np.equal(nu, 0)):
            raise ValueError('A population size is 0. Has the model been '
                             'mis-specified?')
        recs.append((this_dt, inject_value(D, this_dt, xx, th), list(nu), [list(m) for m in mm],
                     list(ga), list(hh), bt))
        t = next_t
    seg = _blank(len(recs))
    for i, (d, inj, nu_, mm_, ga_, hh_, bt) in enumerate(recs):
        seg["nsteps"][i] = 1
        seg["dt"][i] = d
        seg["inject"][i] = inj
        for k in range(D):
            seg["nu"][i, k] = nu_[k]
            seg["gamma"][i, k] = ga_[k]
            seg["h"][i, k] = hh_[k]
            for j, m in enumerate(mm_[k]):
                seg["mig"][i, k, j] = m
        if D == 1:
            seg["beta_p"][i] = (bt + 1.) ** 2
            seg["beta_q"][i] = 4 * bt
    return compress(seg)


def masks(D, frozen, nomut):
    fz = 0
    inj = 0
    for k in range(D):
        if frozen[k]:
            fz |= 1 << k
        elif not nomut[k]:
            inj |= 1 << k
    return fz, inj
