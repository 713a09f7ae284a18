"""Access to tests/golden/ref_integrations.npz (see tests/golden/make_golden.py)."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "golden", "ref_integrations.npz")

_data = None


def _load():
    global _data
    if _data is None:
        _data = np.load(PATH, allow_pickle=False)
    return _data


def meta():
    return json.loads(str(_load()["meta"]))


FN = {
    "growth": lambda a, b, T: (lambda t: a * np.exp(np.log(b / a) * t / T)),
    "linear": lambda a, b, T: (lambda t: a + (b - a) * t / T),
}


def make_fn(spec, T):
    if isinstance(spec, (int, float)):
        return spec
    name, a, b = spec
    return FN[name](a, b, T)


def case(i):
    d = _load()
    c = meta()[i]
    want = {"tv": d["case%d_tv" % i]}
    if ("case%d_const" % i) in d.files:
        want["const"] = d["case%d_const" % i]
    return c, d["case%d_phi0" % i], want


def grid_of(c):
    n = c["n"]
    if c["grid"] == "default":
        unif = np.linspace(-1, 1, n)
        g = 1. / (1. + np.exp(-8. * unif))
        return (g - g[0]) / (g[-1] - g[0])
    return np.linspace(0, 1, n)


def params(c):
    T = c["T"]
    nus = [make_fn(v, T) for v in c["nus"]]
    migs = [[make_fn(m, T) for m in ms] for ms in c["migs"]]
    gammas = [make_fn(v, T) for v in c["gammas"]]
    return nus, migs, gammas, c["hs"]


def run_oracle(O, c, phi0, xx, which):
    nus, migs, gammas, hs = params(c)
    args = (phi0, xx, c["T"], nus, migs, gammas, hs, c["theta0"], c["initial_t"], c["frozen"],
            c["nomut"], c["beta"])
    if which == "tv":
        return O.integrate_tv(*args)
    return O.integrate_const(*args)


_POP_KW = {
    2: ["m12", "m21"],
    3: ["m12", "m13", "m21", "m23", "m31", "m32"],
    4: ["m12", "m13", "m14", "m21", "m23", "m24", "m31", "m32", "m34", "m41", "m42", "m43"],
    5: ["m12", "m13", "m14", "m15", "m21", "m23", "m24", "m25", "m31", "m32", "m34", "m35",
        "m41", "m42", "m43", "m45", "m51", "m52", "m53", "m54"],
}


def api_kwargs(c, force_tv):
    """Keyword arguments of dadi's Integration.<n>_pops for this case.  With
    force_tv, numbers are wrapped as constant functions of time so the call
    takes the time-varying path, as dadi would."""
    D = c["D"]
    nus, migs, gammas, hs = params(c)
    wrap = (lambda v: v if callable(v) else (lambda t, v=v: v)) if force_tv else (lambda v: v)
    kw = {"T": c["T"], "theta0": wrap(c["theta0"]), "initial_t": c["initial_t"]}
    if D == 1:
        kw.update(nu=wrap(nus[0]), gamma=wrap(gammas[0]), h=wrap(hs[0]), beta=wrap(c["beta"]))
        return kw
    for k in range(D):
        kw["nu%d" % (k + 1)] = wrap(nus[k])
        kw["gamma%d" % (k + 1)] = wrap(gammas[k])
        kw["h%d" % (k + 1)] = wrap(hs[k])
        kw["frozen%d" % (k + 1)] = c["frozen"][k]
    names = _POP_KW[D]
    flat = [m for ms in migs for m in ms]
    # dadi rejects a function-valued migration rate next to a frozen
    # population (it cannot tell it is zero), so keep those as numbers
    wrap_m = (lambda v: v) if any(c["frozen"]) else wrap
    for name, m in zip(names, flat):
        kw[name] = wrap_m(m)
    if D == 2:
        kw["nomut1"], kw["nomut2"] = c["nomut"]
    return kw


def segments(c, xx, which):
    """Schedule of a golden case built by the package's host code."""
    from dadi_amd import schedule as S, _lib
    from dadi_amd.Misc import ensure_1arg_func as f
    nus, migs, gammas, hs = params(c)
    D = c["D"]
    fz, inj = S.masks(D, c["frozen"], c["nomut"])
    if which == "const":
        dt = min(S.compute_dt(nus[k], migs[k] if D > 1 else [0], gammas[k], hs[k])
                 for k in range(D))
        seg = S.const_segments(D, xx, c["initial_t"], c["T"], dt, nus, migs, gammas, hs,
                               c["theta0"], c["beta"])
        order = _lib.ORDER_PRE1 if D == 1 else _lib.ORDER_PRE
    else:
        seg = S.tv_segments(D, xx, c["initial_t"], c["T"], [f(v) for v in nus],
                            [[f(m) for m in ms] for ms in migs], [f(v) for v in gammas],
                            [f(v) for v in hs], f(c["theta0"]),
                            f(c["beta"]) if D == 1 else None)
        order = _lib.ORDER_TV
    return seg, order, fz, inj
