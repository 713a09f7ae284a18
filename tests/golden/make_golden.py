"""
Writes tests/golden/ref_integrations.npz: inputs and outputs of short
integrations computed by the reference's own C kernels
(oracle/_ref/libdadi_ref.so, compiled from dadi/tridiag.c,
dadi/integration_shared.c and dadi/integration{1..5}D.c), driven by the
oracle's restatement of dadi's Python time loops.

Each case is integrated twice through the reference C:
  * "const": numbers-only parameters, the constant-parameter path
    (numpy coefficient arrays + implicit_precalc_*; 1-3 populations),
  * "tv":    the same parameters through the time-varying path
    (implicit_<D>D<axis>, every population count), plus cases with
    parameters that are functions of time.

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import dadi_oracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_integrations.npz")

# time-dependent parameter shapes used by the "tv_fn" cases (encoded by name so
# the GPU tests can rebuild the same callables)
FN = {
    "growth": lambda a, b, T: (lambda t: a * np.exp(np.log(b / a) * t / T)),
    "linear": lambda a, b, T: (lambda t: a + (b - a) * t / T),
}


def make_fn(spec, T):
    if isinstance(spec, (int, float)):
        return spec
    name, a, b = spec
    return FN[name](a, b, T)


def cases():
    rng = np.random.default_rng(20240611)
    sizes = {1: 40, 2: 18, 3: 11, 4: 7, 5: 5}
    out = []
    for D in (1, 2, 3, 4, 5):
        for grid in ("default", "linspace"):
            for variant in range(3):
                n = sizes[D] + variant
                nus = list(rng.uniform(0.2, 3.0, D))
                migs = [list(rng.uniform(0, 3, D - 1)) for _ in range(D)]
                gammas = list(rng.uniform(-8, 8, D))
                hs = list(rng.uniform(0.05, 0.95, D))
                frozen = [False] * D
                nomut = [False] * D
                if variant == 2 and D > 1:
                    k = int(rng.integers(0, D))
                    frozen[k] = True
                    migs = [[0.0] * (D - 1) for _ in range(D)]
                if variant == 1 and D == 2:
                    nomut[1] = True
                theta0 = float(rng.uniform(0.5, 3))
                initial_t = float(rng.uniform(0, 0.05))
                T = initial_t + float(rng.uniform(0.002, 0.02))
                beta = float(rng.uniform(0.5, 2.0)) if D == 1 else 1
                out.append(dict(D=D, n=n, grid=grid, nus=nus, migs=migs, gammas=gammas, hs=hs,
                                frozen=frozen, nomut=nomut, theta0=theta0, initial_t=initial_t,
                                T=T, beta=beta, seed=int(rng.integers(1 << 30))))
    # time-varying parameter cases
    for D in (1, 2, 3):
        n = sizes[D] + 1
        T = 0.015
        nus = [("growth", 0.5, 3.0)] + [1.3] * (D - 1)
        migs = [[("linear", 0.2, 2.0)] * (D - 1)] + [[0.7] * (D - 1) for _ in range(D - 1)]
        out.append(dict(D=D, n=n, grid="default", nus=nus, migs=migs,
                        gammas=[("linear", -2.0, 3.0)] + [0.5] * (D - 1), hs=[0.3] * D,
                        frozen=[False] * D, nomut=[False] * D, theta0=1.7, initial_t=0.0, T=T,
                        beta=1, seed=int(rng.integers(1 << 30)), fn=True))
    return out


def grid_of(c):
    return O.default_grid(c["n"]) if c["grid"] == "default" else np.linspace(0, 1, c["n"])


def phi0_of(c):
    rng = np.random.default_rng(c["seed"])
    return rng.uniform(0.1, 2.0, size=(c["n"],) * c["D"])


def main():
    store = {}
    meta = []
    for i, c in enumerate(cases()):
        xx = grid_of(c)
        phi0 = phi0_of(c)
        T = c["T"]
        if c.get("fn"):
            nus = [make_fn(v, T) for v in c["nus"]]
            migs = [[make_fn(m, T) for m in ms] for ms in c["migs"]]
            gammas = [make_fn(v, T) for v in c["gammas"]]
            tv = O.integrate_tv(phi0, xx, T, nus, migs, gammas, c["hs"], c["theta0"],
                                c["initial_t"], c["frozen"], c["nomut"], c["beta"], backend="ref")
            store["case%d_tv" % i] = tv
        else:
            args = (phi0, xx, T, c["nus"], c["migs"], c["gammas"], c["hs"], c["theta0"],
                    c["initial_t"], c["frozen"], c["nomut"], c["beta"])
            store["case%d_tv" % i] = O.integrate_tv(*args, backend="ref")
            if c["D"] <= 3:
                store["case%d_const" % i] = O.integrate_const(*args, backend="ref")
        store["case%d_phi0" % i] = phi0
        meta.append(c)
    store["meta"] = np.array(json.dumps(meta))
    np.savez_compressed(OUT, **store)
    print("wrote", OUT, "cases:", len(meta))


if __name__ == "__main__":
    main()
