"""
Writes tests/golden/split_mig_pts80_90_100_ll.npz: the likelihood fixture of
the 2-population split_mig configuration at pts (80, 90, 100), sample sizes
(20, 20).

Parameters are seeded draws of that configuration's distribution
(numpy.random.default_rng(2024): nu1, nu2 ~ logU[0.1, 10], T ~ U[0.01, 2],
m ~ U[0, 5]), keeping the first four with T <= 0.3 so that the CPU oracle
finishes in seconds.  For each, the expected spectrum is the oracle's
(dadi's split_mig, dadi/Demographics2D.py:133-167: phi_1D -> phi_1D_to_2D ->
two_pops with m12 = m21 = m, sampled, extrapolated over the three grids on
log values), and the data are a Poisson sample of 1000 x that spectrum
(numpy.random.default_rng(77)), stored as integers.

Run from the repository root:  python tests/golden/make_ll_fixture.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import dadi_oracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "split_mig_pts80_90_100_ll.npz")
PTS = (80, 90, 100)
NS = (20, 20)


def draws(seed=2024, keep=4, tmax=0.3):
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < keep:
        nu1, nu2 = np.exp(rng.uniform(np.log(0.1), np.log(10), 2))
        T = rng.uniform(0.01, 2)
        m = rng.uniform(0, 5)
        if T <= tmax:
            out.append((nu1, nu2, T, m))
    return np.array(out)


def oracle_split_mig(params):
    nu1, nu2, T, m = params
    res, xs = [], []
    for pts in PTS:
        xx = O.default_grid(pts)
        phi = O.phi_1D_to_2D(xx, O.phi_1D(xx))
        phi = O.integrate_const(phi, xx, T, [nu1, nu2], [[m], [m]], [0, 0], [0.5, 0.5])
        res.append(O.from_phi(phi, NS, xx))
        xs.append(xx[1])
    return O.extrapolate(res, xs, log=True)


def main():
    O.lib()
    params = draws()
    rng = np.random.default_rng(77)
    data = []
    for p in params:
        model = oracle_split_mig(p)
        data.append(rng.poisson(1000 * np.where(np.isfinite(model), model, 0)))
    np.savez_compressed(OUT, params=params, data=np.array(data, dtype=np.int64),
                        pts=np.array(PTS), ns=np.array(NS))
    print("wrote", OUT, params.tolist())


if __name__ == "__main__":
    main()
