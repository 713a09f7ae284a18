"""
CPU checks that pin the oracle (oracle/dadi_oracle.{c,py}) before it is used
to judge the GPU path:
  * bit-for-bit against the reference's own C kernels, through the committed
    golden vectors and, when oracle/_ref is built, live on random inputs;
  * against mathematics the implementation cannot fake: dense linear solves,
    mass conservation of a sweep, closed-form sampling integrals;
  * against the reference's stored spectrum tests/test_2D_frozen1.fs.
"""
import json
import os

import numpy as np
import pytest

from tests import golden_cases as G

HERE = os.path.dirname(os.path.abspath(__file__))


def test_tridiag_matches_dense_solve(oracle):
    """Thomas solve against numpy's dense solve (pattern of
    tests/test_Tridiagonal.py:1-36 in the reference)."""
    rng = np.random.default_rng(5)
    for n in (3, 10, 100, 600):
        a = rng.uniform(-1, 0, n)
        c = rng.uniform(-1, 0, n)
        b = 2.5 + rng.uniform(0, 1, n)
        a[0] = 0
        c[-1] = 0
        r = rng.normal(size=n)
        M = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
        u = np.empty(n)
        gam = np.empty(n)
        P = oracle._ptr
        oracle.lib().orc_tridiag(P(a), P(b), P(c), P(r), P(u), n, P(gam))
        np.testing.assert_allclose(u, np.linalg.solve(M, r), rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("case", range(len(G.meta())))
def test_oracle_matches_reference_golden(oracle, case):
    """The generic C restatement reproduces the reference kernels' stored
    outputs bit for bit, for both coefficient orders."""
    c, phi0, want = G.case(case)
    xx = G.grid_of(c)
    got_tv = G.run_oracle(oracle, c, phi0, xx, "tv")
    assert np.array_equal(got_tv, want["tv"])
    if "const" in want:
        got_c = G.run_oracle(oracle, c, phi0, xx, "const")
        assert np.array_equal(got_c, want["const"])


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(HERE), "oracle", "_ref",
                                                    "libdadi_ref.so")),
                    reason="oracle/_ref not built (reference sources absent)")
def test_oracle_sweeps_match_reference_live(oracle):
    rng = np.random.default_rng(11)
    for D, n in [(1, 30), (2, 15), (3, 9), (4, 6), (5, 5)]:
        for grid in ("default", "lin"):
            xx = oracle.default_grid(n) if grid == "default" else np.linspace(0, 1, n)
            for axis in range(D):
                phi = rng.uniform(size=(n,) * D)
                nu = rng.uniform(0.1, 3)
                mig = rng.uniform(0, 3, D - 1)
                g = rng.uniform(-50, 50)
                h = rng.uniform(0, 1)
                beta = rng.uniform(0.3, 3) if D == 1 else 1
                dt = 10 ** rng.uniform(-5, -1)
                p1, p2 = phi.copy(), phi.copy()
                oracle.sweep(oracle.ORC_TV, p1, axis, xx, nu, mig, g, h, beta, dt)
                oracle.ref_sweep_tv(p2, axis, xx, nu, mig, g, h, beta, dt)
                assert np.array_equal(p1, p2)
                if D <= 3:
                    p1, p2 = phi.copy(), phi.copy()
                    order = oracle.ORC_PRE1 if D == 1 else oracle.ORC_PRE
                    oracle.sweep(order, p1, axis, xx, nu, mig, g, h, beta, dt)
                    arr = oracle.precalc_arrays(D, xx, axis, nu, mig, g, h, beta)
                    oracle.ref_sweep_pre(p2, axis, arr, dt)
                    assert np.array_equal(p1, p2)


def test_sweep_conserves_trapezoid_mass(oracle):
    """With no migration, selection or boundary flux the implicit step is
    conservative in the trapezoid measure dfactor is built for: the interior
    mass changes only through the boundary rows."""
    n = 50
    xx = oracle.default_grid(n)
    rng = np.random.default_rng(3)
    phi = np.zeros(n)
    phi[10:40] = rng.uniform(1, 2, 30)   # compactly supported away from the edges
    w = np.zeros(n)
    dx = np.diff(xx)
    w[:-1] += dx / 2
    w[1:] += dx / 2
    before = np.dot(w, phi)
    oracle.sweep(oracle.ORC_TV, phi, 0, xx, 1.0, [], 0.0, 0.5, 1, 1e-4)
    after = np.dot(w, phi)
    # only the tiny amount that reaches the absorbing-ish boundaries is lost
    assert abs(after - before) / before < 1e-6


def test_sampler_exact_for_linear_phi(oracle):
    """For phi = a + b x the analytic sampler is exact:
    E[k] = C(n,k) int x^k (1-x)^(n-k) (a + b x) dx = a/(n+1) + b (k+1)/((n+1)(n+2))."""
    xx = oracle.default_grid(37)
    a, b = 0.7, -0.3
    for n in (5, 20, 40):
        fs = oracle.from_phi(a + b * xx, [n], xx)
        k = np.arange(n + 1)
        want = a / (n + 1) + b * (k + 1) / ((n + 1) * (n + 2))
        np.testing.assert_allclose(fs, want, rtol=1e-12, atol=1e-15)
    # 2D: separable product phi = (a + b x)(c + d y)
    c_, d_ = 1.1, 0.4
    phi2 = (a + b * xx)[:, None] * (c_ + d_ * xx)[None, :]
    fs2 = oracle.from_phi(phi2, [6, 9], xx)
    e1 = a / 7 + b * (np.arange(7) + 1) / (7 * 8)
    e2 = c_ / 10 + d_ * (np.arange(10) + 1) / (10 * 11)
    np.testing.assert_allclose(fs2, np.outer(e1, e2), rtol=1e-12, atol=1e-15)


def test_snm_spectrum_is_theta_over_i(oracle):
    """Standard neutral model: the extrapolated spectrum is 1/i
    (reference tests/test_Results.py:5-12, atol 1e-3 there)."""
    n = 20
    res = []
    xs = []
    for pts in (100, 120, 140):
        xx = oracle.default_grid(pts)
        res.append(oracle.from_phi(oracle.phi_1D(xx), [n], xx))
        xs.append(xx[1])
    ex = oracle.extrapolate(res, xs)
    i = np.arange(1, n)
    np.testing.assert_allclose(ex[1:-1], 1.0 / i, atol=1e-3)


def _read_fs(path):
    lines = [l for l in open(path).read().split("\n") if l and not l.startswith("#")]
    shape = tuple(int(v) for v in lines[0].split()[:-1])
    data = np.array([float(v) for v in lines[1].split()]).reshape(shape)
    mask = np.array([int(v) for v in lines[2].split()]).reshape(shape).astype(bool)
    return data, mask


def test_frozen1_reference_golden(oracle):
    """tests/test_2D_frozen1.fs of the reference (tests/test_Freezing.py:142-153):
    2D frozen population 1, pts 30, n (10,10).  The reference checks rtol 1e-2;
    the oracle reproduces the stored digits to ~1e-7."""
    data, mask = _read_fs(os.path.join(HERE, "golden", "test_2D_frozen1.fs"))
    xx = oracle.default_grid(30)
    phi = oracle.phi_1D_to_2D(xx, oracle.phi_1D(xx))
    phi = oracle.integrate_const(phi, xx, 0.1, [1, 0.3], [[0], [0]], [0, 0.5], [0.5, 0.5],
                                 frozen=[True, False])
    fs = oracle.from_phi(phi, [10, 10], xx)
    rel = np.abs(fs - data)[~mask] / np.abs(data[~mask])
    assert rel.max() < 1e-6


def test_direct_sampler_matches_beta_integrals(oracle):
    """The oracle's trapezoid samplers (dadi/Spectrum_mod.py:1263-1279,
    1350-1380, 1550-1592) integrate comb(n,k) x^k (1-x)^(n-k) phi(x): for
    polynomial phi on a fine grid they reach the closed-form Beta integrals,
    with and without the x(1-x) weight of heterozygote ascertainment."""
    from scipy.special import beta as B, comb
    xx = np.linspace(0, 1, 4001)
    n = 9

    def exact(a, het):
        k = np.arange(n + 1)
        e = het * 1
        return comb(n, k) * B(k + a + 1 + e, n - k + 1 + e)

    for a in (0, 1, 3):
        for het in (False, True):
            got = oracle.from_phi_direct(xx ** a, [n], [xx], 'xx' if het else None)
            assert np.allclose(got, exact(a, het), rtol=1e-5, atol=1e-12)
    # 2D and 3D separable densities are products of 1D integrals; the
    # ascertainment weight applies to the named axis only
    m = 6
    x2 = np.linspace(0, 1, 1201)
    phi2 = (x2 ** 2)[:, None] * (1 - x2)[None, :]
    got = oracle.from_phi_direct(phi2, [m, m], [x2, x2], 'yy')
    k = np.arange(m + 1)
    ex = comb(m, k) * B(k + 3, m - k + 1)
    ey = comb(m, k) * B(k + 2, m - k + 3)   # (1-y) * y(1-y)
    assert np.allclose(got, np.outer(ex, ey), rtol=1e-5, atol=1e-12)
    x3 = np.linspace(0, 1, 61)   # coarse: the trapezoid error is ~1e-3 here
    phi3 = ((x3 ** 2)[:, None] * (1 - x3)[None, :])[:, :, None] * np.ones(61)[None, None, :]
    got3 = oracle.from_phi_direct(phi3, [m, m, 4], [x3, x3, x3])
    ez = comb(4, np.arange(5)) * B(np.arange(5) + 1, 4 - np.arange(5) + 1)
    ey0 = comb(m, k) * B(k + 1, m - k + 2)
    assert np.allclose(got3, ex[:, None, None] * ey0[None, :, None] * ez[None, None, :],
                       rtol=5e-3, atol=1e-6)
    # a planted mistake (dropping the binomial coefficient) does not pass
    bad = got / comb(m, k)[None, :]
    assert not np.allclose(bad, np.outer(ex, ey), rtol=1e-3)


@pytest.mark.parametrize("nu,gamma,h,beta", [(1.0, 0.0, 0.5, 1.0), (0.6, -3.0, 0.2, 2.5),
                                             (2.0, 4.0, 0.8, 0.4)])
def test_x_chromosome_step_equals_reparametrised_autosome(oracle, nu, gamma, h, beta):
    """One X-chromosome step (dadi/Integration.py:1244-1352) is the autosomal
    1D step (dadi/Integration.py:986-1035) with the breeding-ratio factor
    (b+1)^2/(4b) = (2 beta+4)(beta+1)/(9 beta), gamma' = 4 gamma/3 and
    h' = (0.5+h)/2: V and M agree as functions of x and both boundary terms
    use 0.5/nu.  So the oracle's X restatement must agree with the pinned
    autosomal sweep up to rounding; without injection (theta0 = 0) to isolate
    the operator."""
    xx = oracle.default_grid(60)
    rng = np.random.default_rng(7)
    phi = rng.uniform(0.5, 2.0, 60)
    dt = 1e-3
    x_step = oracle.one_pop_X(phi, xx, dt, nu, gamma, h, beta, 1.0, theta0=0.0, dt=dt)
    K = (2 * beta + 4) * (beta + 1) / (9 * beta)      # >= 1.29 for every beta > 0
    b_a = ((4 * K - 2) + np.sqrt((4 * K - 2) ** 2 - 4)) / 2
    assert np.isclose((b_a + 1) ** 2 / (4 * b_a), K, rtol=1e-14)
    auto = phi.copy()
    oracle.sweep(oracle.ORC_PRE1, auto, 0, xx, nu, [], 4 * gamma / 3, (0.5 + h) / 2, b_a, dt)
    assert np.allclose(x_step, auto, rtol=1e-11, atol=0)
    # a planted mistake in the selection term is caught
    wrong = phi.copy()
    oracle.sweep(oracle.ORC_PRE1, wrong, 0, xx, nu, [], 2 * gamma, (0.5 + h) / 2, b_a, dt)
    assert gamma == 0 or not np.allclose(x_step, wrong, rtol=1e-8, atol=0)


def test_x_chromosome_injection_factor(oracle):
    """2/(1+2 beta) (1/(alpha+1) + beta) is 1 at beta = alpha = 1 (the X
    injection then equals the autosomal one) and 4/5 at beta = 2, alpha = 1 ...
    checked through the package's schedule against the oracle loop."""
    from dadi_amd import schedule as S
    xx = oracle.default_grid(30)
    for beta, alpha, want in ((1.0, 1.0, 1.0), (2.0, 1.0, 2 / 5 * 2.5), (0.5, 3.0, 1.0 * 0.75)):
        auto = oracle.inject_value(1, 1e-3, xx, 1.0)
        assert np.isclose(S.inject_value_X(1e-3, xx, 1.0, beta, alpha), auto * want,
                          rtol=1e-14)


def test_delj_numpy_path_matches_reference_c(oracle):
    """With use_delj_trick the oracle's numpy coefficient arrays
    (dadi/Integration.py:960-984, 997-1095) and the reference C sweeps
    (dadi/integration_shared.c:52-70) agree to rounding (they differ only in
    exp and in how b and r are associated); and the trick matters: results
    move away from delta = 1/2 where M dx / V is large."""
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built")
    xx = oracle.default_grid(40)
    phi0 = oracle.phi_1D_to_2D(xx, oracle.phi_1D(xx))
    args = (phi0, xx, 0.05, [0.3, 2.0], [[4.0], [1.5]], [-8.0, 5.0], [0.3, 0.7])
    pre = oracle.integrate_const(*args, backend="ref", delj=True)
    tv = oracle.integrate_tv(*args, backend="ref", delj=True)
    plain = oracle.integrate_const(*args, backend="ref")
    assert np.allclose(pre, tv, rtol=1e-9, atol=1e-12 * np.abs(pre).max())
    assert not np.allclose(pre, plain, rtol=1e-6)


@pytest.mark.parametrize("gamma,h,beta", [(2.0, 0.2, 1.0), (-3.0, 0.2, 1.0), (-3.0, 0.8, 1.0),
                                          (4.0, 0.7, 1.0), (1.5, 0.3, 2.0), (0.0, 0.3, 1.0)])
def test_phi_1D_dominance_branch_matches_closed_form(gamma, h, beta):
    """phi_1D for h != 0.5 (dadi/PhiManip.py:45-108: numerical quadrature of
    exp(-Q), Q(x) = 4 g h x + 2 g (1-2h) x^2, g = gamma 4 beta/(beta+1)^2)
    against the closed form of the same integrals through erf / erfi,
    evaluated with mpmath at 40 digits:
    phi(x) = e^Q(x) int_x^1 e^-Q / int_0^1 e^-Q / (x(1-x)) * nu theta0 4 beta/(beta+1)^2,
    with the end values phi[0] = phi[1] and phi[-1] = 1/int_0^1 e^-Q.  For
    gamma = 0 this is the neutral 1/x."""
    import mpmath
    from dadi_amd import Numerics, PhiManip
    mpmath.mp.dps = 40
    xx = Numerics.default_grid(31)
    nu, theta0 = 1.7, 0.9
    got = PhiManip.phi_1D(xx, nu=nu, theta0=theta0, gamma=gamma, h=h, beta=beta)
    g = mpmath.mpf(gamma) * 4 * beta / (beta + 1) ** 2
    a, b = 4 * g * h, 2 * g * (1 - 2 * mpmath.mpf(h))

    def prim(v):   # an antiderivative of exp(-a v - b v^2)
        v = mpmath.mpf(v)
        if b > 0:
            c = a / (2 * b)
            return mpmath.exp(c * c * b) * mpmath.sqrt(mpmath.pi / b) / 2 * mpmath.erf(
                mpmath.sqrt(b) * (v + c))
        if b < 0:
            c = a / (2 * b)
            return mpmath.exp(c * c * b) * mpmath.sqrt(mpmath.pi / -b) / 2 * mpmath.erfi(
                mpmath.sqrt(-b) * (v + c))
        return v if a == 0 else -mpmath.exp(-a * v) / a

    scale = nu * theta0 * 4 * beta / (beta + 1) ** 2
    int0 = prim(1) - prim(0)
    want = np.empty(len(xx))
    for i, x in enumerate(xx):
        if 0 < i < len(xx) - 1:
            q = mpmath.mpf(float(x))
            Q = a * q + b * q * q
            want[i] = float(mpmath.exp(Q) * (prim(1) - prim(q)) / int0 / (q * (1 - q)) * scale)
    want[0] = want[1]
    want[-1] = float(1 / int0 * scale)
    np.testing.assert_allclose(got, want, rtol=1e-9)
    # the check separates a plausible mistake: the dominance orientation (h <-> 1-h)
    if gamma != 0:
        other = PhiManip.phi_1D(xx, nu=nu, theta0=theta0, gamma=gamma, h=1 - h, beta=beta)
        assert np.max(np.abs(other / want - 1)) > 1e-3
