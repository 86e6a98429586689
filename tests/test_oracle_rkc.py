"""Pins of the oracle's RKC core (solve_rkc.cu restated in oracle/or_rkc.c)
against mathematics that does not come from the reference text:

* or_det_sum: exact on integers, within the pairwise bound of math.fsum;
* step (solve_rkc.cu:815-906): one step on a linear two-compartment cable
  equals the published RKC stability polynomial R_m(z) = a_m + b_m T_m(w0 +
  w1 z) (Sommeijer, Shampine & Verwer, J. Comput. Appl. Math. 88 (1998)
  315-326), evaluated here with numpy's Chebyshev class; the stage
  recurrence integrates a constant right-hand side exactly (c_m = 1) and
  conserves the capacitance-weighted mean of a closed cable;
* rkcrho (:562-754): 1.2 x the dominant eigenvalue magnitude of C^-1 L from
  a dense eigensolve;
* rkcint (:794-813): the Hermite cubic reproduces any cubic from its end
  values and slopes;
* rkclow (:908-1096) with the driver: the global error on a problem with a
  known solution stays within a small multiple of the tolerance, and the
  step count grows like tol^(-1/3) (local error O(h^3) control); the
  spectral radius is re-estimated only after a rejected step (nstsig is
  never advanced in the reference).
"""
import math

import numpy as np
import pytest
from numpy.polynomial import chebyshev as cheb

import orc


def rkc_stability(m, z):
    w0 = 1.0 + 2.0 / (13.0 * m * m)
    Tm = cheb.Chebyshev.basis(m)
    d1, d2 = Tm.deriv(1), Tm.deriv(2)
    w1 = d1(w0) / d2(w0)
    bm = d2(w0) / d1(w0) ** 2
    am = 1.0 - bm * Tm(w0)
    return am + bm * Tm(w0 + w1 * z)


def test_det_sum_exact_on_integers():
    rng = np.random.default_rng(1)
    for n in (1, 7, 255, 256, 257, 1000, 65536 + 3):
        x = rng.integers(-2**20, 2**20, size=n).astype(np.float64)
        assert orc.det_sum(x) == float(np.sum(x.astype(np.int64)))


def test_det_sum_within_pairwise_bound():
    rng = np.random.default_rng(2)
    for n in (10, 4097, 100000):
        x = rng.standard_normal(n) * 10.0 ** rng.uniform(-3, 3, n)
        exact = math.fsum(x)
        bound = math.ceil(math.log2(n) + 9) * np.finfo(float).eps * np.sum(np.abs(x))
        assert abs(orc.det_sum(x) - exact) <= bound


def two_comp(g=1.0, c=(1.0, 1.0), y=(1.0, -1.0), b=(0.0, 0.0), tol=1e-6):
    L = np.array([[0.0, g], [g, 0.0]])
    return orc.Pop.generic(L, np.array(c), np.array(b), np.array(y), tol, tol)


@pytest.mark.parametrize("m", [2, 3, 5, 9, 17, 40])
def test_step_is_the_rkc_stability_polynomial(m):
    g, c = 1.0, (1.0, 1.0)
    lam = -g * (1.0 / c[0] + 1.0 / c[1])
    # real stability boundary: the argument of T_m reaches -1 at z = -(1 + w0) / w1
    w0 = 1.0 + 2.0 / (13.0 * m * m)
    Tm = cheb.Chebyshev.basis(m)
    beta = (1.0 + w0) * Tm.deriv(2)(w0) / Tm.deriv(1)(w0)
    for z in (-0.01, -0.5, -0.3 * beta, -0.9 * beta):
        p = two_comp(g, c, y=(0.7, -1.3))
        yn = p.get(orc.YN)
        p.rkc_step(z / lam, m)
        y = p.get(orc.Y)
        d0, d1 = yn[0] - yn[1], y[0] - y[1]
        assert d1 == pytest.approx(rkc_stability(m, z) * d0, rel=1e-9, abs=1e-11 * abs(d0))
        # the capacitance-weighted mean of a closed cable does not move
        assert c[0] * y[0] + c[1] * y[1] == pytest.approx(c[0] * yn[0] + c[1] * yn[1], abs=1e-13)
        assert abs(rkc_stability(m, z)) <= 1.0 + 1e-12   # inside the stability region


@pytest.mark.parametrize("m", [2, 4, 11, 30])
def test_step_integrates_constant_rhs_exactly(m):
    # one compartment, no neighbours: y' = -b / c, so Y_m = y_n + h F exactly
    p = orc.Pop.generic(np.zeros((1, 1)), np.array([2.0]), np.array([3.0]), np.array([-65.0]))
    h = 0.037
    p.rkc_step(h, m)
    assert p.get(orc.Y)[0] == pytest.approx(-65.0 + h * (-3.0 / 2.0), rel=0, abs=1e-12)


def test_rkcrho_matches_dense_eigensolve():
    g = np.array([1.0, 1.0, 0.5])
    L = np.zeros((4, 4))
    for i, gi in enumerate(g):
        L[i, i + 1] = L[i + 1, i] = gi
    cm = np.array([1.0, 0.8, 1.2, 0.05])
    p = orc.Pop.generic(L, cm, np.zeros(4), np.array([-65.0, -64.0, -66.0, -63.0]))
    rho = p.rkcrho()
    J = (L - np.diag(L.sum(axis=1))) / cm[:, None]
    lam = np.max(np.abs(np.linalg.eigvals(J)))
    assert rho == pytest.approx(1.2 * lam, rel=5e-3)


def test_rkcint_is_the_hermite_cubic():
    p = two_comp()
    t0, t1 = 0.3, 0.3 + 0.05
    coef = np.array([[1.5, -2.0, 0.7, 3.0], [-65.0, 4.0, -1.0, 0.25]])   # p(t) per compartment

    def val(t):
        return np.array([np.polyval(c[::-1], t) for c in coef])

    def der(t):
        return np.array([np.polyval(np.polyder(c[::-1]), t) for c in coef])

    p.set(orc.VT1, val(t0))
    p.set(orc.YN, val(t1))
    p.set(orc.VT2, der(t0))
    p.set(orc.FN, der(t1))
    p.set_hwork(0, t1 - t0)
    p.set_hwork(1, t1)
    for t in (t0, t0 + 0.013, t0 + 0.031, t1):
        np.testing.assert_allclose(p.rkcint(t), val(t), rtol=1e-12, atol=1e-12)
    assert np.array_equal(p.rkcint(t1), val(t1))   # s = 1: y_{n+1} exactly


def run_decay(tol, g=1.0, nsteps=8):
    p = two_comp(g, (1.0, 1.0), y=(1.0, -1.0), tol=tol)
    for _ in range(nsteps):
        p.advance()
    t = nsteps * 0.125
    d = p.get(orc.VNEW)
    exact = 2.0 * math.exp(-2.0 * g * t)
    return d[0] - d[1], exact, p.stats()


def test_controller_error_and_step_growth():
    e6, x6, s6 = run_decay(1e-6)
    e9, x9, s9 = run_decay(1e-9)
    # local errors of size ~tol accumulate over the accepted steps
    assert abs(e6 - x6) <= 2.0 * s6["steps"] * 1e-6 * 2.0
    assert abs(e9 - x9) <= 2.0 * s9["steps"] * 1e-9 * 2.0
    ratio = s9["steps"] / s6["steps"]
    assert 4.0 <= ratio <= 25.0, ratio     # ~ (1e3)^(1/3) = 10 for a 3rd-order local error


def test_rejected_step_below_hmin_is_accepted():
    # a stiff closed cable at a tight tolerance: every shrink lands below
    # hmin, the reference then sets idid = 4 and accepts the step
    # (solve_rkc.cu:1006-1019), so rejects are counted but never repeated
    # and the spectral radius is not re-estimated after them
    p = two_comp(25.0, tol=1e-9)
    init_iters = p.stats()["rho_iters"]
    for _ in range(8):
        p.advance()
    s = p.stats()
    assert s["rejects"] > 0
    assert s["rho_iters"] == init_iters + 2   # only the first rkclow call runs rkcrho
