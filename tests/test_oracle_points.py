"""Pins the oracle's point-source and receiver pieces:

* ``orc_series_integrate`` (Series<1>::integrate, src/linalg/Series.hpp:59-133)
  against every call of the reference's Series test (src/linalg/Series.test.cpp,
  extracted to tests/golden/ref_series_fixtures.json), and additivity over
  split intervals;
* ``orc_integrate`` (TimePred::integrate) against the [0, dt] and [0, dt/2]
  Taylor integrals of the reference's advection CK test
  (src/impl/advection/kernels/TimePred.test.inc; same formula, one quantity);
* ``orc_eval_time_pred`` (evalTimePrediction) through calculus: at t = 0 it is
  the zeroth derivative, and its Gauss-Legendre integral over [0, t] is the
  pinned Taylor integral.
"""
import json
import os

import numpy as np
import pytest

import oracle as orc

FX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_series_fixtures.json")


@pytest.fixture(scope="module")
def fx():
    with open(FX) as f:
        return json.load(f)


def test_series_integrate_matches_reference_tests(fx):
    vals = fx["series"]["vals"]
    assert len(fx["series"]["cases"]) == 13
    for c in fx["series"]["cases"]:
        got = orc.series_integrate(c["dx"], c["x0"], vals[:c["n"]], c["x1"], c["x2"], c["na"])
        assert got == pytest.approx(c["expected"], rel=1e-12, abs=1e-14), c


def test_series_integrate_is_additive():
    """Exact for a piecewise-linear series whenever the interval ends lie in
    different sampling intervals (inside one interval the reference
    extrapolates the part before x1 with the previous interval's slope, which
    the restatement keeps), so splitting [a, b] at such an m adds up."""
    rng = np.random.default_rng(4)
    for _ in range(200):
        n = int(rng.integers(8, 40))
        v = rng.normal(size=n)
        dx, x0 = rng.uniform(0.01, 0.3), rng.uniform(-1, 1)
        ia, im, ib = np.sort(rng.choice(n - 1, 3, replace=False))
        a, m, b = (x0 + (i + rng.uniform(0.05, 0.95)) * dx for i in (ia, im, ib))
        whole = orc.series_integrate(dx, x0, v, a, b)
        parts = orc.series_integrate(dx, x0, v, a, m) + orc.series_integrate(dx, x0, v, m, b)
        assert whole == pytest.approx(parts, rel=1e-12, abs=1e-12)
        # the integral of the linear interpolant: trapezoids between its kinks
        xk = x0 + dx * np.arange(n)
        xs = np.concatenate([[a], xk[(xk > a) & (xk < b)], [b]])
        assert whole == pytest.approx(np.trapezoid(np.interp(xs, xk, v), xs), rel=1e-12, abs=1e-12)


def test_integrate_matches_reference_advection_test(fx):
    adv = fx["advection_time_pred"]
    dg = orc.Dg(3, 1)
    der = np.zeros((3, 9, 10, 1))
    der[:, 0, :, 0] = np.array(adv["ders"])
    for dt, exp in ((adv["dt"], adv["tint"]), (0.5 * adv["dt"], adv["tint12"])):
        got = orc.integrate(dg, dt, der)[0, :, 0]
        np.testing.assert_allclose(got, exp, rtol=2e-6, atol=1e-9)


def test_eval_time_pred_is_the_derivative_of_the_integral():
    rng = np.random.default_rng(6)
    dg = orc.Dg(5, 2)
    der = rng.normal(size=(5, 9, 35, 2))
    np.testing.assert_array_equal(orc.eval_time_pred(dg, 0.0, der), der[0])
    xg, wg = np.polynomial.legendre.leggauss(4)  # exact for degree 7 >= O - 1
    for t in (0.3, 1.0, 2.5):
        quad = sum(w * 0.5 * t * orc.eval_time_pred(dg, 0.5 * t * (x + 1), der) for x, w in zip(xg, wg))
        np.testing.assert_allclose(quad, orc.integrate(dg, t, der), rtol=1e-12, atol=1e-12)
