"""Pins the oracle's point-source update and the CFL time step.

* ``orc_src_apply`` (PointSources::apply, src/impl/seismic/setups/
  PointSources.hpp:887-930, with the basis evaluations bEvals of :470-494:
  the basis at the source's reference point times the inverse diagonal mass
  over the Jacobian determinant).  A point impulse of strength I = the
  source's time integral, projected onto the element's polynomials, is the
  polynomial u with  int_el psi u dx = I sca_q psi(x_s)  for every psi of
  degree <= O - 1.  The test checks that identity for physical monomials psi
  on a quadrature of its own (a collapsed Gauss-Legendre rule built here,
  mapped to the physical element), so the update's scaling (mass matrix,
  Jacobian), its placement (element, quantity, mode, run) and its time
  factor are each pinned: a dropped 1/|J|, a transposed index or a wrong run
  fails it.  A unit series gives I = t2 - t1 exactly; splitting the window
  adds up; nothing but (el, :, :, run) changes.
* the in-diameter of tools/edge_v/src/geom/Tet4.cpp:49-78 (12 V / sum of the
  face normals' lengths) against the reference's own test value
  (Tet4.test.cpp:44-55, Geom.test.cpp:391-396) and the closed forms of the
  regular tet (edge / sqrt 6) and of the box mesh's tets.
* ``orc_dt_cfl`` (common.hpp:84-98 with SCALE_CFL = 1/2, const.inc:135-137)
  against the reference's getTimeStepCFL compiled into oracle/_ref.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as orc
from edge_amd import mesh as M
from edge_amd import setups as S


def _tet_rule(n):
    """Collapsed (Duffy) Gauss-Legendre rule on the unit tet, n points per
    direction: exact for polynomials of degree <= 2 n - 3."""
    x, w = np.polynomial.legendre.leggauss(n)
    x, w = 0.5 * (x + 1), 0.5 * w
    pts, wts = [], []
    for a, wa in zip(x, w):
        for b, wb in zip(x, w):
            for c, wc in zip(x, w):
                pts.append([a, b * (1 - a), c * (1 - a) * (1 - b)])
                wts.append(wa * wb * wc * (1 - a) ** 2 * (1 - b))
    return np.array(pts), np.array(wts)


def _monomials(deg):
    return [(i, j, k) for i in range(deg + 1) for j in range(deg + 1 - i) for k in range(deg + 1 - i - j)]


def _bev(order, ve, xi):
    """bEvals of PointSources.hpp:470-494: basis / mass / det J."""
    jac = np.array([ve[1] - ve[0], ve[2] - ve[0], ve[3] - ve[0]])  # jac[i][j] = dx_j / dxi_i
    return S.eval_basis(order, xi[None])[0] / S.dg_matrices(order)["mass"] / np.linalg.det(jac)


def _problem(order=4, runs=3, seed=11):
    m = M.box(3, jitter=0.15, seed=seed)
    rng = np.random.default_rng(seed)
    els = rng.choice(m.n_el, 4, replace=False)
    xi = rng.dirichlet(np.ones(4), 4)[:, 1:] * 0.9 + 0.02
    run = rng.integers(0, runs, 4).astype(np.int32)
    ve = m.el_crds()
    bev = np.array([_bev(order, ve[e], x) for e, x in zip(els, xi)])
    xs = np.array([ve[e][0] + x @ np.array([ve[e][1] - ve[e][0], ve[e][2] - ve[e][0], ve[e][3] - ve[e][0]])
                   for e, x in zip(els, xi)])
    return m, els, run, bev, xs


def _moments(order, ve, u, deg):
    """int_el psi u dx for the physical monomials psi (about the centroid)."""
    pts, w = _tet_rule(order + 2)
    jac = np.array([ve[1] - ve[0], ve[2] - ve[0], ve[3] - ve[0]])
    x = ve[0] + pts @ jac
    c = ve.mean(axis=0)
    vals = S.eval_basis(order, pts) @ u  # u at the quadrature points
    wj = w * abs(np.linalg.det(jac))
    return np.array([np.sum(wj * np.prod((x - c) ** np.array(e), axis=1) * vals) for e in _monomials(deg)])


def test_src_apply_projects_a_unit_impulse():
    order, runs = 4, 3
    m, els, run, bev, xs = _problem(order, runs)
    n = S.n_modes(order)
    t1, t2 = 0.13, 0.71
    for q in (0, 4, 8):
        sca = np.zeros((len(els), 9))
        sca[:, q] = 1.0
        dofs = np.zeros((m.n_el, 9, n, runs))
        series = [np.ones(40) for _ in els]
        orc.src_apply(dofs, els, run, np.zeros(len(els)), np.full(len(els), 0.05), bev, sca, series, t1, t2)
        for i, (e, r) in enumerate(zip(els, run)):
            ve = m.el_crds()[e]
            got = _moments(order, ve, dofs[e, q, :, r], order - 1)
            c = ve.mean(axis=0)
            want = (t2 - t1) * np.array([np.prod((xs[i] - c) ** np.array(k)) for k in _monomials(order - 1)])
            np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-13)
            others = np.delete(dofs[e, :, :, r], q, axis=0)
            assert not others.any()
        mask = np.ones(dofs.shape, bool)
        for e, r in zip(els, run):
            mask[e, q, :, r] = False
        assert not dofs[mask].any()


def test_src_apply_scales_by_the_series_integral_and_splits():
    order, runs = 3, 2
    m, els, run, bev, xs = _problem(order, runs, seed=12)
    n = S.n_modes(order)
    rng = np.random.default_rng(5)
    sca = rng.normal(size=(len(els), 9))
    series = [rng.normal(size=60) for _ in els]
    t0, dts = rng.uniform(-0.1, 0.0, len(els)), rng.uniform(0.02, 0.04, len(els))
    t1, tm, t2 = 0.21, 0.55, 0.93
    whole = np.zeros((m.n_el, 9, n, runs))
    orc.src_apply(whole, els, run, t0, dts, bev, sca, series, t1, t2)
    split = np.zeros_like(whole)
    orc.src_apply(split, els, run, t0, dts, bev, sca, series, t1, tm)
    orc.src_apply(split, els, run, t0, dts, bev, sca, series, tm, t2)
    np.testing.assert_allclose(split, whole, rtol=1e-12, atol=1e-14)
    for i, (e, r) in enumerate(zip(els, run)):
        # the integral of the linearly interpolated series over [t1, t2]
        xk = t0[i] + dts[i] * np.arange(len(series[i]))
        xq = np.concatenate([[t1], xk[(xk > t1) & (xk < t2)], [t2]])
        inte = np.trapezoid(np.interp(xq, xk, series[i]), xq)
        ve = m.el_crds()[e]
        c = ve.mean(axis=0)
        psi = np.array([np.prod((xs[i] - c) ** np.array(k)) for k in _monomials(order - 1)])
        for q in range(9):
            got = _moments(order, ve, whole[e, q, :, r], order - 1)
            np.testing.assert_allclose(got, inte * sca[i, q] * psi, rtol=1e-10, atol=1e-12)


def test_in_diameter_matches_the_reference_test_and_closed_forms():
    # Tet4.test.cpp:44-55 and Geom.test.cpp:391-396
    ve = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.5, 1.0, 0.0], [0.5, 0.5, 1.0]])
    m = M.from_tets(ve, [[0, 1, 2, 3]])
    assert m.in_dia[0] == pytest.approx(0.45358449083204, rel=1e-12)
    # regular tet of edge a: insphere radius a / (2 sqrt 6)
    a = 1.7
    reg = a * np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]]) / np.sqrt(8.0)
    m = M.from_tets(reg, [[0, 1, 2, 3]])
    assert m.in_dia[0] == pytest.approx(a / np.sqrt(6.0), rel=1e-13)
    # the box mesh cuts each hex of side h into 6 tets of volume h^3 / 6 with
    # two right-isosceles faces (legs h) and two right faces (legs h, h sqrt 2):
    # in-diameter 6 V / area = h / (1 + sqrt 2)
    h = 0.25
    m = M.box(4)
    np.testing.assert_allclose(m.in_dia, h / (1 + np.sqrt(2.0)), rtol=1e-13)
    np.testing.assert_allclose(m.volume, h ** 3 / 6, rtol=1e-13)


def test_dt_cfl_matches_the_reference():
    ref = orc.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (no reference tree)")
    f = ref.ref_dt_cfl
    f.restype = C.c_double
    f.argtypes = [C.c_double] * 5
    rng = np.random.default_rng(2)
    for _ in range(50):
        rho, vp, d = rng.uniform(1, 3), rng.uniform(2, 6), rng.uniform(0.01, 1)
        mu = rho * (vp / np.sqrt(3)) ** 2
        lam = rho * vp ** 2 - 2 * mu
        want = f(rho, lam, mu, d, 0.5)  # the reference is built for order 5
        assert orc.dt_cfl(rho, lam, mu, d, 5) == pytest.approx(want, rel=1e-15)
        m_ = M.box(3)
        lam_a, mu_a, rho_a = np.full(m_.n_el, lam), np.full(m_.n_el, mu), np.full(m_.n_el, rho)
        assert S.dt_cfl(m_, rho_a, lam_a, mu_a, 5) == pytest.approx(
            orc.dt_cfl(rho, lam, mu, m_.in_dia.min(), 5), rel=1e-15)
