"""The generated reference-element matrices against the reference's own
fixtures, against independent numerical quadrature, and against the
integration-by-parts identity that ties stiffness and flux matrices."""
import numpy as np
import pytest

from edge_amd.dg import gen_tet4
from edge_amd.setups import dg_matrices, eval_basis, quad_tet


def test_stiffness_degree3_matches_reference_fixture(kernel_fx):
    """stiffT / stiffV at degree 3 equal the tables of
    src/impl/seismic/kernels/TimePred.test.inc:29-90 / VolInt.test.inc (bit-exact)."""
    m = dg_matrices(4)
    for d in range(3):
        st = np.array(kernel_fx["timepred"]["stiffT"][d]).reshape(20, 20)
        sv = np.array(kernel_fx["volint"]["stiffV"][d]).reshape(20, 20)
        assert np.array_equal(m["stiffT"][d], st)
        assert np.array_equal(m["stiffV"][d], sv)


def test_flux_degree2_matches_reference_fixture(kernel_fx):
    """fluxL, fluxN[6], fluxT at degree 2 equal src/impl/seismic/kernels/SurfInt.test.inc."""
    m = dg_matrices(3)
    si = kernel_fx["surfint"]
    assert np.array_equal(m["fluxL"], np.array(si["fluxL"]).reshape(4, 10, 6))
    assert np.array_equal(m["fluxN"][6], np.array(si["fluxN6"]).reshape(10, 6))
    assert np.array_equal(m["fluxT"], np.array(si["fluxT"]).reshape(4, 6, 10))


def test_raw_stiffness_degree2_matches_edge_pre_test(matrix_fx):
    """K_d = stiffV_d M at degree 2 equals the exact matrices of
    tools/edge_pre/edge_pre/int/test_Matrices.py:125-190."""
    m = dg_matrices(3)
    for d in range(3):
        k = m["stiffV"][d] * m["mass"][None, :]
        np.testing.assert_allclose(k, np.array(matrix_fx["stiff_deg2"][d]), rtol=0, atol=2e-16)


def test_flux_products_degree3_match_edge_pre_test(matrix_fx):
    """fm_fa = P_fa M_F P_fa^T and fp_{f1 f2 v} = P_f2 Ori_v P_f1^T of
    tools/edge_pre/edge_pre/int/test_Matrices.py:467-1668 (all 4 + 48)."""
    m = dg_matrices(4)
    g = gen_tet4.generate(3)
    mf = np.array([float(x) for x in g["mass_face"]])
    proj = [m["fluxL"][fa] / mf[None, :] for fa in range(4)]
    for fa in range(4):
        fm = proj[fa] @ np.diag(mf) @ proj[fa].T
        np.testing.assert_allclose(fm, np.array(matrix_fx["fm_deg3"][fa]), rtol=0, atol=1e-15)
    for f1 in range(4):
        for f2 in range(4):
            for v in range(3):
                fp = m["fluxN"][v * 4 + f2] @ proj[f1].T
                ref = np.array(matrix_fx["fp_deg3"]["fp%d%d%d" % (f1, f2, v)])
                np.testing.assert_allclose(fp, ref, rtol=0, atol=1e-15)


@pytest.mark.parametrize("order", [2, 3, 4, 5])
def test_basis_matches_reference_generated_basis(order, matrix_fx):
    """Library basis evaluation vs src/dg/generated/basis_tet.inc values."""
    pts = np.array(matrix_fx["basis_tet_pts"])
    ref = np.array(matrix_fx["basis_tet_val"])
    n = order * (order + 1) * (order + 2) // 6
    val = eval_basis(order, pts)
    np.testing.assert_allclose(val.T, ref[:n], rtol=1e-12, atol=1e-12)


def test_generator_face_basis_matches_reference(matrix_fx):
    pts = np.array(matrix_fx["basis_tria_pts"])
    ref = np.array(matrix_fx["basis_tria_val"])
    psi = gen_tet4.tria_basis(5)
    val = np.array([[sum(float(c) * x ** e[0] * y ** e[1] for e, c in p.items()) for x, y in pts] for p in psi])
    np.testing.assert_allclose(val, ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("order", [2, 3, 4, 5])
def test_stiffT_reproduces_reference_basis_derivatives(order, matrix_fx):
    """d/dxi_d phi_j = sum_m stiffT_d[j][m] phi_m: the derivative polynomials of
    src/dg/generated/basis_tet_der_xi*.inc pin stiffT (and with the mass,
    K_d and stiffV) at every generated degree."""
    m = dg_matrices(order)
    n = m["mass"].size
    phi = np.array(matrix_fx["basis_tet_val"])[:n]
    for d in range(3):
        der = np.array(matrix_fx["basis_tet_der%d" % d])[:n]
        np.testing.assert_allclose(m["stiffT"][d] @ phi, der, rtol=1e-11, atol=1e-10)
        # stiffV_d = K_d M^-1 with K_d = (stiffT_d M)^T
        k = (m["stiffT"][d] * m["mass"][None, :]).T
        np.testing.assert_allclose(m["stiffV"][d], k / m["mass"][None, :], rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("order", [2, 3, 4, 5])
def test_mass_by_quadrature(order):
    pts, w = quad_tet(order + 2)
    b = eval_basis(order, pts)
    mq = (b * w[:, None]).T @ b
    m = dg_matrices(order)["mass"]
    np.testing.assert_allclose(mq, np.diag(m), rtol=0, atol=1e-14)


def _tri_rule(n):
    x, w = np.polynomial.legendre.leggauss(n)
    pts, wts = [], []
    for i in range(n):
        for j in range(n):
            a, b = x[i], x[j]
            pts.append(((1 + a) * (1 - b) / 4, (1 + b) / 2))
            wts.append(w[i] * w[j] * (1 - b) / 8)
    return np.array(pts), np.array(wts)


@pytest.mark.parametrize("order", [2, 3, 4, 5])
def test_flux_matrices_by_quadrature(order):
    """fluxL / fluxN / fluxT from numerical face quadrature of the (fixture-
    pinned) bases and the face maps of tools/edge_pre/edge_pre/types/Tet.py."""
    m = dg_matrices(order)
    psi = gen_tet4.tria_basis(order - 1)
    cp, cw = _tri_rule(order + 2)

    def ev_psi(xy):
        return np.array([[sum(float(c) * x ** e[0] * y ** e[1] for e, c in p.items()) for x, y in xy] for p in psi])

    P = ev_psi(cp)                                   # [f][qp]
    mf = (P * cw[None, :]) @ P.T
    assert np.allclose(mf, np.diag(np.diag(mf)), atol=1e-14)
    mf = np.diag(mf)
    c1, c2 = cp[:, 0], cp[:, 1]
    z = np.zeros_like(c1)
    fa_to_el = [np.stack([c2, c1, z], 1), np.stack([c1, z, c2], 1), np.stack([z, c2, c1], 1),
                np.stack([1 - c1 - c2, c1, c2], 1)]
    fa_to_fa = [np.stack([c2, c1], 1), np.stack([1 - c1 - c2, c2], 1), np.stack([c1, 1 - c1 - c2], 1)]
    fl = []
    for fa in range(4):
        phi = eval_basis(order, fa_to_el[fa])         # [qp][N]
        fl.append((phi * cw[:, None]).T @ P.T)        # [N][NF]
        np.testing.assert_allclose(m["fluxL"][fa], fl[-1], rtol=0, atol=1e-13)
        np.testing.assert_allclose(m["fluxT"][fa], (fl[-1] / mf[None, :]).T / m["mass"][None, :], rtol=1e-12,
                                   atol=1e-11)
    for v in range(3):
        ori = (ev_psi(fa_to_fa[v]) * cw[None, :]) @ P.T   # [g][f]
        for fa in range(4):
            np.testing.assert_allclose(m["fluxN"][v * 4 + fa], (fl[fa] / mf[None, :]) @ ori, rtol=0, atol=1e-13)


@pytest.mark.parametrize("order", [2, 3, 4, 5])
def test_integration_by_parts(order):
    """K_d + K_d^T = sum_fa n_{fa,d} 2|fa| (fluxL_fa M_F^-1 fluxL_fa^T): a
    transposed stiffness, a wrong face map or a sign slip breaks it."""
    m = dg_matrices(order)
    g = gen_tet4.tria_basis(order - 1)
    mf = np.array([float(gen_tet4.int_tria(gen_tet4.p_mul(p, p))) for p in g])
    fm = [m["fluxL"][fa] @ np.diag(1 / mf) @ m["fluxL"][fa].T for fa in range(4)]
    minus_face = {0: 2, 1: 1, 2: 0}  # face with outward normal -e_d
    for d in range(3):
        k = (m["stiffT"][d] * m["mass"][None, :]).T
        np.testing.assert_allclose(k + k.T, fm[3] - fm[minus_face[d]], rtol=0, atol=1e-13)
