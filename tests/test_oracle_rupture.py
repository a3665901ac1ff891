"""Pins the oracle's dynamic-rupture restatement: linear slip-weakening
friction (src/impl/seismic/solvers/FrictionLaws.hpp:309-456) and the
internal-boundary net updates with their face solvers
(InternalBoundary.hpp:184-280, :743-923).

* Jump and flux solvers against the reference's generated sources
  (generated/MiddleState{JumpL,JumpR,Flux}3D.inc, built into oracle/_ref).
* The integrated net update against the expected perturbed middle states of
  the reference's 3D internal-boundary test (InternalBoundary.test.cpp:339-628).
* Physical invariants: a locked fault gives the welded-interface fluxes; a
  failing fault's traction magnitude equals its strength; the boundary
  rotation maps one normal onto the other."""
import ctypes as C

import numpy as np
import pytest

import oracle as orc


def _frame(rng):
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    s = np.cross(n, rng.normal(size=3))
    s /= np.linalg.norm(s)
    return n, s, np.cross(n, s)


def _mats(rng):
    rho = rng.uniform(2000, 3000)
    vp = rng.uniform(4000, 6500)
    vs = vp / rng.uniform(1.6, 2.0)
    mu = rho * vs * vs
    return rho * vp * vp - 2 * mu, mu, rho


def test_ms_solvers_match_reference_generated():
    r = orc.ref_lib()
    if r is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(21)
    P = lambda a: C.c_void_p(a.ctypes.data)
    d = C.c_double
    for _ in range(30):
        (lL, mL, rL), (lR, mR, rR) = _mats(rng), _mats(rng)
        for side, f in (("l", r.ref_ms_jump_l), ("r", r.ref_ms_jump_r)):
            ref = np.zeros(81)
            f(d(lL), d(lR), d(mL), d(mR), d(rL), d(rR), P(ref))
            np.testing.assert_allclose(orc.ms_jump(lL, lR, mL, mR, rL, rR, side), ref.reshape(9, 9), rtol=1e-13,
                                       atol=1e-13 * np.abs(ref).max())
        n, s, t = _frame(rng)
        ref = np.zeros(81)
        r.ref_ms_flux(P(n), P(s), P(t), d(lL), d(mL), d(rL), P(ref))
        np.testing.assert_allclose(orc.ms_flux(n, s, t, lL, mL, rL), ref.reshape(9, 9), rtol=1e-13,
                                   atol=1e-13 * np.abs(ref).max())


def _test3d(rupture_fx):
    """Inputs of the reference's 3D internal-boundary test (identity frame)."""
    lL, mL, rL = 32080749540.0, 32019625230.0, 2670.0
    lR, mR, rR = lL / 2, mL / 2, 2600.0
    e = np.eye(3)
    solv = np.zeros((4, 9, 9))
    solv[0] = np.eye(9)
    solv[1] = orc.ms_jump(lL, lR, mL, mR, rL, rR, "l")
    solv[2] = orc.ms_flux(e[0], e[1], e[2], lL, mL, rL)
    solv[3] = orc.ms_flux(e[0], e[1], e[2], lR, mR, rR)
    csL, csR = np.sqrt(mL / rL), np.sqrt(mR / rR)
    ss0 = (40e6, np.sqrt(81.6e6 ** 2 - 40e6 ** 2))
    state = orc.lsw_state(1, sn0=-120e6, ss0_0=ss0[0], ss0_1=ss0[1], ss0A=81.6e6, muf=0.59, dd_0=0.2, dd_1=-0.1)
    glob = np.array([[0.677], [0.55], [2.5]])
    dL = np.array(rupture_fx["qL"]).reshape(9, 1, 1)
    dR = np.array(rupture_fx["qR"]).reshape(9, 1, 1)
    # the left element is the plus side (lEqM = false): csDmuP from the left material
    return solv, dL, dR, glob, state, csR / mR, csL / mL


def test_net_updates_match_reference_3d_test(rupture_fx):
    solv, dL, dR, glob, state, cm, cp = _test3d(rupture_fx)
    netL, netR, per = orc.net_updates(solv, dL, dR, 1.0, False, cm, cp, glob, state)
    assert per.all()
    expL = -(solv[2] @ np.array(rupture_fx["msLp"]))
    expR = solv[3] @ np.array(rupture_fx["msRp"])
    np.testing.assert_allclose(netL[:, 0, 0], expL, rtol=1e-12, atol=1e-12 * np.abs(expL).max())
    np.testing.assert_allclose(netR[:, 0, 0], expR, rtol=1e-12, atol=1e-12 * np.abs(expR).max())


def test_locked_fault_gives_welded_interface_fluxes(rupture_fx):
    """Below the strength the middle state is not perturbed: the net updates
    are the Godunov fluxes of the welded middle state, and slip and friction
    stay put."""
    solv, dL, dR, glob, state, cm, cp = _test3d(rupture_fx)
    state[:, 2:5] = [1e6, 0.0, 1e6]  # far below the strength mu_f |sigma_n|
    st0 = state.copy()
    netL, netR, per = orc.net_updates(solv, dL, dR, 0.01, False, cm, cp, glob, state)
    assert not per.any()
    ms = dL[:, 0, 0] + solv[1] @ (dR[:, 0, 0] - dL[:, 0, 0])
    np.testing.assert_allclose(netL[:, 0, 0], -0.01 * solv[2] @ ms, rtol=1e-13, atol=1e-6)
    np.testing.assert_allclose(netR[:, 0, 0], 0.01 * solv[3] @ ms, rtol=1e-13, atol=1e-6)
    np.testing.assert_array_equal(state[:, 11:13], st0[:, 11:13])  # no slip


def test_failing_fault_traction_equals_strength():
    rng = np.random.default_rng(3)
    for _ in range(50):
        ms = rng.normal(size=9) * np.array([1e7] * 6 + [1.0] * 3)
        ms[0] = -abs(ms[0])
        st = orc.lsw_state(1, sn0=-rng.uniform(50e6, 150e6), ss0_0=rng.uniform(50e6, 90e6),
                           ss0_1=rng.uniform(-20e6, 20e6), muf=rng.uniform(0.5, 0.7), dd_0=rng.uniform(0, 0.3))[0]
        msM, msP = np.zeros(9), np.zeros(9)
        per = C.c_int()
        P = lambda a: C.c_void_p(a.ctypes.data)
        orc.lib().orc_lsw(C.c_double(0.01), C.c_double(2e-7), C.c_double(3e-7), C.c_double(0.677), C.c_double(0.525),
                          C.c_double(2.5), P(st), P(ms), P(msM), P(msP), C.byref(per))
        if per.value:
            assert np.hypot(st[7], st[8]) == pytest.approx(st[6], rel=1e-12)  # |traction| = strength
            assert 0.525 <= st[5] <= 0.677


def test_boundary_rotation_maps_normals():
    rng = np.random.default_rng(5)
    for _ in range(20):
        a, b = rng.normal(size=3), rng.normal(size=3)
        R = orc.rot_mat(a, b)
        np.testing.assert_allclose(R @ (a / np.linalg.norm(a)), b / np.linalg.norm(b), atol=1e-14)
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-14)
        assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-14)


def _face_case(rng, n_sf=2, R=3):
    n, s, t = _frame(rng)
    bnd = np.array(_frame(rng))
    if n @ bnd[0] < 0:  # the fault normal points from the left element into the right one
        n, t = -n, -t
    mA, mB = _mats(rng), _mats(rng)
    matA, matB = (mA[2], mA[0], mA[1]), (mB[2], mB[0], mB[1])  # (rho, lambda, mu)
    area, vA, vB = rng.uniform(0.5, 2), rng.uniform(0.5, 2), rng.uniform(0.5, 2)
    dA = rng.normal(size=(9, n_sf, R)) * np.array([1e6] * 6 + [1.0] * 3)[:, None, None]
    dB = rng.normal(size=(9, n_sf, R)) * np.array([1e6] * 6 + [1.0] * 3)[:, None, None]
    glob = np.array([[0.677] * R, [0.525] * R, [2.5] * R])
    st = orc.lsw_state(n_sf * R, sn0=-rng.uniform(50e6, 60e6), ss0_0=rng.uniform(20e6, 40e6, n_sf * R),
                       ss0_1=rng.uniform(-5e6, 5e6, n_sf * R), muf=0.677, dd_0=rng.uniform(0, 0.1, n_sf * R))
    return n, s, t, bnd, matA, matB, area, vA, vB, dA, dB, glob, st


def test_swapped_sides_give_the_same_updates():
    """A fault face seen from either element: the face normal of the second
    view opposes the fault normal, which selects the non-preferred branch of
    the solver set-up (flipped normal, right-side jump solver negated, flux
    solvers negated, the left element the minus side).  Both views must give
    each element the same update and the same friction state."""
    rng = np.random.default_rng(8)
    n_fail = 0
    for _ in range(20):
        n, s, t, bnd, matA, matB, area, vA, vB, dA, dB, glob, st = _face_case(rng)
        n_sf = dA.shape[1]
        sA, pA = orc.ib_solvers(n, s, t, bnd, matA, matB, area, vA, vB, n_sf, 4)
        sB, pB = orc.ib_solvers(-n, s, t, bnd, matB, matA, area, vB, vA, n_sf, 4)
        assert pA and not pB
        cd = lambda m: np.sqrt(1.0 / (m[0] * m[2]))
        stA, stB = st.copy(), st.copy()
        # view A: A left and plus; view B: B left and minus
        aL, aR, perA = orc.net_updates(sA, dA, dB, 1e-3, False, cd(matB), cd(matA), glob, stA)
        bL, bR, perB = orc.net_updates(sB, dB, dA, 1e-3, True, cd(matB), cd(matA), glob, stB)
        np.testing.assert_allclose(aL, bR, rtol=1e-10, atol=1e-10 * np.abs(aL).max())
        np.testing.assert_allclose(aR, bL, rtol=1e-10, atol=1e-10 * np.abs(aR).max())
        assert (np.abs(stA - stB) <= 1e-10 * (np.abs(stA).max(axis=0) + 1e-30)).all()
        np.testing.assert_array_equal(perA, perB)
        n_fail += int(perA.sum())
    assert n_fail > 0  # the cases exercise failing sub-faces
