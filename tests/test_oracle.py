"""Pins the oracle (oracle/edge_oracle.c) before it is trusted as the checker:
kernels against the reference's golden vectors, set-up against the
reference's own Elasticity.cpp (oracle/_ref) and against closed-form
properties of the Godunov flux, and the whole scheme against the analytic
plane wave (convergence order p+1)."""
import ctypes as C

import numpy as np
import pytest

import oracle as orc
from edge_amd import mesh as M
from edge_amd import setups as S


def _a(x, *shape):
    a = np.array(x, dtype=np.float64)
    return a.reshape(shape) if shape else a


# --------------------------------------------------------------- kernels
def test_ck_elastic_matches_timepred_fixture(kernel_fx):
    """TimePredVanilla.test.cpp:31-128 with the fixture's degree-3 inputs."""
    tp = kernel_fx["timepred"]
    dg = orc.Dg(4, 1, mats={"stiffT": _a(tp["stiffT"], 3, 20, 20)})
    _, tint, _, _ = orc.ck(dg, tp["dt"], _a(tp["starE"], 3, 81), _a(tp["dofsE"])[..., None])
    np.testing.assert_allclose(tint[..., 0], _a(tp["refEtDofs"]), rtol=1e-13, atol=1e-13)


def test_ck_viscoelastic_matches_timepred_fixture(kernel_fx):
    tp = kernel_fx["timepred"]
    dg = orc.Dg(4, 1, mats={"stiffT": _a(tp["stiffT"], 3, 20, 20)}, rfs=_a(tp["rfs"]))
    _, tint, _, tint_a = orc.ck(dg, tp["dt"], _a(tp["starE"], 3, 81), _a(tp["dofsE"])[..., None],
                                _a(tp["starA"], 3, 18), _a(tp["srcA"], 2, 36), _a(tp["dofsA"])[..., None])
    np.testing.assert_allclose(tint[..., 0], _a(tp["refVtDofsE"]), rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(tint_a[..., 0], _a(tp["refVtDofsA"]), rtol=1e-13, atol=1e-15)


def test_volume_matches_volint_fixture(kernel_fx):
    vi = kernel_fx["volint"]
    dg = orc.Dg(4, 1, mats={"stiffV": _a(vi["stiffV"], 3, 20, 20)})
    out, _ = orc.vol(dg, _a(vi["starE"], 3, 81), _a(vi["tDofsE"])[..., None], _a(vi["dofsE"])[..., None])
    np.testing.assert_allclose(out[..., 0], _a(vi["refEdofs"]), rtol=1e-13, atol=1e-12)
    dg = orc.Dg(4, 1, mats={"stiffV": _a(vi["stiffV"], 3, 20, 20)}, rfs=_a(vi["rfs"]))
    out, out_a = orc.vol(dg, _a(vi["starE"], 3, 81), _a(vi["tDofsE"])[..., None], _a(vi["dofsE"])[..., None],
                         _a(vi["starA"], 3, 18), _a(vi["srcA"], 2, 36), _a(vi["tDofsA"])[..., None],
                         _a(vi["dofsA"])[..., None])
    np.testing.assert_allclose(out[..., 0], _a(vi["refVdofsE"]), rtol=1e-13, atol=1e-12)
    np.testing.assert_allclose(out_a[..., 0], _a(vi["refVdofsA"]), rtol=1e-13, atol=1e-12)


def _surf_dg(si, rfs=None):
    fn = np.zeros((12, 10, 6))
    fn[6] = _a(si["fluxN6"], 10, 6)
    return orc.Dg(3, 1, mats={"fluxL": _a(si["fluxL"], 4, 10, 6), "fluxN": fn, "fluxT": _a(si["fluxT"], 4, 6, 10)},
                  rfs=rfs)


def test_surface_matches_surfint_fixture(kernel_fx):
    """SurfIntVanilla.test.cpp:29-211: local, neighbour (fa 3, vId 1, fId 2) and free surface (fa 2)."""
    si = kernel_fx["surfint"]
    dg = _surf_dg(si)
    fs = _a(si["fSolvE"], 4, 81)
    t, d = _a(si["tDofsE"])[..., None], _a(si["dofsE"])[..., None]
    out, _ = orc.surf_local(dg, fs, t, d)
    np.testing.assert_allclose(out[..., 0], _a(si["refEdofs"]), rtol=1e-13, atol=1e-12)
    out, _ = orc.surf_neigh(dg, 3, 1, 2, fs[0], t, d)
    np.testing.assert_allclose(out[..., 0], _a(si["refEneighDofs"]), rtol=1e-13, atol=1e-12)
    out, _ = orc.surf_neigh(dg, 2, orc.NO_ID16, orc.NO_ID16, fs[0], t, d)
    np.testing.assert_allclose(out[..., 0], _a(si["refEneighFdofs"]), rtol=1e-13, atol=1e-12)


def test_surface_viscoelastic_matches_surfint_fixture(kernel_fx):
    si = kernel_fx["surfint"]
    dg = _surf_dg(si, rfs=_a(si["rfs"]))
    fs, fsa = _a(si["fSolvE"], 4, 81), _a(si["fSolvA"], 4, 54)
    t, d = _a(si["tDofsE"])[..., None], _a(si["dofsE"])[..., None]
    _, out_a = orc.surf_local(dg, fs, t, d, fsa, _a(si["dofsA"])[..., None])
    np.testing.assert_allclose(out_a[..., 0], _a(si["refVdofsA"]), rtol=1e-13, atol=1e-12)
    _, up = orc.surf_neigh(dg, 3, 1, 2, fs[0], t, d, fsa[0])
    np.testing.assert_allclose(up[..., 0], _a(si["refVneighDofsA"]), rtol=1e-13, atol=1e-12)
    _, up = orc.surf_neigh(dg, 2, orc.NO_ID16, orc.NO_ID16, fs[0], t, d, fsa[0])
    np.testing.assert_allclose(up[..., 0], _a(si["refVneighFdofsA"]), rtol=1e-13, atol=1e-12)


def test_fused_runs_are_independent():
    """R fused runs equal R single runs; permuting runs permutes outputs (the
    reference's fused tests duplicate one run and cannot see this)."""
    rng = np.random.default_rng(7)
    order, R = 4, 5
    st = rng.normal(size=(3, 81))
    dofs = rng.normal(size=(9, 20, R))
    dg1, dgR = orc.Dg(order, 1), orc.Dg(order, R)
    _, tR, _, _ = orc.ck(dgR, 0.01, st, dofs)
    for r in range(R):
        _, t1, _, _ = orc.ck(dg1, 0.01, st, dofs[..., r:r + 1])
        np.testing.assert_allclose(tR[..., r], t1[..., 0], rtol=1e-14, atol=1e-14)
    perm = rng.permutation(R)
    _, tp, _, _ = orc.ck(dgR, 0.01, st, dofs[..., perm])
    np.testing.assert_allclose(tp, tR[..., perm], rtol=1e-14, atol=1e-14)


# --------------------------------------------------------------- set-up
def _frame(rng):
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    s = np.cross(n, rng.normal(size=3))
    s /= np.linalg.norm(s)
    return n, s, np.cross(n, s)


def _ref():
    r = orc.ref_lib()
    if r is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    return r


def _P(a):
    return C.c_void_p(a.ctypes.data)


def test_star_matches_reference_elasticity():
    r = _ref()
    rng = np.random.default_rng(11)
    for _ in range(20):
        ji = rng.normal(size=(3, 3))
        rho, lam, mu = rng.uniform(1, 3), rng.uniform(0.5, 3), rng.uniform(0.5, 2)
        out = np.zeros(243)
        r.ref_star(C.c_double(rho), C.c_double(lam), C.c_double(mu), _P(ji), _P(out))
        np.testing.assert_array_equal(orc.star(rho, lam, mu, ji), out.reshape(3, 81))


@pytest.mark.parametrize("free_surface", [False, True])
def test_flux_solvers_match_reference_elasticity(free_surface):
    """oracle fsL / fsR vs Elasticity::fs with the reference's own trafos."""
    r = _ref()
    rng = np.random.default_rng(12 + free_surface)
    for _ in range(50):
        n, s, t = _frame(rng)
        if rng.random() < 0.5:  # the right element's frame (-n, s, t)
            n = -n
        mats = [rng.uniform(1, 3), rng.uniform(1, 3), rng.uniform(0.5, 3), rng.uniform(0.5, 3),
                rng.uniform(0.3, 2), rng.uniform(0.3, 2)]
        a, b = np.zeros(81), np.zeros(81)
        r.ref_fs(*[C.c_double(x) for x in mats], _P(n), _P(s), _P(t), C.c_int(int(free_surface)), _P(a), _P(b))
        fl, fr = orc.fs_face(*mats, n, s, t, free_surface)
        np.testing.assert_allclose(fl, a.reshape(9, 9), rtol=1e-13, atol=1e-14)
        np.testing.assert_allclose(fr, b.reshape(9, 9), rtol=1e-13, atol=1e-14)


def _jacobian_n(rho, lam, mu, n):
    """A_n = sum_k n_k A_k: with jinv = diag(n) in the first column, S_0 = A_n."""
    ji = np.zeros((3, 3))
    ji[:, 0] = n
    return orc.star(rho, lam, mu, ji)[0].reshape(9, 9)


def test_flux_solver_closed_form_properties():
    """Equal materials: fsL + fsR = A_n (consistency) and fsL / fsR are the
    positive / negative parts of A_n (eigenvalues c_p, c_s, c_s, 0 / their
    negatives).  Free surface: the right state is the traction mirror."""
    rng = np.random.default_rng(13)
    for _ in range(20):
        n, s, t = _frame(rng)
        rho, lam, mu = rng.uniform(1, 3), rng.uniform(0.5, 3), rng.uniform(0.3, 2)
        fl, fr = orc.fs_face(rho, rho, lam, lam, mu, mu, n, s, t)
        an = _jacobian_n(rho, lam, mu, n)
        np.testing.assert_allclose(fl + fr, an, rtol=1e-13, atol=1e-13)
        cp, cs = np.sqrt((lam + 2 * mu) / rho), np.sqrt(mu / rho)
        ev = np.sort(np.linalg.eigvals(fl).real)
        np.testing.assert_allclose(ev, np.sort([cp, cs, cs, 0, 0, 0, 0, 0, 0]), atol=1e-12)
        ev = np.sort(np.linalg.eigvals(fr).real)
        np.testing.assert_allclose(ev, np.sort([-cp, -cs, -cs, 0, 0, 0, 0, 0, 0]), atol=1e-12)
        # free surface: fsR_free = fsR_equal * T D T^-1 with D the traction mirror
        _, ffs = orc.fs_face(rho, rho, lam, lam, mu, mu, n, s, t, True)
        T, Ti = orc.trafo(n, s, t)
        D = np.diag([-1, 1, 1, -1, 1, -1, 1, 1, 1.0])
        np.testing.assert_allclose(ffs, fr @ T @ D @ Ti, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(T @ Ti, np.eye(9), atol=1e-14)


def test_flux_solvers_have_rank_three():
    """Different materials on both sides, free surfaces included: each flux
    solver has rank exactly 3 (the flux depends on a side's state only
    through its three waves travelling towards the face), the structure the
    MFMA kernels factor as U V^T.  A product of a rank-3 factor pair
    reproduces each solver to rounding."""
    rng = np.random.default_rng(17)
    for _ in range(50):
        n, s, t = _frame(rng)
        rl, rr = rng.uniform(1, 3, 2)
        vp = rng.uniform(3, 6, 2)
        vs = vp / rng.uniform(1.5, 2.5, 2)
        ml, mr = rl * vs[0] ** 2, rr * vs[1] ** 2
        ll, lr = rl * vp[0] ** 2 - 2 * ml, rr * vp[1] ** 2 - 2 * mr
        for free in (False, True):
            for f in orc.fs_face(rl, rr, ll, lr, ml, mr, n, s, t, free):
                sv = np.linalg.svd(f, compute_uv=False)
                assert sv[2] > 1e-3 * sv[0] and sv[3] < 1e-13 * sv[0], sv
                u, w, vt = np.linalg.svd(f)
                np.testing.assert_allclose((u[:, :3] * w[:3]) @ vt[:3], f, atol=1e-13 * sv[0])


# --------------------------------------------------------------- the scheme
@pytest.mark.parametrize("order", [2, 3, 4])
def test_oracle_plane_wave_convergence(order):
    """Errors of the P plane wave shrink like h^(p+1) (p = order - 1)."""
    pw = S.PlaneWave(1.0, 2.0, 1.0, k=(2 * np.pi,) * 3)
    errs = []
    for n in (4, 8):
        m = M.box(n)
        ne = m.n_el
        rho, lam, mu = np.ones(ne), 2 * np.ones(ne), np.ones(ne)
        mo = orc.MeshOracle(m, order, rho=rho, lam=lam, mu=mu)
        dofs = np.ascontiguousarray(S.project(m, order, pw))
        tint = np.zeros_like(dofs)
        T = 0.1
        dt = S.dt_cfl(m, rho, lam, mu, order)
        ns = int(np.ceil(T / dt))
        mo.step(dofs, tint, T / ns, ns)
        _, l2, _ = S.error_norms(m, order, dofs, lambda x: pw(x, T), n1d=order + 2)
        errs.append(np.sqrt(l2.sum()))
    rate = np.log2(errs[0] / errs[1])
    assert rate > order - 0.3, (errs, rate)


# --------------------------------------------------------------- viscoelastic set-up
# Inputs of the reference's viscoelastic set-up tests
# (src/impl/seismic/setups/ViscoElasticity.test.cpp:147-330): 3 mechanisms,
# central frequency 2, frequency ratio 10, Q_p = 200, Q_s = 300 (the Lame-form
# tests; the p/s test fits Q = 200 then Q = 300), lambda = 20.8e9, mu = 10.4e9.
def _golden(fx, key, name):
    return [(tuple(i), v) for n, i, v in fx[key] if n == name]


def test_visco_frequencies_and_fits_match_reference_tests(visco_fx):
    f = orc.visco_frequencies(3, 2, 10)
    for (i,), v in _golden(visco_fx, "frequencies", "freqs1"):
        assert f[i] == pytest.approx(v, rel=1e-12)
    fits = _golden(visco_fx, "coeff_ps", "fit1")
    for q, part in ((200.0, fits[:3]), (300.0, fits[3:])):
        c = orc.anelastic_coeff_ps(3, q, f)
        for (i,), v in part:
            assert c[i] == pytest.approx(v, rel=1e-11)


def test_visco_lame_and_source_match_reference_tests(visco_fx):
    f = [v for _, v in _golden(visco_fx, "frequencies", "freqs1")]
    fits = [v for _, v in _golden(visco_fx, "coeff_ps", "fit1")]
    le, me = orc.elastic_lame(3, 2 * np.pi * 2, f, fits[:3], fits[3:], 20.8e9, 10.4e9)
    gold = dict((n, v) for n, _, v in visco_fx["elastic_lame"])
    assert le == pytest.approx(gold["lamElastic"], rel=1e-12)
    assert me == pytest.approx(gold["muElastic"], rel=1e-12)
    le, me, co = orc.anelastic_coeff_lame(3, 2, 10, 200.0, 300.0, 20.8e9, 10.4e9)
    for (i,), v in _golden(visco_fx, "coeff_lame", "lameE"):
        assert (le, me)[i] == pytest.approx(v, rel=1e-12)
    for (i,), v in _golden(visco_fx, "coeff_lame", "coeffs1"):
        assert co[i] == pytest.approx(v, rel=1e-11)
    le, me, src = orc.src_a(3, 2, 10, 200.0, 300.0, 20.8e9, 10.4e9)
    checked = np.zeros((3, 36), bool)
    for (m, i), v in _golden(visco_fx, "src3d", "srcElastic"):
        assert src[m].ravel()[i] == pytest.approx(v, rel=1e-11, abs=1e-3)
        checked[m, i] = True
    assert checked.sum() >= 36  # every non-zero of the three mechanisms and the zeros the test lists


def test_visco_star_matches_reference_test(visco_fx):
    ji = np.array([[4, 8, 0], [2, 8, 4], [8, 0, 4]], dtype=float)
    st = orc.star_a(ji)  # [3][6][3]
    gold = _golden(visco_fx, "star3d", "starA")
    assert len(gold) == 9
    for (d, i), v in gold:
        assert st[d].ravel()[i] == pytest.approx(v, rel=1e-15)
    # the other directions follow the same pattern with column d of J^-1
    for d in range(3):
        np.testing.assert_allclose(st[d], orc.star_a(ji[:, [d, d, d]] * 1.0)[0], rtol=0, atol=0)


def test_visco_flux_mid_reduces_to_reference_2d(visco_fx):
    """The 3D anelastic middle states reduce to the 2D ones the reference
    tests (P: sigma_nn / u_n, S: sigma_ns / u_s and sigma_nt / u_t)."""
    midL, midR = orc.fs_mid_a(1.2e3, 1.1e3, 20.8e9, 19.4e9, 10.4e9, 13.1e9)
    g = {(n, i): v for n, i, v in ((n, tuple(i), v) for n, i, v in visco_fx["fs_mid2d"])}
    pos = {(0, 0): (0, 0), (0, 3): (0, 6), (2, 2): (3, 3), (2, 4): (3, 7)}
    for side, mid in (("fsMidL1", midL), ("fsMidR1", midR)):
        expect = np.zeros((6, 9))
        for (r2, c2), (r3, c3) in pos.items():
            expect[r3, c3] = g[(side, (r2, c2))]
        expect[5, 5], expect[5, 8] = expect[3, 3], expect[3, 7]  # second shear direction
        np.testing.assert_allclose(mid, expect, rtol=1e-12, atol=0)


def test_visco_flux_solvers_rotate_the_middle_state():
    """In the global frame (n, s, t) = (x, y, z) the rotation is the identity:
    the solvers are the middle states; a free surface negates the right
    solver's traction columns (0, 3, 5).  Rotated frames: T_A mid T^-1."""
    args = (1.2, 1.5, 2.0, 1.7, 1.1, 0.9)
    e = np.eye(3)
    l, r = orc.fs_a(*args, e[0], e[1], e[2])
    ml, mr = orc.fs_mid_a(*args)
    np.testing.assert_allclose(l, ml, atol=1e-15)
    np.testing.assert_allclose(r, mr, atol=1e-15)
    _, rf = orc.fs_a(*args, e[0], e[1], e[2], True)
    np.testing.assert_allclose(rf, mr * np.array([-1, 1, 1, -1, 1, -1, 1, 1, 1.0]), atol=1e-15)
    rng = np.random.default_rng(4)
    n, s, t = _frame(rng)
    T, Ti = orc.trafo(n, s, t)
    l, r = orc.fs_a(*args, n, s, t)
    np.testing.assert_allclose(l, T[:6, :6] @ ml @ Ti, rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(r, T[:6, :6] @ mr @ Ti, rtol=1e-13, atol=1e-14)


def test_plane_wave_polarised_per_element():
    """project(..., materials=...) gives each element the P eigenvector of its
    own material: the same DOFs as a homogeneous wave of that material, and
    the eigenvector of the normal Jacobian A_n for c_p."""
    m = M.box(3, jitter=0.1, seed=4)
    rho, lam, mu = S.random_materials(m.n_el, seed=4)
    pw = S.PlaneWave(k=(2 * np.pi, 0.0, 2 * np.pi))
    q = S.project(m, 3, pw, materials=(rho, lam, mu))
    for e in (0, 7, m.n_el - 1):
        qe = S.project(m, 3, S.PlaneWave(rho[e], lam[e], mu[e], k=pw.k))[e]
        np.testing.assert_allclose(q[e], qe, rtol=1e-14, atol=1e-14)
    r = pw.polarisations(rho, lam, mu)
    n = pw.k / np.linalg.norm(pw.k)
    for e in (0, 5):
        cp = np.sqrt((lam[e] + 2 * mu[e]) / rho[e])
        a = sum(n[d] * orc.star(rho[e], lam[e], mu[e], np.eye(3))[d].reshape(9, 9) for d in range(3))
        w = np.linalg.eigvals(a)
        assert np.min(np.abs(np.abs(w) - cp)) < 1e-9
        # r_P is an eigenvector of A_n for an eigenvalue of magnitude c_p
        lam_r = (a @ r[e]) @ r[e] / (r[e] @ r[e])
        np.testing.assert_allclose(a @ r[e], lam_r * r[e], rtol=1e-12, atol=1e-10)
        assert abs(abs(lam_r) - cp) < 1e-9
