"""Dynamic rupture on the GPU against the FP64 oracle, through the C ABI:
the face solvers built on the device (fault-frame orientation, jump and flux
solvers, scaling) and several time steps of net updates with linear
slip-weakening friction (state, failure flags, both face orientations,
R = 1 and 16 fused runs), plus the reference's 3D internal-boundary test.

Bar: FP64, relative 1e-12 per quantity.
"""
import numpy as np
import pytest

import oracle as orc
from edge_amd._lib import EdgeError
from edge_amd.rupture import Rupture

pytestmark = pytest.mark.gpu

SCALE = np.array([1e6] * 6 + [1.0] * 3)  # stresses, velocities


def _frame(rng):
    n = rng.normal(size=3)
    n /= np.linalg.norm(n)
    s = np.cross(n, rng.normal(size=3))
    s /= np.linalg.norm(s)
    return n, s, np.cross(n, s)


def _faces(rng, nf, n_sf, R):
    bnd = np.array(_frame(rng))
    nrm, tan = np.zeros((nf, 3)), np.zeros((nf, 2, 3))
    for f in range(nf):
        n, s, t = _frame(rng)
        if abs(n @ bnd[0]) < 0.05:  # keep away from faces across the fault plane
            n = n + 0.2 * np.sign(n @ bnd[0] + 1e-3) * bnd[0]
            n /= np.linalg.norm(n)
            s = np.cross(n, t)
            s /= np.linalg.norm(s)
            t = np.cross(n, s)
        nrm[f], tan[f, 0], tan[f, 1] = n, s, t
    mats = []
    for _ in range(2):
        rho = rng.uniform(2000, 3000, nf)
        vp = rng.uniform(4000, 6500, nf)
        vs = vp / rng.uniform(1.6, 2.0, nf)
        mu = rho * vs * vs
        mats.append(np.stack([rho, rho * vp * vp - 2 * mu, mu], axis=1))
    glob = np.stack([rng.uniform(0.6, 0.7, R), rng.uniform(0.45, 0.55, R), rng.uniform(1, 5, R)])
    ns = nf * n_sf * R
    st = orc.lsw_state(ns, sn0=-rng.uniform(50e6, 120e6, ns), co0=rng.uniform(0, 1e6, ns),
                       ss0_0=rng.uniform(10e6, 60e6, ns), ss0_1=rng.uniform(-10e6, 10e6, ns),
                       muf=np.repeat(glob[0][None, :], nf * n_sf, axis=0).ravel(), dd_0=rng.uniform(0, 0.2, ns))
    st[:, 4] = np.hypot(st[:, 2], st[:, 3])
    return dict(normals=nrm, tangents=tan, areas=rng.uniform(0.5, 2, nf), fault_frame=bnd, mat_l=mats[0],
                mat_r=mats[1], vol_l=rng.uniform(0.5, 2, nf), vol_r=rng.uniform(0.5, 2, nf), glob=glob,
                state0=st.reshape(nf, n_sf, R, 14))


def _oracle_solvers(p, n_sf, n_sc):
    out, pref = [], []
    for f in range(p["areas"].shape[0]):
        s, pr = orc.ib_solvers(p["normals"][f], p["tangents"][f, 0], p["tangents"][f, 1], p["fault_frame"],
                               p["mat_l"][f], p["mat_r"][f], p["areas"][f], p["vol_l"][f], p["vol_r"][f], n_sf, n_sc)
        out.append(s)
        pref.append(pr)
    return np.array(out), np.array(pref)


def _close(a, b, tol=1e-12):
    a, b = np.asarray(a), np.asarray(b)
    assert np.abs(a - b).max() <= tol * max(np.abs(b).max(), 1e-300), np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("n_sf,R", [(1, 1), (9, 1), (9, 16)])
def test_net_updates_match_oracle(gpu, n_sf, R):
    import torch
    rng = np.random.default_rng(100 + n_sf + R)
    nf, n_sc, dt = 37, 16, 2e-3
    p = _faces(rng, nf, n_sf, R)
    rup = Rupture(n_sf, n_sc, R, **p)
    solv, pref = _oracle_solvers(p, n_sf, n_sc)
    assert pref.any() and not pref.all()  # both orientations occur
    g_solv, g_st, g_per = rup.get()
    for k in range(4):
        _close(g_solv[:, k], solv[:, k], 1e-12)
    np.testing.assert_array_equal(g_st, p["state0"])
    state = p["state0"].reshape(nf, n_sf * R, 14).copy()
    cd = lambda m: np.sqrt(1.0 / (m[:, 0] * m[:, 2]))
    cdl, cdr = cd(p["mat_l"]), cd(p["mat_r"])
    net_l, net_r = rup.alloc(), rup.alloc()
    n_fail = 0
    for step in range(3):
        dl = rng.normal(size=(nf, 9, n_sf, R)) * SCALE[None, :, None, None]
        dr = rng.normal(size=(nf, 9, n_sf, R)) * SCALE[None, :, None, None]
        # faces [1, nf-1) only on the first step: the others keep their state
        first, n = (1, nf - 2) if step == 0 else (0, nf)
        rup.net_updates(torch.tensor(dl, device="cuda"), torch.tensor(dr, device="cuda"), net_l, net_r, dt, first, n)
        exp_l, exp_r = np.zeros_like(dl), np.zeros_like(dl)
        exp_per = np.zeros((nf, n_sf * R), dtype=bool)
        for f in range(first, first + n):
            l_eq_m = not pref[f]
            cm, cp = (cdl[f], cdr[f]) if l_eq_m else (cdr[f], cdl[f])
            exp_l[f], exp_r[f], pe = orc.net_updates(solv[f], dl[f], dr[f], dt, l_eq_m, cm, cp, p["glob"], state[f])
            exp_per[f] = pe.ravel()
        got_l, got_r = net_l.cpu().numpy(), net_r.cpu().numpy()
        for q in range(9):  # per quantity: stresses and velocities differ by orders of magnitude
            _close(got_l[first:first + n, q], exp_l[first:first + n, q])
            _close(got_r[first:first + n, q], exp_r[first:first + n, q])
        _, g_st, g_per = rup.get()
        g_st = g_st.reshape(nf, n_sf * R, 14)
        for c in range(14):
            _close(g_st[:, :, c], state[:, :, c])
        np.testing.assert_array_equal(g_per.reshape(nf, -1)[first:first + n], exp_per[first:first + n])
        n_fail += int(exp_per.sum())
    assert n_fail > 0


def test_reference_3d_internal_boundary(gpu, rupture_fx):
    """The reference's 3D internal-boundary test (InternalBoundary.test.cpp:339-628):
    one face with the identity frame, the left element the plus side."""
    import torch
    lL, mL, rL = 32080749540.0, 32019625230.0, 2670.0
    lR, mR, rR = lL / 2, mL / 2, 2600.0
    e = np.eye(3)
    ss0 = (40e6, np.sqrt(81.6e6 ** 2 - 40e6 ** 2))
    st = orc.lsw_state(1, sn0=-120e6, ss0_0=ss0[0], ss0_1=ss0[1], ss0A=81.6e6, muf=0.59, dd_0=0.2, dd_1=-0.1)
    rup = Rupture(1, 1, 1, normals=e[:1], tangents=e[None, 1:], areas=[1.0], fault_frame=e, mat_l=[[rL, lL, mL]],
                  mat_r=[[rR, lR, mR]], vol_l=[1.0], vol_r=[1.0], glob=[[0.677], [0.55], [2.5]],
                  state0=st.reshape(1, 1, 1, 14))
    solv, _, _ = rup.get()
    dl = torch.tensor(np.array(rupture_fx["qL"]).reshape(1, 9, 1, 1), device="cuda")
    dr = torch.tensor(np.array(rupture_fx["qR"]).reshape(1, 9, 1, 1), device="cuda")
    net_l, net_r = rup.alloc(), rup.alloc()
    rup.net_updates(dl, dr, net_l, net_r, 1.0)
    _, _, per = rup.get()
    assert per.all()
    exp_l = -(solv[0, 2] @ np.array(rupture_fx["msLp"]))
    exp_r = solv[0, 3] @ np.array(rupture_fx["msRp"])
    np.testing.assert_allclose(net_l.cpu().numpy().ravel(), exp_l, rtol=1e-12, atol=1e-12 * np.abs(exp_l).max())
    np.testing.assert_allclose(net_r.cpu().numpy().ravel(), exp_r, rtol=1e-12, atol=1e-12 * np.abs(exp_r).max())


def test_rejects_bad_ranges(gpu):
    import torch
    rng = np.random.default_rng(1)
    rup = Rupture(1, 1, 1, **_faces(rng, 4, 1, 1))
    a = rup.alloc()
    with pytest.raises(EdgeError, match="out of bounds"):
        rup.net_updates(a, a, a, a, 1e-3, first=2, n=3)
    with pytest.raises(ValueError):
        rup.net_updates(a, a, a, torch.zeros(3, device="cuda", dtype=torch.float64), 1e-3)
