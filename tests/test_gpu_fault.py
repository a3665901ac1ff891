"""The fault phase of the coupled step on the GPU against the FP64 oracle's
coupled step (local, fault, neighbour; tests/test_oracle_fault.py pins the
oracle), through the C ABI: FP64 single run (FP64 MFMA and generic kernels)
at relative L2 <= 1e-12, FP32 16 fused runs (MFMA kernels) at <= 1e-5, with a
nucleation patch so that sub-faces slip; the friction state alike; a locked
fault equals the welded step; the graph-captured step equals the eager one."""
import numpy as np
import pytest

import oracle as orc
from edge_amd import _lib
from edge_amd import fault as FL
from edge_amd import mesh as M
from edge_amd import setups as S
from edge_amd._lib import EdgeError
from edge_amd.mesh import FA_OUTFLOW, FA_PERIODIC
from edge_amd.solver import KERNEL_EC64, KERNEL_GENERIC, KERNEL_MFMA, KERNEL_MFMA64, KERNEL_MFMA_TILES, ElasticSolver

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def _dev(a, prec):
    import torch
    return torch.tensor(a, dtype=torch.float64 if prec == 64 else torch.float32, device="cuda").contiguous()


def _problem(order, R, mus=0.677, seed=3, n=4):
    m = M.box(n, bc=(FA_PERIODIC, FA_PERIODIC, FA_OUTFLOW), seed=seed)
    faces = FL.tag_plane(m, 2, 0.5)
    rho, lam, mu = S.random_materials(m.n_el, seed=seed)
    frame = np.array([[0, 0, 1.0], [1, 0, 0], [0, 1, 0]])
    glob = np.array([[mus] * R, [0.525] * R, [1 / 0.4] * R])
    rng = np.random.default_rng(seed)
    nsf = FL.n_sub_faces(order)
    ss0 = np.broadcast_to(rng.uniform(70, 81.24, R)[None, None, :], (len(faces), nsf, R)).copy()
    pts = FL.sub_face_points(m, order)
    inside = (np.abs(pts[..., 0] - 0.5) < 0.3) & (np.abs(pts[..., 1] - 0.5) < 0.3)
    ss0[inside] = 81.6
    st0 = FL.lsw_state(len(faces), nsf, R, sn0=-120.0, ss0_0=ss0, ss0_1=0.0, mus=glob[0])
    return m, rho, lam, mu, frame, glob, st0


CASES = [(64, 1, KERNEL_MFMA64, 1e-12, 4), (64, 1, KERNEL_GENERIC, 1e-12, 4), (64, 1, KERNEL_EC64, 1e-12, 4),
         (32, 16, KERNEL_MFMA, 1e-5, 4), (32, 16, KERNEL_MFMA_TILES, 1e-5, 4), (32, 1, KERNEL_EC64, 1e-5, 4),
         # the MFMA fault phase of FP32 16-run solvers at the other orders
         # (order 5: 133 KB of LDS per workgroup)
         (32, 16, KERNEL_MFMA, 1e-5, 2), (32, 16, KERNEL_MFMA, 1e-5, 3), (32, 16, KERNEL_MFMA, 1e-5, 5)]


@pytest.mark.parametrize("prec,R,kernel,tol,order", CASES)
def test_coupled_step_matches_oracle(gpu, prec, R, kernel, tol, order):
    n_steps = 3
    m, rho, lam, mu, frame, glob, st0 = _problem(order, R)
    dt = S.dt_cfl(m, rho, lam, mu, order)
    s = ElasticSolver(m, order, R, prec, rho, lam, mu, kernel=kernel)
    f = FL.Fault(s, m, rho, lam, mu, frame, glob, st0)
    fo = orc.FaultOracle(m, order, R, rho, lam, mu, frame, glob, st0)
    d0 = 0.01 * np.random.default_rng(5).normal(size=(m.n_el, 9, S.n_modes(order), R))
    d, t = d0.copy(), np.zeros_like(d0)
    fo.step(d, t, dt, n_steps)
    gd, gt = _dev(d0, prec), _dev(np.zeros_like(d0), prec)
    s.step(gd, gt, dt, n_steps)
    assert rel(gd.cpu().numpy(), d) <= tol
    st, per = f.get()
    ost = fo.state.reshape(st.shape)
    np.testing.assert_array_equal(per, fo.per)
    assert fo.per.any() and not fo.per.all()
    for c in (5, 6, 7, 8, 9, 10, 11, 12):  # friction coefficient, strength, traction, slip rate, slip
        scale = np.abs(ost[..., c]).max()
        assert np.abs(st[..., c] - ost[..., c]).max() <= max(tol, 1e-12) * 10 * scale, c
    f.close()
    s.close()


def test_locked_fault_equals_welded_step(gpu):
    order, R = 4, 16
    m, rho, lam, mu, frame, glob, st0 = _problem(order, R, mus=1e6)
    dt = S.dt_cfl(m, rho, lam, mu, order)
    # small enough that the normal stress stays compressive (-120 + O(1)): a
    # tensile sub-face has zero strength and opens however large mu_s is
    d0 = 0.02 * np.random.default_rng(2).normal(size=(m.n_el, 9, S.n_modes(order), R))
    s = ElasticSolver(m, order, R, 64, rho, lam, mu, kernel=KERNEL_GENERIC)
    f = FL.Fault(s, m, rho, lam, mu, frame, glob, st0)
    gd, gt = _dev(d0, 64), _dev(np.zeros_like(d0), 64)
    s.step(gd, gt, dt, 2)
    welded = M.box(4, bc=(FA_PERIODIC, FA_PERIODIC, FA_OUTFLOW), seed=3)
    w = ElasticSolver(welded, order, R, 64, rho, lam, mu, kernel=KERNEL_GENERIC)
    wd, wt = _dev(d0, 64), _dev(np.zeros_like(d0), 64)
    w.step(wd, wt, dt, 2)
    ref = wd.cpu().numpy()
    assert rel(gd.cpu().numpy() - d0, ref - d0) <= 1e-13
    assert not f.get()[1].any()
    f.close()
    s.close()
    w.close()


def test_graph_step_equals_eager_and_checks_ownership(gpu):
    import torch
    order, R = 4, 16
    m, rho, lam, mu, frame, glob, st0 = _problem(order, R)
    dt = S.dt_cfl(m, rho, lam, mu, order)
    d0 = 0.01 * np.random.default_rng(4).normal(size=(m.n_el, 9, S.n_modes(order), R))
    outs = []
    for graph in (False, True):
        s = ElasticSolver(m, order, R, 32, rho, lam, mu)
        f = FL.Fault(s, m, rho, lam, mu, frame, glob, st0)
        gd, gt = _dev(d0, 32), _dev(np.zeros_like(d0), 32)
        for _ in range(3):  # replays with eager host work in between
            s.step(gd, gt, dt, 1, graph=graph)
            torch.cuda.synchronize()
        outs.append((gd.cpu().numpy(), f.get()[0]))
        other = ElasticSolver(m, order, R, 32, rho, lam, mu)
        with pytest.raises(EdgeError):  # a fault belongs to the solver it was built for
            _lib.check(_lib.lib().edge_el_set_fault(other._h, f._h))
        other.close()
        f.close()
        s.close()
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])
