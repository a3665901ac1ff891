"""The autotuner: it times every kernel variant a solver can switch to on
its own mesh, keeps the fastest, and the solver still meets the parity bars
against the FP64 oracle afterwards (FP64 relative L2 <= 1e-12, FP32 <= 1e-5).
Each candidate is a kernel variant the parity tests check on its own through
the C ABI's kernel selector (test_gpu_parity.py, test_gpu_visco.py)."""
import numpy as np
import pytest

import oracle as orc
from edge_amd import mesh as M
from edge_amd import setups as S
from edge_amd.mesh import FA_FREE_SURFACE, FA_OUTFLOW, FA_PERIODIC
from edge_amd.solver import (KERNEL_GENERIC, KERNEL_MFMA, KERNEL_MFMA64, PATH_FUSED, PATH_HI_OCC,
                             PATH_PROJECTIONS, ElasticSolver, ViscoSolver)

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def _dev(a, prec):
    import torch
    return torch.tensor(a, dtype=torch.float64 if prec == 64 else torch.float32, device="cuda").contiguous()


@pytest.mark.parametrize("prec,runs,expect", [(32, 16, {(KERNEL_MFMA, True), (KERNEL_MFMA, False)}),
                                              (64, 1, {(KERNEL_MFMA64, False), (KERNEL_GENERIC, False)})])
def test_autotune_elastic(gpu, prec, runs, expect):
    order = 4
    m = M.box(4, bc=(FA_FREE_SURFACE, FA_OUTFLOW, FA_PERIODIC), jitter=0.1, seed=4)
    rho, lam, mu = S.random_materials(m.n_el, seed=4)
    s = ElasticSolver(m, order, runs, prec, rho, lam, mu)
    dt = S.dt_cfl(m, rho, lam, mu, order)
    cands = s.autotune(dt, n_reps=2)
    assert expect <= {(k, p) for k, p, *_ in cands}
    assert all(c[3] > 0 for c in cands)
    k, p, _, _, f, _ = min(cands, key=lambda c: c[3])
    kernel, flags = s.kernel_info()
    assert bool(flags & PATH_FUSED) == f
    assert kernel == k and bool(flags & PATH_PROJECTIONS) == p
    rng = np.random.default_rng(1)
    d0 = rng.normal(size=(m.n_el, 9, S.n_modes(order), runs))
    mo = orc.MeshOracle(m, order, runs, rho=rho, lam=lam, mu=mu)
    d, t = d0.copy(), np.zeros_like(d0)
    mo.step(d, t, dt, 2)
    gd, gt = _dev(d0, prec), _dev(np.zeros_like(d0), prec)
    s.step(gd, gt, dt, 2)
    assert rel(gd.cpu().numpy(), d) <= (1e-12 if prec == 64 else 1e-5)
    s.close()


def test_autotune_visco(gpu):
    order, runs, nr = 4, 8, 3
    m = M.box(3, bc=(FA_FREE_SURFACE, FA_OUTFLOW, FA_PERIODIC), jitter=0.1, seed=6)
    rho, lam, mu = S.random_materials(m.n_el, seed=6)
    qp = np.random.default_rng(6).uniform(20, 60, m.n_el)
    s = ViscoSolver(m, order, runs, 32, rho, lam, mu, qp, qp / 2, nr, 20.0, 10.0)
    visco = orc.setup_mesh_visco(m, rho, lam, mu, qp, qp / 2, nr, 20.0, 10.0)
    dt = S.dt_cfl(m, rho, visco["lam_e"], visco["mu_e"], order)
    cands = s.autotune(dt, n_reps=2)
    assert {(KERNEL_MFMA, True), (KERNEL_MFMA, False), (KERNEL_GENERIC, False)} <= {(k, p) for k, p, *_ in cands}
    assert {x for k, _, x, *_ in cands if k == KERNEL_MFMA} == {True, False}  # both element schedules
    assert {c[5] for c in cands if c[0] == KERNEL_MFMA} == {True, False}  # both occupancy variants
    k, *_, o = min(cands, key=lambda c: c[3])
    assert s.kernel_info()[0] == k and bool(s.kernel_info()[1] & PATH_HI_OCC) == o
    rng = np.random.default_rng(2)
    n = S.n_modes(order)
    d0 = rng.normal(size=(m.n_el, 9, n, runs))
    a0 = 0.1 * rng.normal(size=(m.n_el, nr, 6, n, runs))
    mo = orc.MeshOracle(m, order, runs, visco=visco)
    d, t, a = d0.copy(), np.zeros_like(d0), a0.copy()
    mo.step(d, t, dt, 2, dofs_a=a)
    gd, gt, ga = _dev(d0, 32), _dev(np.zeros_like(d0), 32), _dev(a0, 32)
    s.step(gd, gt, dt, 2, ga)
    assert rel(gd.cpu().numpy(), d) <= 1e-5
    assert rel(ga.cpu().numpy(), a) <= 1e-5
    s.close()


def _cases():
    return [("elastic", 32, 16), ("elastic", 64, 1), ("elastic", 32, 1), ("visco", 32, 8), ("visco", 64, 1)]


@pytest.mark.parametrize("kind,prec,runs", _cases())
def test_every_candidate_matches_oracle(gpu, kind, prec, runs):
    """Each variant the autotuner may keep, pinned through edge_el_set_variant
    and stepped against the oracle (FP64 1e-12, FP32 1e-5)."""
    order, nr = 4, 3
    m = M.box(3, bc=(FA_FREE_SURFACE, FA_OUTFLOW, FA_PERIODIC), jitter=0.1, seed=8)
    rho, lam, mu = S.random_materials(m.n_el, seed=8)
    n = S.n_modes(order)
    rng = np.random.default_rng(9)
    d0 = rng.normal(size=(m.n_el, 9, n, runs))
    a0 = 0.1 * rng.normal(size=(m.n_el, nr, 6, n, runs)) if kind == "visco" else None
    if kind == "visco":
        qp = rng.uniform(20, 60, m.n_el)
        s = ViscoSolver(m, order, runs, prec, rho, lam, mu, qp, qp / 2, nr, 20.0, 10.0)
        visco = orc.setup_mesh_visco(m, rho, lam, mu, qp, qp / 2, nr, 20.0, 10.0)
        dt = S.dt_cfl(m, rho, visco["lam_e"], visco["mu_e"], order)
        mo = orc.MeshOracle(m, order, runs, visco=visco)
    else:
        s = ElasticSolver(m, order, runs, prec, rho, lam, mu)
        dt = S.dt_cfl(m, rho, lam, mu, order)
        mo = orc.MeshOracle(m, order, runs, rho=rho, lam=lam, mu=mu)
    # three steps: a fused candidate runs local, two fused launches (both
    # projection buffers read and written) and the last neighbour step
    d, t = d0.copy(), np.zeros_like(d0)
    a = a0.copy() if a0 is not None else None
    mo.step(d, t, dt, 3, dofs_a=a) if a is not None else mo.step(d, t, dt, 3)
    cands = s.autotune(dt, n_reps=1)
    assert len(cands) >= 2
    if prec == 32 and runs in (16, 8):
        assert any(c[4] for c in cands)  # a fused candidate
    tol = 1e-12 if prec == 64 else 1e-5
    for k, p, x, _, f, o in cands:
        s.set_variant(k, p, x, f, o)
        assert s.kernel_info()[0] == k
        assert bool(s.kernel_info()[1] & PATH_FUSED) == f
        assert bool(s.kernel_info()[1] & PATH_HI_OCC) == o
        for graph in (False, True):
            gd, gt = _dev(d0, prec), _dev(np.zeros_like(d0), prec)
            if a is not None:
                ga = _dev(a0, prec)
                s.step(gd, gt, dt, 3, ga, graph=graph)
                assert rel(ga.cpu().numpy(), a) <= tol, (k, p, x, f, o, graph)
            else:
                s.step(gd, gt, dt, 3, graph=graph)
            assert rel(gd.cpu().numpy(), d) <= tol, (k, p, x, f, o, graph)
    with pytest.raises(Exception):
        s.set_variant(7)
    s.close()


@pytest.mark.parametrize("kind", ["elastic", "visco"])
def test_neigh_checks_the_tint_the_projections_stand_for(gpu, kind):
    """With face projections exchanged the neighbour step reads them instead
    of t_int, so it refuses a t_int other than the one the local step wrote
    (and a step after the autotuner overwrote them); the tInt-reading variant
    takes a substituted copy and gives the same DOFs."""
    import torch
    from edge_amd._lib import EdgeError
    order, runs, nr = 4, (16 if kind == "elastic" else 8), 3
    m = M.box(3, jitter=0.1, seed=5)
    rho, lam, mu = S.random_materials(m.n_el, seed=5)
    if kind == "visco":
        qp = np.full(m.n_el, 40.0)
        s = ViscoSolver(m, order, runs, 32, rho, lam, mu, qp, qp / 2, nr, 20.0, 10.0)
    else:
        s = ElasticSolver(m, order, runs, 32, rho, lam, mu)
    dt = 0.5 * S.dt_cfl(m, rho, lam, mu, order)
    s.set_variant(KERNEL_MFMA, projections=True)
    assert s.kernel_info()[1] & PATH_PROJECTIONS
    rng = np.random.default_rng(3)
    d0 = rng.normal(size=(m.n_el, 9, S.n_modes(order), runs))
    ga = _dev(np.zeros((m.n_el, nr, 6, S.n_modes(order), runs)), 32) if kind == "visco" else None

    def local(d, t):
        s.local(d, t, dt, dofs_a=ga) if ga is not None else s.local(d, t, dt)

    def neigh(d, t):
        s.neigh(d, t, dofs_a=ga) if ga is not None else s.neigh(d, t)

    gd, gt = _dev(d0, 32), _dev(np.zeros_like(d0), 32)
    with pytest.raises(EdgeError):  # no local step yet
        neigh(gd, gt)
    local(gd, gt)
    with pytest.raises(EdgeError):  # a substituted tInt
        neigh(gd, gt.clone())
    neigh(gd, gt)
    ref = gd.cpu().numpy()
    local(gd, gt)
    s.autotune(dt, n_reps=1)
    s.set_variant(KERNEL_MFMA, projections=True)
    with pytest.raises(EdgeError):  # the autotuner overwrote the projections
        neigh(gd, gt)
    # the tInt-reading variant takes the substituted copy
    s.set_variant(KERNEL_MFMA, projections=False)
    gd2, gt2 = _dev(d0, 32), _dev(np.zeros_like(d0), 32)
    if ga is not None:
        ga.zero_()
    local(gd2, gt2)
    neigh(gd2, gt2.clone())
    torch.cuda.synchronize()
    assert rel(gd2.cpu().numpy(), ref) <= 1e-6
    s.close()
