"""Viscoelastic solvers (N_RMS relaxation mechanisms) on the GPU against the
FP64 oracle, through the C ABI: the set-up (constant-Q fit, unrelaxed Lame
parameters, anelastic source / star matrices and flux solvers, elastic
star matrices and flux solvers from the unrelaxed parameters) and the
local / neighbour steps with anelastic DOFs.

Bars: FP64 relative L2 <= 1e-12, FP32 (8 fused runs) <= 1e-5.
"""
import numpy as np
import pytest

import oracle as orc
from edge_amd import mesh as M
from edge_amd import setups as S
from edge_amd.mesh import FA_FREE_SURFACE, FA_OUTFLOW, FA_PERIODIC
from edge_amd.solver import (KERNEL_AUTO, KERNEL_GENERIC, KERNEL_MFMA, KERNEL_MFMA64, KERNEL_MFMA_TILES,
                             PATH_PROJECTIONS, PATH_VISCO, ViscoSolver)

pytestmark = pytest.mark.gpu

# a band where the relaxation frequencies times dt are O(0.1): the anelastic
# terms change the solution visibly within a few steps
F_CEN, F_RAT, NR = 20.0, 10.0, 3


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def _problem(order, n=3, bc=(FA_FREE_SURFACE, FA_OUTFLOW, FA_PERIODIC), seed=31):
    m = M.box(n, bc=bc, jitter=0.1, seed=seed)
    rho, lam, mu = S.random_materials(m.n_el, seed=seed)
    rng = np.random.default_rng(seed)
    qp = rng.uniform(20, 60, m.n_el)
    return m, rho, lam, mu, qp, qp / 2


def _dev(a, prec):
    import torch
    return torch.tensor(a, dtype=torch.float64 if prec == 64 else torch.float32, device="cuda").contiguous()


def test_visco_setup_matches_oracle(gpu):
    m, rho, lam, mu, qp, qs = _problem(4)
    s = ViscoSolver(m, 4, 1, 64, rho, lam, mu, qp, qs, NR, F_CEN, F_RAT)
    assert s.kernel_info()[1] & PATH_VISCO
    ref = orc.setup_mesh_visco(m, rho, lam, mu, qp, qs, NR, F_CEN, F_RAT)
    got = s.setup_a()
    for k in ("lam_e", "mu_e", "rfs", "srcA", "starA", "fsA"):
        assert rel(got[k], ref[k]) <= 1e-12, k
    star, fs = s.setup()
    assert rel(star, ref["starE"]) <= 1e-12
    assert rel(fs, ref["fsE"]) <= 1e-12


@pytest.mark.parametrize("prec,order,runs,kernel,nr", [
    (64, 2, 1, KERNEL_AUTO, NR), (64, 3, 1, KERNEL_AUTO, NR), (64, 4, 1, KERNEL_AUTO, NR),
    (64, 5, 1, KERNEL_AUTO, NR), (64, 4, 1, KERNEL_GENERIC, NR), (64, 5, 1, KERNEL_GENERIC, NR),
    (64, 5, 1, KERNEL_MFMA64, 1), (64, 3, 1, KERNEL_MFMA64, 2),
    (32, 4, 8, KERNEL_GENERIC, NR), (32, 5, 8, KERNEL_GENERIC, NR),
    (32, 3, 8, KERNEL_MFMA, NR), (32, 4, 8, KERNEL_MFMA, NR), (32, 5, 8, KERNEL_AUTO, NR),
    (32, 4, 8, KERNEL_MFMA, 1), (32, 5, 8, KERNEL_MFMA, 2), (32, 5, 8, KERNEL_MFMA_TILES, NR)])
def test_visco_steps_match_oracle(gpu, prec, order, runs, kernel, nr):
    m, rho, lam, mu, qp, qs = _problem(order, seed=40 + order)
    s = ViscoSolver(m, order, runs, prec, rho, lam, mu, qp, qs, nr, F_CEN, F_RAT, kernel=kernel)
    if prec == 64 and kernel != KERNEL_GENERIC:  # auto picks the FP64 MFMA kernels for one run
        assert s.kernel_info()[0] == KERNEL_MFMA64
    if prec == 32 and kernel != KERNEL_GENERIC:  # auto picks the MFMA kernels for FP32 with 8 runs
        k, flags = s.kernel_info()
        assert k == KERNEL_MFMA
        # face projections exchanged unless the neighbour step is asked to read tInt
        assert bool(flags & PATH_PROJECTIONS) == (kernel != KERNEL_MFMA_TILES)
    visco = orc.setup_mesh_visco(m, rho, lam, mu, qp, qs, nr, F_CEN, F_RAT)
    mo = orc.MeshOracle(m, order, runs, visco=visco)
    rng = np.random.default_rng(order)
    n = S.n_modes(order)
    d0 = rng.normal(size=(m.n_el, 9, n, runs))
    a0 = 0.1 * rng.normal(size=(m.n_el, nr, 6, n, runs))
    dt = S.dt_cfl(m, rho, visco["lam_e"], visco["mu_e"], order)
    tol = 1e-12 if prec == 64 else 1e-5
    d, t, a = d0.copy(), np.zeros_like(d0), a0.copy()
    gd, gt, ga = _dev(d0, prec), _dev(np.zeros_like(d0), prec), _dev(a0, prec)
    mo.local(d, t, dt, dofs_a=a)
    s.local(gd, gt, dt, ga)
    assert rel(gt.cpu().numpy(), t) <= tol
    assert rel(gd.cpu().numpy(), d) <= tol
    assert rel(ga.cpu().numpy(), a) <= tol
    mo.neigh(d, t, dofs_a=a)
    s.neigh(gd, gt, ga)
    assert rel(gd.cpu().numpy(), d) <= tol
    assert rel(ga.cpu().numpy(), a) <= tol
    mo.step(d, t, dt, 4, dofs_a=a)
    s.step(gd, gt, dt, 4, ga)
    got, got_a = gd.cpu().numpy(), ga.cpu().numpy()
    for r in range(runs):
        assert rel(got[..., r], d[..., r]) <= tol
        assert rel(got_a[..., r], a[..., r]) <= tol
    # the anelastic terms matter here: the elastic update of the same DOFs differs
    assert rel(d, orc_elastic_step(m, order, runs, rho, visco, d0, dt)) > 1e-3


def orc_elastic_step(m, order, runs, rho, visco, d0, dt):
    """Five elastic steps (unrelaxed Lame parameters, no memory variables)."""
    mo = orc.MeshOracle(m, order, runs, star=visco["starE"], fs=visco["fsE"])
    d, t = d0.copy(), np.zeros_like(d0)
    mo.step(d, t, dt, 5)
    return d


def test_elastic_generic_eight_runs_match_oracle(gpu):
    """The generic elastic kernels with 8 fused runs (FP64)."""
    from edge_amd.solver import KERNEL_GENERIC, ElasticSolver
    m, rho, lam, mu, _, _ = _problem(4, seed=9)
    s = ElasticSolver(m, 4, 8, 64, rho, lam, mu, kernel=KERNEL_GENERIC)
    mo = orc.MeshOracle(m, 4, 8, rho=rho, lam=lam, mu=mu)
    rng = np.random.default_rng(2)
    d0 = rng.normal(size=(m.n_el, 9, 20, 8))
    dt = S.dt_cfl(m, rho, lam, mu, 4)
    d, t = d0.copy(), np.zeros_like(d0)
    mo.step(d, t, dt, 3)
    gd, gt = _dev(d0, 64), _dev(np.zeros_like(d0), 64)
    s.step(gd, gt, dt, 3)
    assert rel(gd.cpu().numpy(), d) <= 1e-12
