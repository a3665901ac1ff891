"""Rate-2 local time stepping on the GPU against the FP64 oracle, through the
C ABI: two and three time groups (contiguous element ranges, twice the step
per group), the tDofs[1] / tDofs[2] buffers and the interface faces, with the
generic kernels (FP64, 1 run) and the MFMA-tiled kernels with face
projections (FP32, 16 runs).  Bars: FP64 relative L2 <= 1e-12, FP32 <= 1e-5."""
import numpy as np
import pytest

import oracle as orc
from edge_amd import mesh as M
from edge_amd import setups as S
from edge_amd.lts import Rate2Lts
from edge_amd.mesh import FA_FREE_SURFACE, FA_OUTFLOW, FA_PERIODIC
from edge_amd.solver import (KERNEL_AUTO, KERNEL_EC64, KERNEL_GENERIC, KERNEL_MFMA, KERNEL_MFMA64, KERNEL_MFMA_TILES,
                             PATH_PROJECTIONS, ElasticSolver)

pytestmark = pytest.mark.gpu


def _run_oracle(o, dofs, ranges, dt0, cycles):
    bufs = [np.zeros_like(dofs) for _ in range(3)]
    counts = [0] * len(ranges)

    def step(g):
        ft = counts[g] % 2 == 0
        first, n = ranges[g]
        o.local_lts(dofs, *bufs, dt0 * 2 ** g, ft, first, n)
        if g > 0:
            step(g - 1)
            step(g - 1)
        o.neigh_lts(dofs, *bufs, ft, first, n)
        counts[g] += 1

    for _ in range(cycles):
        step(len(ranges) - 1)


@pytest.mark.parametrize("prec,order,runs,n_groups,kernel", [(64, 4, 1, 2, KERNEL_AUTO), (64, 3, 1, 3, KERNEL_AUTO),
                                                             (32, 5, 16, 2, KERNEL_AUTO), (32, 4, 16, 3, KERNEL_AUTO),
                                                             (64, 5, 1, 2, KERNEL_EC64), (64, 4, 1, 3, KERNEL_EC64),
                                                             (32, 5, 1, 2, KERNEL_AUTO), (32, 3, 1, 3, KERNEL_EC64)])
def test_lts_matches_oracle(gpu, prec, order, runs, n_groups, kernel):
    import torch
    m = M.box(4, jitter=0.1, seed=17)
    rho, lam, mu = S.random_materials(m.n_el, seed=17)
    s = ElasticSolver(m, order, runs, prec, rho, lam, mu, kernel=kernel)
    if prec == 32:
        assert s.kernel_info()[1] & PATH_PROJECTIONS
    cut = np.linspace(0, m.n_el, n_groups + 1).astype(int)
    ranges = [(cut[g], cut[g + 1] - cut[g]) for g in range(n_groups)]
    dt0 = S.dt_cfl(m, rho, lam, mu, order) / 2 ** (n_groups - 1)
    lts = Rate2Lts(s, ranges, dt0)
    assert (lts.el_lts & 1).any() and (lts.el_lts & 2).any()
    q0 = S.project(m, order, S.PlaneWave(k=(2 * np.pi, 2 * np.pi, 0.0)), n_runs=runs)
    q0 *= np.linspace(0.5, 1.0, runs)
    d = torch.tensor(q0, dtype=s.dtype, device="cuda")
    lts.cycle(d, 2)
    star, fs = s.setup()
    o = orc.MeshOracle(m, order, runs, star=star, fs=fs)
    o.set_lts(lts.el_lts, lts.fa_lts)
    ref = q0.copy()
    _run_oracle(o, ref, ranges, dt0, 2)
    got = d.double().cpu().numpy()
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err <= (1e-12 if prec == 64 else 1e-5), err
    # the interface mattered: GTS at the finest step differs
    gts = q0.copy()
    o.step(gts, np.zeros_like(gts), dt0, 2 ** (n_groups - 1) * 2)
    assert np.linalg.norm(gts - ref) / np.linalg.norm(ref) > 10 * err
    s.close()


@pytest.mark.parametrize("prec,runs,kernel,tol", [(64, 1, KERNEL_MFMA64, 1e-12), (64, 1, KERNEL_GENERIC, 1e-12),
                                                  (32, 8, KERNEL_MFMA, 1e-5), (32, 8, KERNEL_MFMA_TILES, 1e-5)])
def test_visco_rate2_matches_oracle(gpu, prec, runs, kernel, tol):
    """Viscoelastic solvers with three time groups (AderDg with TL_N_RMS > 0
    and local time stepping): tDofs[2] from the CK with the anelastic terms,
    interface faces with the anelastic update."""
    import torch
    from edge_amd.solver import ViscoSolver
    order, nr = 4, 3
    m = M.box(4, bc=(FA_PERIODIC, FA_FREE_SURFACE, FA_OUTFLOW), jitter=0.1, seed=12)
    rho, lam, mu = S.random_materials(m.n_el, seed=12)
    qp = np.random.default_rng(12).uniform(20, 60, m.n_el)
    s = ViscoSolver(m, order, runs, prec, rho, lam, mu, qp, qp / 2, nr, 2.0, 10.0, kernel=kernel)
    assert s.kernel_info()[0] == (KERNEL_MFMA if kernel == KERNEL_MFMA_TILES else kernel)
    visco = orc.setup_mesh_visco(m, rho, lam, mu, qp, qp / 2, nr, 2.0, 10.0)
    mo = orc.MeshOracle(m, order, runs, visco=visco)
    dt0 = S.dt_cfl(m, rho, visco["lam_e"], visco["mu_e"], order) / 4
    n3 = m.n_el // 3
    ranges = [(0, n3), (n3, n3), (2 * n3, m.n_el - 2 * n3)]
    lts = Rate2Lts(s, ranges, dt0)
    mo.set_lts(lts.el_lts, lts.fa_lts)
    rng = np.random.default_rng(4)
    n = S.n_modes(order)
    d0 = rng.normal(size=(m.n_el, 9, n, runs))
    a0 = 0.1 * rng.normal(size=(m.n_el, nr, 6, n, runs))
    d, a = d0.copy(), a0.copy()
    bufs = [np.zeros_like(d0) for _ in range(3)]
    counts = [0] * 3

    def step(g):
        ft = counts[g] % 2 == 0
        first, cnt = ranges[g]
        mo.local_lts(d, *bufs, dt0 * 2 ** g, ft, first, cnt, dofs_a=a)
        if g > 0:
            step(g - 1)
            step(g - 1)
        mo.neigh_lts(d, *bufs, ft, first, cnt, dofs_a=a)
        counts[g] += 1

    step(2)
    dt_t = torch.float64 if prec == 64 else torch.float32
    gd = torch.tensor(d0, dtype=dt_t, device="cuda")
    ga = torch.tensor(a0, dtype=dt_t, device="cuda")
    lts.cycle(gd, 1, dofs_a=ga)
    for got, ref in ((gd, d), (ga, a)):
        g = got.cpu().double().numpy()
        assert np.linalg.norm(g - ref) / np.linalg.norm(ref) <= tol
    s.close()
