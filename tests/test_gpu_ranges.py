"""Element ranges (the reference's work regions, AderDg::local / neigh over
[first, first + n)): local over uneven ranges, then neigh over other uneven
ranges, equals the full-range step bit for bit for every kernel family; the
face projections one range's local step writes reach the neighbour steps of
the other ranges."""
import numpy as np
import pytest

from edge_amd import mesh as M
from edge_amd import setups as S
from edge_amd.mesh import FA_FREE_SURFACE, FA_OUTFLOW, FA_PERIODIC
from edge_amd.solver import (KERNEL_EC64, KERNEL_GENERIC, KERNEL_MFMA, KERNEL_MFMA64, KERNEL_MFMA_TILES,
                             ElasticSolver, ViscoSolver)

pytestmark = pytest.mark.gpu

CUTS_L = [0, 37, 100, 161]   # local ranges [a, b)
CUTS_N = [0, 80, 81, 161]    # neighbour ranges


def _mesh():
    m = M.box((3, 3, 3), bc=(FA_FREE_SURFACE, FA_OUTFLOW, FA_PERIODIC), jitter=0.1, seed=12)
    return m, *S.random_materials(m.n_el, seed=12)


def _ranges(cuts, n_el):
    c = list(cuts[:-1]) + [n_el]
    return [(a, b - a) for a, b in zip(c[:-1], c[1:]) if b > a]


@pytest.mark.parametrize("prec,runs,kernel", [(64, 1, KERNEL_MFMA64), (64, 1, KERNEL_GENERIC), (32, 16, KERNEL_MFMA),
                                              (32, 16, KERNEL_MFMA_TILES), (64, 1, KERNEL_EC64), (32, 1, KERNEL_EC64)])
def test_elastic_ranges_equal_full(gpu, prec, runs, kernel):
    import torch
    m, rho, lam, mu = _mesh()
    s = ElasticSolver(m, 4, runs, prec, rho, lam, mu, kernel=kernel)
    dt = S.dt_cfl(m, rho, lam, mu, 4)
    gen = torch.Generator(device="cuda").manual_seed(1)
    d0 = torch.randn(s.dof_shape(), dtype=s.dtype, device="cuda", generator=gen)
    full, part = d0.clone(), d0.clone()
    tf, tp = torch.zeros_like(d0), torch.zeros_like(d0)
    for _ in range(2):
        s.local(full, tf, dt)
        s.neigh(full, tf)
        for a, n in _ranges(CUTS_L, m.n_el):
            s.local(part, tp, dt, first=a, n=n)
        for a, n in _ranges(CUTS_N, m.n_el):
            s.neigh(part, tp, first=a, n=n)
    torch.cuda.synchronize()
    assert torch.equal(full, part)
    s.close()


@pytest.mark.parametrize("prec,runs,kernel", [(32, 8, KERNEL_MFMA), (32, 8, KERNEL_MFMA_TILES), (64, 1, KERNEL_MFMA64),
                                              (32, 8, KERNEL_GENERIC)])
def test_visco_ranges_equal_full(gpu, prec, runs, kernel):
    import torch
    m, rho, lam, mu = _mesh()
    qp = np.random.default_rng(3).uniform(20, 60, m.n_el)
    s = ViscoSolver(m, 4, runs, prec, rho, lam, mu, qp, qp / 2, 3, 20.0, 10.0, kernel=kernel)
    dt = 0.5 * S.dt_cfl(m, rho, lam, mu, 4)
    gen = torch.Generator(device="cuda").manual_seed(2)
    d0 = torch.randn(s.dof_shape(), dtype=s.dtype, device="cuda", generator=gen)
    a0 = 0.1 * torch.randn(s.dof_a_shape(), dtype=s.dtype, device="cuda", generator=gen)
    df, dp, af, ap = d0.clone(), d0.clone(), a0.clone(), a0.clone()
    tf, tp = torch.zeros_like(d0), torch.zeros_like(d0)
    for _ in range(2):
        s.local(df, tf, dt, af)
        s.neigh(df, tf, af)
        for a, n in _ranges(CUTS_L, m.n_el):
            s.local(dp, tp, dt, ap, first=a, n=n)
        for a, n in _ranges(CUTS_N, m.n_el):
            s.neigh(dp, tp, ap, first=a, n=n)
    torch.cuda.synchronize()
    assert torch.equal(df, dp) and torch.equal(af, ap)
    s.close()
