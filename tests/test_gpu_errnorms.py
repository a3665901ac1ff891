"""L1, squared L2 and Linf errors on the GPU (edge_el_error_norms) against
plane waves: against the host restatement of InitialDofs::err
(setups.error_norms: the same quadrature, numpy), and, with zero DOFs, against
the closed-form norms of sin(2 pi x) over the periodic unit cube
(L2^2 = r^2 / 2, L1 = 2 |r| / pi, Linf = |r| up to where the points fall)."""
import numpy as np
import pytest

from edge_amd import mesh as M
from edge_amd import setups as S
from edge_amd.solver import KERNEL_MFMA, KERNEL_MFMA64, ElasticSolver

pytestmark = pytest.mark.gpu


def _dev(a, prec):
    import torch
    return torch.tensor(a, dtype=torch.float64 if prec == 64 else torch.float32, device="cuda").contiguous()


@pytest.mark.parametrize("prec,runs,kernel", [(64, 1, KERNEL_MFMA64), (32, 16, KERNEL_MFMA)])
def test_error_norms_match_host(gpu, prec, runs, kernel):
    order = 5
    m = M.box(3, jitter=0.1, seed=8)
    rho, lam, mu = np.ones(m.n_el), 2 * np.ones(m.n_el), np.ones(m.n_el)
    s = ElasticSolver(m, order, runs, prec, rho, lam, mu, kernel=kernel)
    rng = np.random.default_rng(4)
    waves = [S.PlaneWave(k=2 * np.pi * rng.integers(1, 3, size=3), amp=rng.uniform(0.5, 2.0)) for _ in range(runs)]
    q0 = np.stack([S.project(m, order, w)[..., 0] for w in waves], axis=-1)
    d, t_int = _dev(q0, prec), _dev(np.zeros_like(q0), prec)
    dt = S.dt_cfl(m, rho, lam, mu, order)
    s.step(d, t_int, dt, 5)
    t = 5 * dt
    l1, l2, li = s.error_norms(d, [[w] for w in waves], t)

    def fun(x):
        return np.stack([w(x, t) for w in waves], axis=-1)

    h1, h2, hi = S.error_norms(m, order, d.double().cpu().numpy(), fun)
    assert np.all(h2 > 0)
    np.testing.assert_allclose(l1, h1, rtol=1e-11)
    np.testing.assert_allclose(l2, h2, rtol=1e-11)
    np.testing.assert_allclose(li, hi, rtol=1e-11)
    s.close()


def test_error_norms_closed_form(gpu):
    m = M.box(4)
    n = m.n_el
    s = ElasticSolver(m, 4, 1, 64, np.ones(n), 2 * np.ones(n), np.ones(n), kernel=KERNEL_MFMA64)
    pw = S.PlaneWave(k=(2 * np.pi, 0.0, 0.0), amp=1.3)
    l1, l2, li = s.error_norms(s.alloc(), [[pw]], 0.0, n1d=8)
    r = np.abs(pw.r)
    np.testing.assert_allclose(l2[:, 0], r ** 2 / 2, rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(l1[:, 0], 2 * r / np.pi, rtol=1e-6, atol=1e-14)
    assert np.all(li[:, 0] <= r * (1 + 1e-14)) and np.all(li[:, 0] >= 0.98 * r)
    s.close()
