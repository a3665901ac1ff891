"""The fused neighbour + local launch of the FP32 MFMA elastic kernels
(edge_el_neigh_local, and edge_el_step's fused sequence) through the C ABI:
the same DOFs bit for bit as separate neighbour and local steps, and the FP64
oracle's within 1e-5; taps hold the step-start DOFs the separate neighbour
step leaves; the argument checks."""
import numpy as np
import pytest

import oracle as orc
from edge_amd import mesh as M
from edge_amd import setups as S
from edge_amd._lib import EdgeError
from edge_amd.mesh import FA_FREE_SURFACE, FA_OUTFLOW, FA_PERIODIC
from edge_amd.solver import KERNEL_MFMA, PATH_FUSED, ElasticSolver

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def _dev(a, prec=32):
    import torch
    return torch.tensor(a, dtype=torch.float64 if prec == 64 else torch.float32, device="cuda").contiguous()


@pytest.mark.parametrize("order", [3, 4, 5])
def test_neigh_local_equals_separate_steps_and_oracle(gpu, order):
    import torch
    runs, n_steps = 16, 4
    m = M.box(4, bc=(FA_FREE_SURFACE, FA_OUTFLOW, FA_PERIODIC), jitter=0.1, seed=order)
    rho, lam, mu = S.random_materials(m.n_el, seed=order)
    dt = S.dt_cfl(m, rho, lam, mu, order)
    s = ElasticSolver(m, order, runs, 32, rho, lam, mu)
    assert s.kernel_info()[1] & PATH_FUSED
    d0 = np.random.default_rng(order).normal(size=(m.n_el, 9, S.n_modes(order), runs))
    # separate steps
    a, ta = _dev(d0), _dev(np.zeros_like(d0))
    for _ in range(n_steps):
        s.local(a, ta, dt)
        s.neigh(a, ta)
    # local, neigh_local x (n - 1), neigh; taps on a third of the elements
    b, tb = _dev(d0), _dev(np.zeros_like(d0))
    taps = torch.full_like(b, float("nan"))
    tap_el = np.arange(0, m.n_el, 3)
    s.set_taps(tap_el, taps)
    s.local(b, tb, dt)
    for _ in range(n_steps - 1):
        s.neigh_local(b, tb, dt)
    s.neigh(b, tb)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())
    np.testing.assert_array_equal(ta.cpu().numpy(), tb.cpu().numpy())
    # the taps hold the DOFs the last fused launch's local step started from:
    # those after n - 1 separate steps
    c, tc = _dev(d0), _dev(np.zeros_like(d0))
    for _ in range(n_steps - 1):
        s.local(c, tc, dt)
        s.neigh(c, tc)
    tp = taps.cpu().numpy()
    np.testing.assert_array_equal(tp[tap_el], c.cpu().numpy()[tap_el])
    others = np.setdiff1d(np.arange(m.n_el), tap_el)
    assert np.isnan(tp[others]).all()
    s.set_taps(None)
    # the fused multi-step call, eager and graph-captured (replayed twice)
    for graph in (False, True):
        e, te = _dev(d0), _dev(np.zeros_like(d0))
        s.step(e, te, dt, n_steps, graph=graph)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(a.cpu().numpy(), e.cpu().numpy())
    e2, te2 = _dev(d0), _dev(np.zeros_like(d0))
    s.step(e2, te2, dt, n_steps, graph=True)  # new arrays: a new capture
    s.step(e2, te2, dt, n_steps, graph=True)  # replay: 2 n steps in all
    torch.cuda.synchronize()
    f, tf = _dev(d0), _dev(np.zeros_like(d0))
    for _ in range(2 * n_steps):
        s.local(f, tf, dt)
        s.neigh(f, tf)
    np.testing.assert_array_equal(e2.cpu().numpy(), f.cpu().numpy())
    # the oracle
    mo = orc.MeshOracle(m, order, runs, rho=rho, lam=lam, mu=mu)
    d, t = d0.copy(), np.zeros_like(d0)
    mo.step(d, t, dt, n_steps)
    assert rel(b.cpu().numpy(), d) <= 1e-5
    s.close()


def test_neigh_local_argument_checks(gpu):
    order, runs = 4, 16
    m = M.box(3, jitter=0.1, seed=2)
    rho, lam, mu = S.random_materials(m.n_el, seed=2)
    dt = S.dt_cfl(m, rho, lam, mu, order)
    s = ElasticSolver(m, order, runs, 32, rho, lam, mu)
    d0 = np.random.default_rng(1).normal(size=(m.n_el, 9, S.n_modes(order), runs))
    d, t = _dev(d0), _dev(np.zeros_like(d0))
    with pytest.raises(EdgeError):  # no local step pending
        s.neigh_local(d, t, dt)
    s.local(d, t, dt)
    with pytest.raises(EdgeError):  # another tInt than the local step wrote
        s.neigh_local(d, t.clone(), dt)
    s.neigh_local(d, t, dt)
    s.neigh(d, t)
    with pytest.raises(EdgeError):  # the neighbour step has run
        s.neigh_local(d, t, dt)
    with pytest.raises(EdgeError):  # taps need an array
        s.set_taps([0], None)
    s.set_variant(KERNEL_MFMA, projections=False)  # tInt-reading kernels: no fused launch
    assert not s.kernel_info()[1] & PATH_FUSED
    s.local(d, t, dt)
    with pytest.raises(EdgeError):
        s.neigh_local(d, t, dt)
    s.close()
    s64 = ElasticSolver(m, order, 1, 64, rho, lam, mu)
    d64 = _dev(d0[..., :1], 64)
    t64 = _dev(np.zeros_like(d0[..., :1]), 64)
    s64.local(d64, t64, dt)
    with pytest.raises(EdgeError):  # FP64: unsupported
        s64.neigh_local(d64, t64, dt)
    s64.close()
