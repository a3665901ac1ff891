"""edge_el_l2_norms_noalloc (no device allocation, a static scratch array)
against edge_el_l2_norms, through the C ABI: bit-equal FP64 norms on two calls
in a row, the second on a smaller solver (fewer elements, fewer runs), so a
stale scratch from the first call would show.
"""
import ctypes as C

import numpy as np
import pytest

from edge_amd import _lib
from edge_amd import mesh as M
from edge_amd import setups as S
from edge_amd.solver import ElasticSolver

pytestmark = pytest.mark.gpu


def _norms(fn, s, a, b, runs):
    import torch
    d = np.full((9, runs), np.nan)
    n = np.full((9, runs), np.nan)
    pa = 64 if a.dtype == torch.float64 else 32
    pb = 0 if b is None else (64 if b.dtype == torch.float64 else 32)
    torch.cuda.synchronize()
    rc = fn(s._h, C.c_void_p(a.data_ptr()), pa, None if b is None else C.c_void_p(b.data_ptr()), pb,
            d.ctypes.data_as(C.c_void_p), n.ctypes.data_as(C.c_void_p), None)
    assert rc == 0, _lib.lib().edge_last_error()
    return d, n


@pytest.mark.parametrize("prec", [64, 32])
def test_noalloc_norms_equal_allocating_norms(gpu, prec):
    import torch
    lib = _lib.lib()
    cases = [(5, 16, 3), (3, 8, 2)]  # (hexes per axis, runs, order): the second call is smaller
    for n_hex, runs, order in cases:
        print(f"[no_alloc] prec={prec} hex={n_hex} runs={runs} order={order}", flush=True)
        m = M.box(n_hex, jitter=0.1, seed=n_hex)
        rho, lam, mu = S.random_materials(m.n_el, seed=n_hex)
        s = ElasticSolver(m, order, runs, prec, rho, lam, mu)
        dt = torch.float64 if prec == 64 else torch.float32
        g = torch.Generator(device="cpu").manual_seed(n_hex * 100 + runs)
        a = torch.randn(s.dof_shape(), generator=g, dtype=torch.float64).to(dt).cuda()
        b = torch.randn(s.dof_shape(), generator=g, dtype=torch.float64).to(dt).cuda()
        for bb in (b, None):
            ref = _norms(lib.edge_el_l2_norms, s, a, bb, runs)
            got = _norms(lib.edge_el_l2_norms_noalloc, s, a, bb, runs)
            assert np.all(np.isfinite(got[0])) and np.all(np.isfinite(got[1]))
            assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])
        s.close()


def test_noalloc_norms_reject_bad_arguments(gpu):
    lib = _lib.lib()
    d = np.zeros(9)
    assert lib.edge_el_l2_norms_noalloc(None, None, 64, None, 0, d.ctypes.data_as(C.c_void_p), None, None) != 0
