"""Host logic of receivers and point sources (no GPU): point location in the
mesh (reference coordinates of the affine tet map), the receivers' output
times per step (Receivers::getRecvTimeRel, src/io/Receivers.cpp:235-248),
and the C ABI's argument checks."""
import ctypes as C

import numpy as np
import pytest

from edge_amd import _lib
from edge_amd import mesh as M
from edge_amd.points import locate, output_times


def test_locate_round_trip():
    m = M.box(4, jitter=0.15, seed=5)
    rng = np.random.default_rng(0)
    ve = m.el_crds()
    el = rng.choice(m.n_el, 200, replace=False)
    bary = rng.dirichlet(np.ones(4), 200) * 0.8 + 0.05  # away from faces
    bary /= bary.sum(axis=1, keepdims=True)
    x = np.einsum("nv,nvd->nd", bary, ve[el])
    got_el, got_xi = locate(m, x)
    np.testing.assert_array_equal(got_el, el)
    np.testing.assert_allclose(got_xi, bary[:, 1:], atol=1e-12)


def test_locate_rejects_outside_points():
    m = M.box(3)
    with pytest.raises(ValueError):
        locate(m, [[1.5, 0.5, 0.5]])


def test_output_times_follow_the_reference_stepping():
    # every 0.3 from 0.1, steps of 0.25: outputs land in the step containing them
    nxt = np.array([0.1, 0.1])
    got = []
    for s in range(8):
        t = 0.25 * s
        tau = output_times(nxt, t, 0.25, 0.3, 1e-9)
        got += [t + x for x in tau[0] if x >= 0]
        np.testing.assert_array_equal(tau[0], tau[1])
    np.testing.assert_allclose(got, [0.1, 0.4, 0.7, 1.0, 1.3, 1.6, 1.9], atol=1e-12)
    # several outputs in one step, capped at 8 per call
    tau = output_times(np.array([0.0]), 0.0, 1.0, 0.1, 1e-9)
    assert tau.shape == (1, 8)
    np.testing.assert_allclose(tau[0], 0.1 * np.arange(8), atol=1e-12)


def test_points_c_abi_checks_arguments():
    L = _lib.lib()
    h = C.c_void_p()
    el = np.zeros(1, np.int64)
    z = np.zeros(64)
    assert L.edge_recv_create(None, 1, el.ctypes.data, z.ctypes.data, C.byref(h)) == _lib.EDGE_ERR_ARG
    assert not h.value
    assert L.edge_src_create(None, 1, el.ctypes.data, el.ctypes.data, z.ctypes.data, z.ctypes.data, z.ctypes.data,
                             z.ctypes.data, el.ctypes.data, z.ctypes.data, C.byref(h)) == _lib.EDGE_ERR_ARG
    assert L.edge_recv_eval(None, None, None, 1, None, None, None) == _lib.EDGE_ERR_ARG
    assert L.edge_src_apply(None, None, 0.0, 1.0, None, None) == _lib.EDGE_ERR_ARG
    assert L.edge_el_point_basis(None, 1, el.ctypes.data, z.ctypes.data, 0, z.ctypes.data) == _lib.EDGE_ERR_ARG
    assert L.edge_points_destroy(None) == 0
