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


def test_receiver_csv_layout(tmp_path):
    """io::Receivers' file layout (src/io/Receivers.cpp:178-220, 292-318):
    comment header with name and coordinates, columns time,Q<q>_C<run>
    quantity-major, one row per output time; values read back exactly to the
    printed precision."""
    from types import SimpleNamespace

    from edge_amd.points import Receivers
    r = Receivers.__new__(Receivers)
    r.solver = SimpleNamespace(n_runs=2)
    r.n = 2
    r.points = np.array([[0.25, 0.5, 0.75], [0.1, 0.2, 0.3]])
    rng = np.random.default_rng(1)
    r.times = [[0.0, 0.125, 0.25], [0.5]]
    r.values = [[rng.normal(size=(9, 2)) for _ in range(3)], [rng.normal(size=(9, 2))]]
    paths = r.write_csv(str(tmp_path), names=["a", "b"])
    lines = open(paths[0]).read().splitlines()
    assert lines[0] == "# EDGE"
    assert "# receiver name: a" in lines
    assert "# specified coordinates: 0.25 0.5 0.75" in lines
    cols = [ln for ln in lines if not ln.startswith("#")][0].split(",")
    assert cols[:4] == ["time", "Q0_C0", "Q0_C1", "Q1_C0"] and cols[-1] == "Q8_C1" and len(cols) == 19
    rows = [ln for ln in lines if not ln.startswith("#")][1:]
    assert len(rows) == 3 and rows[1].split(",")[0] == "0.125000"
    got = np.array([[float(x) for x in row.split(",")[1:]] for row in rows]).reshape(3, 9, 2)
    np.testing.assert_allclose(got, np.array(r.values[0]), rtol=1e-6)
