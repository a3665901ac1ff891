"""The C-ABI library: it loads, exports every entry point include/edge_amd.h
declares, and its host-only entry points (matrices, quadrature, meshes)
validate their arguments.  No compute on a GPU here."""
import ctypes as C

import numpy as np
import pytest

from edge_amd import _lib


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _lib.declared_symbols()
    assert len(declared) >= 15
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing


def test_version_and_error_message():
    L = _lib.lib()
    assert b"gfx950" in L.edge_version()
    n, nf = C.c_int(), C.c_int()
    assert L.edge_dg_sizes(9, C.byref(n), C.byref(nf)) != 0
    assert b"order" in L.edge_last_error()
    assert L.edge_dg_sizes(5, C.byref(n), C.byref(nf)) == 0 and (n.value, nf.value) == (35, 15)


def test_bad_mesh_arguments_are_rejected():
    from edge_amd import mesh as M
    with pytest.raises(_lib.EdgeError):
        M.box((2, 4, 4))  # periodic axes need three hexes
    with pytest.raises(_lib.EdgeError):
        M.box(4, jitter=2.0)  # inverted tets


def test_solver_arguments_checked_before_device_use():
    """Configuration errors surface as error codes, before any HIP call."""
    from edge_amd import mesh as M
    m = M.box(3)
    L = _lib.lib()
    mi = m.mesh_in()
    h = C.c_void_p()
    rho = np.ones(m.n_el)
    for cfg in [_lib.ElConfig(7, 1, 64, 0), _lib.ElConfig(4, 3, 64, 0), _lib.ElConfig(4, 1, 16, 0)]:
        rc = L.edge_el_create(C.byref(cfg), C.byref(mi), rho.ctypes.data, rho.ctypes.data, rho.ctypes.data,
                              C.byref(h))
        assert rc != 0 and not h.value
    bad = -np.ones(m.n_el)
    rc = L.edge_el_create(C.byref(_lib.ElConfig(4, 1, 64, 0)), C.byref(mi), bad.ctypes.data, rho.ctypes.data,
                          rho.ctypes.data, C.byref(h))
    assert rc == 1 and b"material" in L.edge_last_error()
    assert L.edge_el_local(None, 0, 1, 0.1, None, None, None) == 1


def test_visco_create_rejects_bad_arguments():
    """Argument checks of edge_el_create_visco that return before any device work."""
    import ctypes as C
    import numpy as np
    from edge_amd import _lib
    from edge_amd import mesh as M
    m = M.box(3)
    ne = m.n_el
    one = np.ones(ne)
    good = [np.full(ne, 2.0), np.full(ne, 3.0), np.full(ne, 1.5), np.full(ne, 50.0), np.full(ne, 25.0)]
    L = _lib.lib()

    def create(order=4, runs=1, prec=64, n_rms=3, fc=1.0, fr=10.0, arrs=good):
        cfg = _lib.ElConfig(order, runs, prec, 0)
        vc = _lib.ViscoConfig(n_rms, fc, fr)
        h = C.c_void_p()
        rc = L.edge_el_create_visco(C.byref(cfg), C.byref(vc), C.byref(m.mesh_in()), *[a.ctypes.data for a in arrs],
                                    C.byref(h))
        return rc, L.edge_last_error().decode()

    assert create(n_rms=0)[0] != _lib.EDGE_OK
    assert create(n_rms=6)[0] != _lib.EDGE_OK
    assert create(fr=1.0)[0] != _lib.EDGE_OK
    assert create(fc=0.0)[0] != _lib.EDGE_OK
    assert create(runs=16, prec=32)[0] != _lib.EDGE_OK  # the anelastic tiles of 16 runs do not fit LDS
    rc, msg = create(arrs=good[:3] + [-one, one])
    assert rc != _lib.EDGE_OK and "quality" in msg


def test_rupture_create_rejects_bad_arguments():
    """Argument checks of edge_rupture_create that return before any device work."""
    import ctypes as C
    import numpy as np
    from edge_amd import _lib
    L = _lib.lib()
    nf, R = 2, 1
    e = np.eye(3)
    good = dict(normals=np.tile(e[0], (nf, 1)), tangents=np.tile(e[1:], (nf, 1, 1)), areas=np.ones(nf),
                fault_frame=e.copy(), mat_l=np.tile([2700.0, 3e10, 3e10], (nf, 1)),
                mat_r=np.tile([2700.0, 3e10, 3e10], (nf, 1)), vol_l=np.ones(nf), vol_r=np.ones(nf),
                glob=np.array([[0.6], [0.5], [2.0]]), state0=np.zeros((nf, 1, R, 14)))

    def create(n_sf=1, n_sc=1, runs=R, **over):
        a = dict(good, **over)
        h = C.c_void_p()
        rc = L.edge_rupture_create(C.byref(_lib.RuptureConfig(n_sf, n_sc, runs)), nf,
                                   *[np.ascontiguousarray(a[k], dtype=np.float64).ctypes.data for k in good],
                                   C.byref(h))
        assert not h.value
        return rc, L.edge_last_error().decode()

    assert create(n_sf=0)[0] == 1
    assert create(n_sc=-1)[0] == 1
    rc, msg = create(fault_frame=2 * e)
    assert rc == 1 and "unit" in msg
    rc, msg = create(glob=np.array([[0.6], [0.5], [0.0]]))
    assert rc == 1 and "1/Dc" in msg
    rc, msg = create(areas=np.array([1.0, -1.0]))
    assert rc == 1 and "area" in msg
    assert L.edge_rupture_net_updates(None, 0, 1, 0.1, None, None, None, None, None) == 1
    assert L.edge_rupture_destroy(None) == 0


def test_null_handles_are_rejected():
    """Entry points return EDGE_ERR_ARG for a missing solver handle before any
    device work."""
    L = _lib.lib()
    n = C.c_int32(4)
    assert L.edge_el_autotune(None, 1e-3, 3, None, C.byref(n), None, None, None) == 1
    assert L.edge_el_energy(None, None, None, None, None) == 1
    assert L.edge_el_error_norms(None, None, 0, 0, None, 0.0, None, None, None, None) == 1
    assert L.edge_el_step_a(None, 1, 1e-3, None, None, None, 0, None) == 1
    assert b"handle" in L.edge_last_error()


def test_missing_library_fails_loudly(monkeypatch):
    """Without the built HIP library every entry point raises; nothing falls
    back to host code."""
    monkeypatch.setattr(_lib, "LIB_PATH", _lib.LIB_PATH + ".absent")
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.EdgeError, match="missing"):
        _lib.lib()
