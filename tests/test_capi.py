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
