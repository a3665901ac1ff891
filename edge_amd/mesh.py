"""Box meshes of tet4 elements (6 Kuhn tets per hex) through the library's
C++ generator, returned as numpy arrays in the reference's entity layout."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import FA_FREE_SURFACE, FA_OUTFLOW, FA_PERIODIC  # noqa: F401


class Mesh:
    """Entity arrays of a tet4 mesh (see include/edge_amd.h, edge_mesh_out)."""

    def __init__(self, **arrays):
        self.__dict__.update(arrays)
        self.n_el = self.el_ve.shape[0]
        self.n_fa = self.fa_el.shape[0]
        self.n_ve = self.ve_crds.shape[0]

    def mesh_in(self):
        """C struct edge_mesh_in pointing at this mesh's arrays."""
        m = _lib.MeshIn()
        m.n_el, m.n_fa = self.n_el, self.n_fa
        for name in ["ve_crds", "el_ve", "el_fa", "fa_el", "el_fa_el", "f_id", "v_id", "fa_type", "normals",
                     "tangents", "areas"]:
            setattr(m, name, self.__dict__[name].ctypes.data)
        return m

    def el_crds(self):
        """Vertex coordinates per element [nEl][4][3]."""
        return self.ve_crds[self.el_ve]


def box(n, extent=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0), bc=(FA_PERIODIC,) * 3, jitter=0.0, seed=3343,
        block=1):
    """Box of n[0] x n[1] x n[2] hexes, each cut into 6 tets.

    bc per axis: FA_PERIODIC, FA_FREE_SURFACE or FA_OUTFLOW.  jitter moves the
    interior vertices by U[-jitter h, jitter h] per axis.  block orders the
    elements in blocks of block^3 hexes."""
    if np.isscalar(n):
        n = (n, n, n)
    p = _lib.BoxParams()
    p.n[:] = [int(x) for x in n]
    p.origin[:] = [float(x) for x in origin]
    p.extent[:] = [float(x) for x in extent]
    p.bc[:] = [int(x) for x in bc]
    p.jitter = float(jitter)
    p.seed = int(seed)
    p.block = int(block)
    L = _lib.lib()
    nve, nel, nfa = C.c_int64(), C.c_int64(), C.c_int64()
    _lib.check(L.edge_mesh_box_sizes(C.byref(p), C.byref(nve), C.byref(nel), C.byref(nfa)))
    nve, nel, nfa = nve.value, nel.value, nfa.value
    a = dict(
        ve_crds=np.zeros((nve, 3)), el_ve=np.zeros((nel, 4), np.int64), el_fa=np.zeros((nel, 4), np.int64),
        fa_el=np.zeros((nfa, 2), np.int64), el_fa_el=np.zeros((nel, 4), np.int64),
        f_id=np.zeros((nel, 4), np.uint16), v_id=np.zeros((nel, 4), np.uint16), fa_type=np.zeros(nfa, np.int32),
        normals=np.zeros((nfa, 3)), tangents=np.zeros((nfa, 2, 3)), areas=np.zeros(nfa), in_dia=np.zeros(nel),
        volume=np.zeros(nel))
    m = _lib.MeshOut()
    for k, v in a.items():
        setattr(m, k, v.ctypes.data)
    _lib.check(L.edge_mesh_box(C.byref(p), C.byref(m)))
    return Mesh(**a)


def _out_arrays(nve, nel, nfa):
    return dict(
        ve_crds=np.zeros((nve, 3)), el_ve=np.zeros((nel, 4), np.int64), el_fa=np.zeros((nel, 4), np.int64),
        fa_el=np.zeros((nfa, 2), np.int64), el_fa_el=np.zeros((nel, 4), np.int64),
        f_id=np.zeros((nel, 4), np.uint16), v_id=np.zeros((nel, 4), np.uint16), fa_type=np.zeros(nfa, np.int32),
        normals=np.zeros((nfa, 3)), tangents=np.zeros((nfa, 2, 3)), areas=np.zeros(nfa), in_dia=np.zeros(nel),
        volume=np.zeros(nel))


def from_tets(ve_crds, el_ve, bf_ve=None, bf_type=None, default_bc=FA_OUTFLOW):
    """Mesh of any tet4 elements: vertex coordinates [nVe][3], element
    vertices [nEl][4] (any orientation), tagged triangles bf_ve [nBf][3] with
    their face types bf_type [nBf] (boundary conditions, internal
    boundaries); untagged boundary faces get default_bc."""
    ve_crds = np.ascontiguousarray(ve_crds, dtype=np.float64)
    el_ve = np.ascontiguousarray(el_ve, dtype=np.int64)
    bf_ve = np.ascontiguousarray(np.zeros((0, 3)) if bf_ve is None else bf_ve, dtype=np.int64)
    bf_type = np.ascontiguousarray(np.zeros(0) if bf_type is None else bf_type, dtype=np.int32)
    t = _lib.TetsIn(len(ve_crds), len(el_ve), ve_crds.ctypes.data, el_ve.ctypes.data, len(bf_ve),
                    bf_ve.ctypes.data, bf_type.ctypes.data, int(default_bc))
    L = _lib.lib()
    nfa = C.c_int64()
    _lib.check(L.edge_mesh_tets_sizes(C.byref(t), C.byref(nfa)))
    a = _out_arrays(len(ve_crds), len(el_ve), nfa.value)
    m = _lib.MeshOut()
    for k, v in a.items():
        setattr(m, k, v.ctypes.data)
    _lib.check(L.edge_mesh_tets(C.byref(t), C.byref(m)))
    return Mesh(**a)


def union(meshes):
    """Disjoint union of meshes as one mesh (element, face and vertex ids of
    member i follow those of members 0..i-1; missing neighbours stay -1), and
    the element range (first, n) of each member.  One solver over the union
    advances every member in one launch per kernel."""
    ve, el, fa = 0, 0, 0
    parts = {k: [] for k in ("ve_crds", "el_ve", "el_fa", "fa_el", "el_fa_el", "f_id", "v_id", "fa_type", "normals",
                             "tangents", "areas", "in_dia", "volume")}
    ranges = []
    for m in meshes:
        parts["ve_crds"].append(m.ve_crds)
        parts["el_ve"].append(m.el_ve + ve)
        parts["el_fa"].append(m.el_fa + fa)
        parts["fa_el"].append(np.where(m.fa_el >= 0, m.fa_el + el, m.fa_el))
        parts["el_fa_el"].append(np.where(m.el_fa_el >= 0, m.el_fa_el + el, m.el_fa_el))
        for k in ("f_id", "v_id", "fa_type", "normals", "tangents", "areas", "in_dia", "volume"):
            parts[k].append(getattr(m, k))
        ranges.append((el, m.n_el))
        ve, el, fa = ve + m.n_ve, el + m.n_el, fa + m.n_fa
    return Mesh(**{k: np.ascontiguousarray(np.concatenate(v)) for k, v in parts.items()}), ranges
