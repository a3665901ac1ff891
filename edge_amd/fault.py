"""Host side of a dynamic-rupture fault coupled to the element update
(fault.hip over libedge_amd's C ABI): tagging fault faces in a mesh, the
sub-face geometry that initial fault stresses are given on, and a handle for
the fault phase of the time step.

The pieces come from the reference's internal boundary
(src/impl/seismic/solvers/InternalBoundary.hpp:184-280, :743-923), linear
slip weakening (src/impl/seismic/solvers/FrictionLaws.hpp:309-392) and the
per-face friction data (src/impl/seismic/setups/RuptureInit.hpp:170-206);
the reference has no call site for them in its time loop, the fault phase
(between the local and the neighbour step) is this library's.
"""
import ctypes as C

import numpy as np

from . import _lib
from .solver import _stream_ptr

FA_RUPTURE = 201  # the reference's RUPTURE face type (src/impl/seismic/const.inc:142-148)
STATE_VALS = 14   # sn0, co0, ss0[2], ss0A, muf, st, tr[2], sr[2], dd[2]

# vertices of each local face of the reference tet, in the order of the face
# map chi -> xi (faces 0: 0-2-1, 1: 0-1-3, 2: 0-3-2, 3: 1-2-3): the face
# point chi lies at v[0] + chi_1 (v[1] - v[0]) + chi_2 (v[2] - v[0])
_FACE_VES = [(0, 2, 1), (0, 1, 3), (0, 3, 2), (1, 2, 3)]


def _face_vertices(mesh, faces):
    """Vertex coordinates [n][3][3] of faces seen from their first element, in
    that element's face-map order."""
    el = mesh.fa_el[faces, 0]
    loc = np.argmax(mesh.el_fa[el] == faces[:, None], axis=1)
    idx = np.array(_FACE_VES)[loc]  # [n][3] local vertex ids
    return mesh.ve_crds[mesh.el_ve[el[:, None], idx]], loc


def tag_plane(mesh, axis, value, tol=1e-9):
    """Sets the face type of every interior face lying in the plane
    x[axis] = value to FA_RUPTURE; returns the tagged face ids."""
    inner = np.nonzero((mesh.fa_el[:, 0] >= 0) & (mesh.fa_el[:, 1] >= 0))[0]
    v, _ = _face_vertices(mesh, inner)
    on = np.all(np.abs(v[:, :, axis] - value) <= tol, axis=1)
    faces = inner[on]
    mesh.fa_type[faces] = FA_RUPTURE
    return faces.astype(np.int64)


def n_sub_faces(order):
    return (2 * order - 1) ** 2


def sub_face_centroids(order):
    """Centroids [n_sf][2] of a face's sub-triangles in face coordinates
    (rows along chi_1, upward then downward triangle per column)."""
    n = 2 * order - 1
    out = []
    for e2 in range(n):
        for e1 in range(n):
            if e1 < n - e2:
                out.append(((e1 + 1 / 3) / n, (e2 + 1 / 3) / n))
            if e1 < n - e2 - 1:
                out.append(((e1 + 2 / 3) / n, (e2 + 2 / 3) / n))
    return np.array(out)


def fault_faces(mesh):
    """(faces, elements [nF][2], local face of the first element) of the
    mesh's fault faces, in the order the fault holds them."""
    faces = np.nonzero((mesh.fa_type & FA_RUPTURE) == FA_RUPTURE)[0]
    els = mesh.fa_el[faces]
    loc = np.argmax(mesh.el_fa[els[:, 0]] == faces[:, None], axis=1).astype(np.int64)
    return faces, els, loc


def sub_face_points(mesh, order):
    """Physical centroids [nF][n_sf][3] of the fault faces' sub-faces."""
    faces, _, _ = fault_faces(mesh)
    c = sub_face_centroids(order)
    v, _ = _face_vertices(mesh, faces)
    return (v[:, None, 0] + c[None, :, :1] * (v[:, None, 1] - v[:, None, 0]) +
            c[None, :, 1:] * (v[:, None, 2] - v[:, None, 0]))


def lsw_state(n_faces, n_sf, n_runs, sn0, ss0_0, ss0_1, mus, co0=0.0):
    """Initial linear slip-weakening state [nF][n_sf][R][14] as RuptureInit
    sets it (RuptureInit.hpp:256-290): initial normal and shear tractions,
    cohesion, the friction coefficient at its static value, zero slip, slip
    rate and traction.  Each value broadcasts to [nF][n_sf][R]; mus per run."""
    st = np.zeros((n_faces, n_sf, n_runs, STATE_VALS))
    shp = (n_faces, n_sf, n_runs)
    st[..., 0] = np.broadcast_to(sn0, shp)
    st[..., 1] = np.broadcast_to(co0, shp)
    st[..., 2] = np.broadcast_to(ss0_0, shp)
    st[..., 3] = np.broadcast_to(ss0_1, shp)
    st[..., 5] = np.broadcast_to(np.asarray(mus)[None, None, :], shp)
    return st


class Fault:
    """The fault of an elastic solver's mesh (faces of type FA_RUPTURE).

    fault_frame [3][3]: fault normal, strike, dip; glob [3][R]: static and
    dynamic friction coefficient and 1/Dc per run; state0 [nF][n_sf][R][14]."""

    def __init__(self, solver, mesh, rho, lam, mu, fault_frame, glob, state0, attach=True):
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        rho, lam, mu, fault_frame, glob, state0 = map(f, (rho, lam, mu, fault_frame, glob, state0))
        self.solver = solver
        faces, _, _ = fault_faces(mesh)
        self.n_faces, self.n_sf, self.n_runs = len(faces), n_sub_faces(solver.order), solver.n_runs
        want = (self.n_faces, self.n_sf, self.n_runs, STATE_VALS)
        if state0.shape != want:
            raise ValueError(f"state0: expected shape {want}, got {state0.shape}")
        if glob.shape != (3, self.n_runs):
            raise ValueError(f"glob: expected shape (3, {self.n_runs})")
        mi = mesh.mesh_in()
        h = C.c_void_p()
        _lib.check(_lib.lib().edge_fault_create(solver._h, C.byref(mi), rho.ctypes.data, lam.ctypes.data,
                                                mu.ctypes.data, fault_frame.ctypes.data, glob.ctypes.data,
                                                state0.ctypes.data, C.byref(h)))
        self._h = h
        if attach:
            _lib.check(_lib.lib().edge_el_set_fault(solver._h, h))
        self.attached = attach

    def apply(self, dofs, t_int, dt, stream=None):
        """The fault phase alone (after the local step, before the neighbour step)."""
        self.solver._check_arr(dofs)
        self.solver._check_arr(t_int)
        _lib.check(_lib.lib().edge_fault_apply(self._h, self.solver._h, float(dt), t_int.data_ptr(), dofs.data_ptr(),
                                               _stream_ptr(stream)))

    def get(self):
        """(state [nF][n_sf][R][14], slipping [nF][n_sf][R]) of the last step."""
        st = np.zeros((self.n_faces, self.n_sf, self.n_runs, STATE_VALS))
        per = np.zeros((self.n_faces, self.n_sf, self.n_runs), dtype=np.uint8)
        _lib.check(_lib.lib().edge_fault_get(self._h, st.ctypes.data, per.ctypes.data))
        return st, per.astype(bool)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            if self.attached and getattr(self.solver, "_h", None) is not None and self.solver._h.value:
                _lib.lib().edge_el_set_fault(self.solver._h, None)
            _lib.lib().edge_fault_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
