"""Host-side mirror of EDGE's dynamic-rupture face update: the internal
boundary solvers (edge::seismic::solvers::InternalBoundarySolvers,
src/impl/seismic/solvers/InternalBoundary.hpp:743-923), the per-face sides of
RuptureInit (src/impl/seismic/setups/RuptureInit.hpp:170-206) and the net
updates with linear slip-weakening friction (InternalBoundary::netUpdates,
InternalBoundary.hpp:184-280; FrictionLaws<3,R>::linSlipWeak,
FrictionLaws.hpp:309-392), over libedge_amd's C ABI.

Sub-cell DOFs and net updates are caller-owned FP64 device tensors
[nFaces][9][n_sf][R]; the friction state lives on the device.
"""
import ctypes as C

import numpy as np

from . import _lib
from .solver import _stream_ptr

STATE_VALS = 14  # sn0, co0, ss0[2], ss0A, muf, st, tr[2], sr[2], dd[2]


class Rupture:
    """Linear slip-weakening rupture faces on one GPU.

    normals [nF][3], tangents [nF][2][3]: the faces' frames w.r.t. the left
    element; areas [nF]; fault_frame [3][3]: fault normal, strike, dip;
    mat_l / mat_r [nF][3]: (rho, lambda, mu) of the left / right element;
    vol_l / vol_r [nF]; glob [3][R]: static and dynamic friction coefficient
    and 1/Dc per run; state0 [nF][n_sf][R][14]."""

    def __init__(self, n_sub_faces, n_sub_cells, n_runs, normals, tangents, areas, fault_frame, mat_l, mat_r, vol_l,
                 vol_r, glob, state0):
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        normals, tangents, areas, fault_frame = f(normals), f(tangents), f(areas), f(fault_frame)
        mat_l, mat_r, vol_l, vol_r, glob, state0 = f(mat_l), f(mat_r), f(vol_l), f(vol_r), f(glob), f(state0)
        nf = areas.shape[0]
        self.n_faces, self.n_sf, self.n_runs = nf, n_sub_faces, n_runs
        want = {"normals": (normals, (nf, 3)), "tangents": (tangents, (nf, 2, 3)), "fault_frame": (fault_frame, (3, 3)),
                "mat_l": (mat_l, (nf, 3)), "mat_r": (mat_r, (nf, 3)), "vol_l": (vol_l, (nf,)),
                "vol_r": (vol_r, (nf,)), "glob": (glob, (3, n_runs)),
                "state0": (state0, (nf, n_sub_faces, n_runs, STATE_VALS))}
        for k, (a, s) in want.items():
            if a.shape != s:
                raise ValueError(f"{k}: expected shape {s}, got {a.shape}")
        cfg = _lib.RuptureConfig(n_sub_faces, n_sub_cells, n_runs)
        h = C.c_void_p()
        _lib.check(_lib.lib().edge_rupture_create(C.byref(cfg), nf, *[a.ctypes.data for a in (
            normals, tangents, areas, fault_frame, mat_l, mat_r, vol_l, vol_r, glob, state0)], C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().edge_rupture_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def dof_shape(self):
        return (self.n_faces, 9, self.n_sf, self.n_runs)

    def alloc(self, device="cuda"):
        import torch
        return torch.zeros(self.dof_shape(), dtype=torch.float64, device=device)

    def _check(self, a):
        import torch
        if tuple(a.shape) != self.dof_shape() or a.dtype != torch.float64 or not a.is_cuda or not a.is_contiguous():
            raise ValueError(f"expected a contiguous float64 device tensor of shape {self.dof_shape()}")

    def net_updates(self, dofs_l, dofs_r, net_l, net_r, dt, first=0, n=None, stream=None):
        """Net updates of faces [first, first+n); advances the friction state."""
        for a in (dofs_l, dofs_r, net_l, net_r):
            self._check(a)
        n = self.n_faces - first if n is None else n
        _lib.check(_lib.lib().edge_rupture_net_updates(self._h, first, n, float(dt), dofs_l.data_ptr(),
                                                       dofs_r.data_ptr(), net_l.data_ptr(), net_r.data_ptr(),
                                                       _stream_ptr(stream)))

    def get(self):
        """(solvers [nF][4][9][9], state [nF][n_sf][R][14], failing [nF][n_sf][R])."""
        solv = np.zeros((self.n_faces, 4, 9, 9))
        st = np.zeros((self.n_faces, self.n_sf, self.n_runs, STATE_VALS))
        per = np.zeros((self.n_faces, self.n_sf, self.n_runs), dtype=np.uint8)
        _lib.check(_lib.lib().edge_rupture_get(self._h, solv.ctypes.data, st.ctypes.data, per.ctypes.data))
        return solv, st, per.astype(bool)
