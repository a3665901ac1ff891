"""Host-side mirror of EDGE's receivers (src/io/Receivers.cpp) and point
sources (src/impl/seismic/setups/PointSources.hpp) over libedge_amd's C ABI.

A time step with both runs, as the reference's steps (src/impl/seismic/
steps.inc: local with receiver output, then sources, then neigh):

    recv.sample(dofs, time, dt)      # before the local update
    solver.local(dofs, t_int, dt)
    src.apply(time, time + dt, dofs)
    solver.neigh(dofs, t_int)

Receivers and sources are located in the mesh on the host: the containing
element (largest smallest barycentric coordinate) and the reference
coordinates of the affine map x = v0 + xi (v1 - v0) + eta (v2 - v0) +
zeta (v3 - v0) (src/linalg/Mappings.hpp:416-535).
"""
import ctypes as C

import numpy as np

from . import _lib
from .solver import _stream_ptr

# quantity order of the elastic DOFs (src/impl/seismic/setups/Elasticity.cpp:172-212)
QUANTITIES = ("sxx", "syy", "szz", "sxy", "syz", "sxz", "u", "v", "w")
MAX_TAU = 8  # output times per receiver and evaluation call


def locate(mesh, points, tol=1e-9):
    """(element ids [n], reference coordinates [n][3]) of physical points [n][3]."""
    pts = np.atleast_2d(np.asarray(points, dtype=np.float64))
    ve = mesh.el_crds()  # [nEl][4][3]
    lo, hi = ve.min(axis=1) - tol, ve.max(axis=1) + tol
    m = np.transpose(ve[:, 1:, :] - ve[:, :1, :], (0, 2, 1))  # columns v_i - v_0
    el = np.zeros(len(pts), dtype=np.int64)
    xi = np.zeros((len(pts), 3))
    for i, x in enumerate(pts):
        cand = np.nonzero(np.all((x >= lo) & (x <= hi), axis=1))[0]
        if len(cand) == 0:
            raise ValueError(f"point {x} lies outside the mesh")
        r = np.linalg.solve(m[cand], (x - ve[cand, 0])[..., None])[..., 0]
        bary = np.concatenate([1 - r.sum(axis=1, keepdims=True), r], axis=1)
        best = int(np.argmax(bary.min(axis=1)))
        if bary[best].min() < -1e-6:
            raise ValueError(f"point {x} lies outside the mesh")
        el[i] = cand[best]
        xi[i] = np.clip(r[best], 0.0, 1.0)
    return el, xi


def point_basis(solver, el, xi, src_scaled=False):
    """Basis functions at the points [n][N] (divided by the mass and |J| for sources)."""
    el = np.ascontiguousarray(el, dtype=np.int64)
    xi = np.ascontiguousarray(xi, dtype=np.float64)
    out = np.zeros((len(el), solver.n_modes))
    _lib.check(_lib.lib().edge_el_point_basis(solver._h, len(el), el.ctypes.data, xi.ctypes.data,
                                              1 if src_scaled else 0, out.ctypes.data))
    return out


def output_times(next_times, time, dt, freq, tol):
    """Relative output times of the step [time, time + dt) per receiver
    (Receivers::getRecvTimeRel, src/io/Receivers.cpp:235-248, and the
    advance by the output frequency in writeRecvAll, :284): tau [n][k], -1
    where a receiver has fewer outputs.  Advances next_times in place."""
    taus = [[] for _ in range(len(next_times))]
    for i in range(len(next_times)):
        while next_times[i] > time - tol and next_times[i] < time + dt - tol and len(taus[i]) < MAX_TAU:
            taus[i].append(abs(next_times[i] - time))
            next_times[i] += freq
    k = max((len(t) for t in taus), default=0)
    tau = np.full((len(next_times), k), -1.0)
    for i, t in enumerate(taus):
        tau[i, :len(t)] = t
    return tau


class Receivers:
    """Receivers sampled every `freq` from time t0 (io::Receivers): the
    element's time prediction at each output time within a step, evaluated
    at the receiver's point.  Values per receiver: times [k] and
    values [k][9][R] (FP64 on the host)."""

    def __init__(self, solver, points, freq, t0=0.0):
        self.solver = solver
        self.points = np.asarray(points, dtype=np.float64).reshape(-1, 3)
        self.el, self.xi = locate(solver.mesh, points)
        self.n = len(self.el)
        self.freq = float(freq)
        self.next = np.full(self.n, float(t0))
        # TOL.TIME of the reference's build in this precision (src/constants.hpp:289-300)
        self.tol = 1e-9 if solver.precision == 64 else 1e-5
        basis = point_basis(solver, self.el, self.xi)
        h = C.c_void_p()
        _lib.check(_lib.lib().edge_recv_create(solver._h, self.n, self.el.ctypes.data, basis.ctypes.data, C.byref(h)))
        self._h = h
        self.times = [[] for _ in range(self.n)]
        self.values = [[] for _ in range(self.n)]

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().edge_points_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sample(self, dofs, time, dt, stream=None, dofs_a=None):
        """Records the outputs of the step [time, time + dt) from the DOFs at
        its start (and the anelastic DOFs for a viscoelastic solver); returns
        the number of output times evaluated."""
        import torch
        tau = output_times(self.next, time, dt, self.freq, self.tol)
        k = tau.shape[1]
        if k == 0 or self.n == 0:
            return 0
        r = self.solver.n_runs
        out = torch.empty((self.n, k, 9, r), dtype=dofs.dtype, device=dofs.device)
        tau = np.ascontiguousarray(tau)
        _lib.check(_lib.lib().edge_recv_eval_a(self._h, self.solver._h, dofs.data_ptr(),
                                               None if dofs_a is None else dofs_a.data_ptr(), k, tau.ctypes.data,
                                               out.data_ptr(), _stream_ptr(stream)))
        vals = out.double().cpu().numpy()
        for i in range(self.n):
            for j in range(k):
                if tau[i, j] >= 0:
                    self.times[i].append(time + tau[i, j])
                    self.values[i].append(vals[i, j])
        return k

    def eval_into(self, dofs, tau, out, stream=None):
        """Evaluates every receiver at relative times tau [n][k] of the coming
        step into the device tensor out [n][k][9][R], without host copies of
        the results (no output-time bookkeeping)."""
        tau = np.ascontiguousarray(tau, dtype=np.float64)
        _lib.check(_lib.lib().edge_recv_eval(self._h, self.solver._h, dofs.data_ptr(), tau.shape[1],
                                             tau.ctypes.data, out.data_ptr(), _stream_ptr(stream)))

    def seismogram(self, i):
        """(times [k], values [k][9][R]) of receiver i."""
        return np.array(self.times[i]), np.array(self.values[i]).reshape(-1, 9, self.solver.n_runs)

    def write_csv(self, out_dir, names=None):
        """One CSV per receiver, in the layout of io::Receivers
        (src/io/Receivers.cpp:178-220, 292-318): comment lines with the
        receiver's name and coordinates (the given point, which is also where
        it is evaluated: it lies in its element), the column names
        time,Q0_C0,Q0_C1,...,Q8_C{R-1} (quantity-major, run innermost), then
        one row per output time: the time with six decimals, the values in
        scientific notation.  Files are out_dir/<name>.csv; returns their
        paths."""
        import os
        r = self.solver.n_runs
        head = ",".join(["time"] + [f"Q{q}_C{c}" for q in range(9) for c in range(r)])
        paths = []
        for i in range(self.n):
            t, v = self.seismogram(i)
            name = names[i] if names is not None else f"receiver_{i}"
            x = " ".join(repr(float(c)) for c in self.points[i])
            path = os.path.join(out_dir, f"{name}.csv")
            with open(path, "w") as f:
                f.write("# EDGE\n# code version: edge_amd\n")
                f.write(f"# receiver name: {name}\n# specified coordinates: {x}\n# projected coordinates: {x}\n")
                f.write(head + "\n")
                for k in range(len(t)):
                    f.write(f"{t[k]:.6f}," + ",".join(f"{y:e}" for y in v[k].reshape(-1)) + "\n")
            paths.append(path)
        return paths


class PointSources:
    """Kinematic point sources (PointSources::t_ps with one fused run per
    source): per source a physical point, its run, quantity scalings [9], a
    series sampled every dts from t0."""

    def __init__(self, solver, points, runs, scas, t0, dts, series):
        self.solver = solver
        self.points = np.asarray(points, dtype=np.float64).reshape(-1, 3)
        self.el, self.xi = locate(solver.mesh, points)
        n = len(self.el)
        self.n = n
        bev = point_basis(solver, self.el, self.xi, src_scaled=True)
        runs = np.ascontiguousarray(np.broadcast_to(runs, (n,)), dtype=np.int32)
        scas = np.ascontiguousarray(np.broadcast_to(scas, (n, 9)), dtype=np.float64)
        t0 = np.ascontiguousarray(np.broadcast_to(t0, (n,)), dtype=np.float64)
        dts = np.ascontiguousarray(np.broadcast_to(dts, (n,)), dtype=np.float64)
        ptr = np.zeros(n + 1, dtype=np.int64)
        ptr[1:] = np.cumsum([len(s) for s in series])
        ts = np.ascontiguousarray(np.concatenate([np.asarray(s, dtype=np.float64) for s in series]))
        h = C.c_void_p()
        _lib.check(_lib.lib().edge_src_create(solver._h, n, self.el.ctypes.data, runs.ctypes.data, t0.ctypes.data,
                                              dts.ctypes.data, bev.ctypes.data, scas.ctypes.data, ptr.ctypes.data,
                                              ts.ctypes.data, C.byref(h)))
        self._h = h
        self.bev, self.runs, self.scas, self.t0, self.dts, self.series = bev, runs, scas, t0, dts, series

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().edge_points_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def apply(self, t1, t2, dofs, stream=None):
        """Adds the contributions of [t1, t2] to the DOFs (device tensor)."""
        _lib.check(_lib.lib().edge_src_apply(self._h, self.solver._h, float(t1), float(t2), dofs.data_ptr(),
                                             _stream_ptr(stream)))


def ricker(f0, t_peak, dt, n):
    """Ricker wavelet of peak frequency f0 centred at t_peak, sampled at
    0, dt, ..., (n - 1) dt: (1 - 2 a) exp(-a), a = (pi f0 (t - t_peak))^2."""
    t = np.arange(n) * dt
    a = (np.pi * f0 * (t - t_peak)) ** 2
    return (1 - 2 * a) * np.exp(-a)


def moment_tensor_scalings(m):
    """Quantity scalings [9] of a symmetric moment tensor m [3][3]: the six
    stress components in the DOFs' order (sxx, syy, szz, sxy, syz, sxz),
    velocities zero."""
    m = np.asarray(m, dtype=np.float64)
    return np.array([m[0, 0], m[1, 1], m[2, 2], m[0, 1], m[1, 2], m[0, 2], 0.0, 0.0, 0.0])
