"""Host-side mirror of edge::seismic::solvers::AderDg for the element update
(src/impl/seismic/solvers/AderDg.hpp:225-722): a solver handle over
libedge_amd's C ABI.  DOF and time-integrated DOF arrays are caller-owned
device tensors in the reference layout [el][9][mode][run].

torch is used only for device memory and streams.
"""
import ctypes as C

import numpy as np

from . import _lib
from .setups import n_modes

KERNEL_AUTO, KERNEL_GENERIC, KERNEL_MFMA, KERNEL_MFMA_TILES, KERNEL_MFMA64, KERNEL_EC64 = 0, 1, 2, 3, 4, 5
PATH_SPARSE_STAR, PATH_RANK3_FS, PATH_PROJECTIONS, PATH_VISCO, PATH_FUSED, PATH_HI_OCC = 1, 2, 4, 8, 16, 32


def _stream_ptr(stream):
    if stream is None:
        return None
    return stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream)


class ElasticSolver:
    """ADER-DG element update (elastic, tet4) on one GPU.

    Built either from mesh geometry and materials (star matrices and flux
    solvers are then set up in FP64 on the GPU), or from caller-given star
    matrices [nEl][3][81] and flux solvers [2][nEl][4][81] (``raw``)."""

    def __init__(self, mesh, order, n_runs=1, precision=64, rho=None, lam=None, mu=None, star=None, fs=None,
                 kernel=KERNEL_AUTO):
        self.mesh = mesh
        self.order = order
        self.n_runs = n_runs
        self.precision = precision
        self.n_modes = n_modes(order)
        cfg = _lib.ElConfig(order, n_runs, precision, kernel)
        mi = mesh.mesh_in()
        h = C.c_void_p()
        L = _lib.lib()
        if star is not None:
            star = np.ascontiguousarray(star, dtype=np.float64)
            fs = np.ascontiguousarray(fs, dtype=np.float64)
            _lib.check(L.edge_el_create_raw(C.byref(cfg), C.byref(mi), star.ctypes.data, fs.ctypes.data,
                                            C.byref(h)))
        else:
            rho, lam, mu = (np.ascontiguousarray(a, dtype=np.float64) for a in (rho, lam, mu))
            _lib.check(L.edge_el_create(C.byref(cfg), C.byref(mi), rho.ctypes.data, lam.ctypes.data,
                                        mu.ctypes.data, C.byref(h)))
        self._h = h

    # ------------------------------------------------------------------
    def kernel_info(self):
        """(kernel, flags): 1 generic / 2 MFMA-tiled / 4 FP64 MFMA / 5 element-batched; flags PATH_* bits."""
        k, f = C.c_int32(), C.c_int32()
        _lib.check(_lib.lib().edge_el_kernel_info(self._h, C.byref(k), C.byref(f)))
        return k.value, f.value

    def energy(self, dofs, stream=None):
        """(kinetic [R], strain [R]) elastic energies of the DOFs, FP64 on the GPU."""
        self._check_arr(dofs)
        k, e = np.zeros(self.n_runs), np.zeros(self.n_runs)
        _lib.check(_lib.lib().edge_el_energy(self._h, dofs.data_ptr(), k.ctypes.data, e.ctypes.data,
                                             _stream_ptr(stream)))
        return k, e

    def error_norms(self, dofs, waves, t=0.0, n1d=0, stream=None):
        """(L1, squared L2, Linf) [9][R] errors, FP64 on the GPU, against plane
        waves: waves[r] lists (r[9], k[3], omega, phase) per run, or objects with
        .r, .k, .omega (phase 0)."""
        self._check_arr(dofs)
        nw = max(len(w) for w in waves) if len(waves) else 0
        arr = np.zeros((self.n_runs, nw, 14))
        for r in range(self.n_runs):
            for i, w in enumerate(waves[r]):
                if hasattr(w, "omega"):
                    arr[r, i, :9], arr[r, i, 9:12], arr[r, i, 12] = w.r, w.k, w.omega
                else:
                    rv, kv, om, ph = w
                    arr[r, i, :9], arr[r, i, 9:12], arr[r, i, 12], arr[r, i, 13] = rv, kv, om, ph
        out = [np.zeros((9, self.n_runs)) for _ in range(3)]
        _lib.check(_lib.lib().edge_el_error_norms(self._h, dofs.data_ptr(), int(n1d), nw, arr.ctypes.data, float(t),
                                                  *[o.ctypes.data for o in out], _stream_ptr(stream)))
        return tuple(out)

    def autotune(self, dt, n_reps=3, stream=None):
        """Times the kernel variants and schedules this solver can switch to on
        its mesh and keeps the fastest; returns [(kernel, projections,
        xcd_schedule, ms per step, fused, occupancy)] (fused: multi-step calls
        run the neighbour step of step k - 1 and the local step of step k as
        one launch; occupancy: the higher-occupancy variant of the kernels)."""
        cap = 32
        n = C.c_int32(cap)
        k, f, ms = (C.c_int32 * cap)(), (C.c_int32 * cap)(), (C.c_double * cap)()
        _lib.check(_lib.lib().edge_el_autotune(self._h, float(dt), int(n_reps), _stream_ptr(stream), C.byref(n), k,
                                               f, ms))
        return [(k[i], bool(f[i] & 1), bool(f[i] & 2), ms[i], bool(f[i] & 4), bool(f[i] & 8))
                for i in range(min(n.value, cap))]

    def set_variant(self, kernel, projections=False, xcd_schedule=True, fused=False, occupancy=False):
        """Switches to one of the autotuner's candidates without timing it."""
        flags = (1 if projections else 0) | (2 if xcd_schedule else 0) | (4 if fused else 0) | (8 if occupancy else 0)
        _lib.check(_lib.lib().edge_el_set_variant(self._h, int(kernel), flags))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().edge_el_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def dtype(self):
        import torch
        return torch.float64 if self.precision == 64 else torch.float32

    def dof_shape(self):
        return (self.mesh.n_el, 9, self.n_modes, self.n_runs)

    def alloc(self, device="cuda"):
        import torch
        return torch.zeros(self.dof_shape(), dtype=self.dtype, device=device)

    def setup(self):
        """(star [nEl][3][81], fs [2][nEl][4][81]) as FP64 host arrays."""
        ne = self.mesh.n_el
        star = np.zeros((ne, 3, 81))
        fs = np.zeros((2, ne, 4, 81))
        _lib.check(_lib.lib().edge_el_get_setup(self._h, star.ctypes.data, fs.ctypes.data))
        return star, fs

    def _check_arr(self, a):
        if tuple(a.shape) != self.dof_shape() or a.dtype != self.dtype or not a.is_cuda or not a.is_contiguous():
            raise ValueError(f"expected a contiguous {self.dtype} device tensor of shape {self.dof_shape()}")

    # ------------------------------------------------------------------
    def local(self, dofs, t_int, dt, first=0, n=None, stream=None):
        """AderDg::local over elements [first, first+n)."""
        self._check_arr(dofs)
        self._check_arr(t_int)
        n = self.mesh.n_el - first if n is None else n
        _lib.check(_lib.lib().edge_el_local(self._h, first, n, float(dt), dofs.data_ptr(), t_int.data_ptr(),
                                            _stream_ptr(stream)))

    def neigh(self, dofs, t_int, first=0, n=None, stream=None):
        """AderDg::neigh over elements [first, first+n)."""
        self._check_arr(dofs)
        self._check_arr(t_int)
        n = self.mesh.n_el - first if n is None else n
        _lib.check(_lib.lib().edge_el_neigh(self._h, first, n, t_int.data_ptr(), dofs.data_ptr(),
                                            _stream_ptr(stream)))

    def neigh_local(self, dofs, t_int, dt, stream=None):
        """The pending step's neigh and the next step's local as one launch
        (solvers with PATH_FUSED)."""
        self._check_arr(dofs)
        self._check_arr(t_int)
        _lib.check(_lib.lib().edge_el_neigh_local(self._h, float(dt), dofs.data_ptr(), t_int.data_ptr(),
                                                  _stream_ptr(stream)))

    def set_taps(self, elements, taps=None):
        """Fused launches also write the step-start DOFs of `elements` to the
        device array `taps` (DOF layout); elements=None switches them off."""
        if elements is None or len(elements) == 0:
            _lib.check(_lib.lib().edge_el_set_taps(self._h, 0, None, None))
            return
        if taps is not None:  # (none: the C ABI refuses it)
            self._check_arr(taps)
        el = np.ascontiguousarray(elements, dtype=np.int64)
        self._taps = taps  # keeps the array alive while the solver writes it
        _lib.check(_lib.lib().edge_el_set_taps(self._h, len(el), el.ctypes.data,
                                               None if taps is None else taps.data_ptr()))

    def step(self, dofs, t_int, dt, n_steps=1, graph=False, stream=None):
        """n_steps global time steps (local, then neigh); with graph=True the
        step (or a fused sequence, PATH_FUSED) is captured into a HIP graph on
        first use and replayed."""
        self._check_arr(dofs)
        self._check_arr(t_int)
        _lib.check(_lib.lib().edge_el_step(self._h, int(n_steps), float(dt), dofs.data_ptr(), t_int.data_ptr(),
                                           1 if graph else 0, _stream_ptr(stream)))

    # ---- rate-2 local time stepping
    def set_lts(self, el_lts, fa_lts):
        """el_lts [nEl] bits (1: a larger-step neighbour, 2: a smaller-step
        one), fa_lts [nEl][4] (0 equal, 1 greater, 2 smaller); None, None: off."""
        if el_lts is None:
            _lib.check(_lib.lib().edge_el_set_lts(self._h, None, None))
            return
        el_lts = np.ascontiguousarray(el_lts, dtype=np.uint8)
        fa_lts = np.ascontiguousarray(fa_lts, dtype=np.uint8)
        _lib.check(_lib.lib().edge_el_set_lts(self._h, el_lts.ctypes.data, fa_lts.ctypes.data))

    def local_lts(self, dofs, t0, t1, t2, dt, first_ts, first=0, n=None, stream=None, dofs_a=None):
        """AderDg::local with local time stepping; dofs_a: the anelastic DOFs
        of a viscoelastic solver."""
        for a in (dofs, t0, t1, t2):
            self._check_arr(a)
        n = self.mesh.n_el - first if n is None else n
        if dofs_a is not None:
            _lib.check(_lib.lib().edge_el_local_lts_a(self._h, first, n, float(dt), 1 if first_ts else 0,
                                                      dofs.data_ptr(), dofs_a.data_ptr(), t0.data_ptr(),
                                                      t1.data_ptr(), t2.data_ptr(), _stream_ptr(stream)))
            return
        _lib.check(_lib.lib().edge_el_local_lts(self._h, first, n, float(dt), 1 if first_ts else 0, dofs.data_ptr(),
                                                t0.data_ptr(), t1.data_ptr(), t2.data_ptr(), _stream_ptr(stream)))

    def neigh_lts(self, dofs, t0, t1, t2, first_ts, first=0, n=None, stream=None, dofs_a=None):
        for a in (dofs, t0, t1, t2):
            self._check_arr(a)
        n = self.mesh.n_el - first if n is None else n
        if dofs_a is not None:
            _lib.check(_lib.lib().edge_el_neigh_lts_a(self._h, first, n, 1 if first_ts else 0, t0.data_ptr(),
                                                      t1.data_ptr(), t2.data_ptr(), dofs.data_ptr(),
                                                      dofs_a.data_ptr(), _stream_ptr(stream)))
            return
        _lib.check(_lib.lib().edge_el_neigh_lts(self._h, first, n, 1 if first_ts else 0, t0.data_ptr(), t1.data_ptr(),
                                                t2.data_ptr(), dofs.data_ptr(), _stream_ptr(stream)))

    def l2_norms(self, a, b=None, stream=None):
        """FP64 (||a-b||^2, ||b||^2) per [9][R] through the mass matrix."""
        import torch
        r = self.n_runs
        d = np.zeros((9, r))
        n = np.zeros((9, r))
        pa = 64 if a.dtype == torch.float64 else 32
        pb = 0 if b is None else (64 if b.dtype == torch.float64 else 32)
        _lib.check(_lib.lib().edge_el_l2_norms(self._h, a.data_ptr(), pa, None if b is None else b.data_ptr(),
                                               pb, d.ctypes.data, n.ctypes.data, _stream_ptr(stream)))
        return d, n


class ViscoSolver(ElasticSolver):
    """Viscoelastic ADER-DG update (N_RMS relaxation mechanisms) on one GPU:
    AderDg with TL_N_RMS > 0.  Built from phase-velocity Lame parameters and
    P / S quality factors per element; the constant-Q fit over the band
    (freq_cen, freq_rat) gives the unrelaxed Lame parameters and the
    anelastic source matrices.  Anelastic DOFs: [nEl][n_rms][6][N][R]."""

    def __init__(self, mesh, order, n_runs, precision, rho, lam, mu, qp, qs, n_rms=3, freq_cen=1.0,
                 freq_rat=10.0, kernel=KERNEL_AUTO):
        self.mesh = mesh
        self.order = order
        self.n_runs = n_runs
        self.precision = precision
        self.n_modes = n_modes(order)
        self.n_rms = n_rms
        cfg = _lib.ElConfig(order, n_runs, precision, kernel)
        vc = _lib.ViscoConfig(n_rms, freq_cen, freq_rat)
        mi = mesh.mesh_in()
        h = C.c_void_p()
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (rho, lam, mu, qp, qs)]
        _lib.check(_lib.lib().edge_el_create_visco(C.byref(cfg), C.byref(vc), C.byref(mi),
                                                   *[a.ctypes.data for a in arrs], C.byref(h)))
        self._h = h

    def dof_a_shape(self):
        return (self.mesh.n_el, self.n_rms, 6, self.n_modes, self.n_runs)

    def alloc_a(self, device="cuda"):
        import torch
        return torch.zeros(self.dof_a_shape(), dtype=self.dtype, device=device)

    def setup_a(self):
        """dict of the anelastic set-up as FP64 host arrays."""
        ne, nr = self.mesh.n_el, self.n_rms
        out = {"starA": np.zeros((ne, 3, 18)), "srcA": np.zeros((ne, nr, 36)), "fsA": np.zeros((2, ne, 4, 54)),
               "rfs": np.zeros(nr), "lam_e": np.zeros(ne), "mu_e": np.zeros(ne)}
        _lib.check(_lib.lib().edge_el_get_setup_a(self._h, *[out[k].ctypes.data for k in
                                                             ("starA", "srcA", "fsA", "rfs", "lam_e", "mu_e")]))
        return out

    def _check_a(self, a):
        if tuple(a.shape) != self.dof_a_shape() or a.dtype != self.dtype or not a.is_cuda or not a.is_contiguous():
            raise ValueError(f"expected a contiguous {self.dtype} device tensor of shape {self.dof_a_shape()}")

    def local(self, dofs, t_int, dt, dofs_a=None, first=0, n=None, stream=None):
        self._check_arr(dofs)
        self._check_arr(t_int)
        self._check_a(dofs_a)
        n = self.mesh.n_el - first if n is None else n
        _lib.check(_lib.lib().edge_el_local_a(self._h, first, n, float(dt), dofs.data_ptr(), dofs_a.data_ptr(),
                                              t_int.data_ptr(), _stream_ptr(stream)))

    def neigh(self, dofs, t_int, dofs_a=None, first=0, n=None, stream=None):
        self._check_arr(dofs)
        self._check_arr(t_int)
        self._check_a(dofs_a)
        n = self.mesh.n_el - first if n is None else n
        _lib.check(_lib.lib().edge_el_neigh_a(self._h, first, n, t_int.data_ptr(), dofs.data_ptr(),
                                              dofs_a.data_ptr(), _stream_ptr(stream)))

    def step(self, dofs, t_int, dt, n_steps=1, dofs_a=None, graph=False, stream=None):
        """n_steps global time steps with the anelastic DOFs; graph=True
        captures one step into a HIP graph on first use and replays it."""
        self._check_arr(dofs)
        self._check_arr(t_int)
        self._check_a(dofs_a)
        _lib.check(_lib.lib().edge_el_step_a(self._h, int(n_steps), float(dt), dofs.data_ptr(), dofs_a.data_ptr(),
                                             t_int.data_ptr(), 1 if graph else 0, _stream_ptr(stream)))
