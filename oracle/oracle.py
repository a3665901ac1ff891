"""Python side of the oracle.  TEST INFRASTRUCTURE ONLY.

ctypes wrappers around liboracle.so (the C restatement in edge_oracle.c)
and around oracle/_ref/libedge_ref.so (the reference's own Elasticity.cpp,
built from the reference tree when it is present).  Imported by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg only.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libedge_ref.so")

NO_ID16 = 0xFFFF
_dp = C.c_void_p


class OrcDg(C.Structure):
    _fields_ = [("order", C.c_int), ("n_crs", C.c_int), ("n_rms", C.c_int), ("stiffT", _dp), ("stiffV", _dp),
                ("fluxL", _dp), ("fluxN", _dp), ("fluxT", _dp), ("rfs", _dp)]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-C", HERE, "oracle"], check=True, capture_output=True)
        _lib = C.CDLL(LIB)
        _lib.orc_dt_cfl.restype = C.c_double
        _lib.orc_dt_cfl.argtypes = [C.c_double] * 4 + [C.c_int]
        for f in ["orc_ck", "orc_vol", "orc_surf_local", "orc_surf_neigh", "orc_neigh_flux_int", "orc_local",
                  "orc_neigh", "orc_star", "orc_fs", "orc_jac", "orc_trafo", "orc_fs_mid", "orc_init_fs",
                  "orc_integrate", "orc_eval_time_pred", "orc_scatter_update_a", "orc_visco_frequencies",
                  "orc_anelastic_coeff_ps", "orc_elastic_lame", "orc_anelastic_coeff_lame", "orc_src_a",
                  "orc_star_a", "orc_fs_mid_a", "orc_fs_a", "orc_init_fs_a", "orc_ms_jump_l", "orc_ms_jump_r",
                  "orc_ms_flux", "orc_rot_mat", "orc_lsw", "orc_net_updates", "orc_src_apply", "orc_local_lts",
                  "orc_neigh_lts"]:
            getattr(_lib, f).restype = None
    return _lib


def ref_lib():
    """The reference's Elasticity.cpp (or None when it was not built)."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        _ref = C.CDLL(REF_LIB)
        for f in ["ref_star", "ref_trafo", "ref_fs", "ref_ms_jump_l", "ref_ms_jump_r", "ref_ms_flux"]:
            getattr(_ref, f).restype = None
    return _ref


def _p(a):
    """Pointer to an array's data that keeps the array alive as long as the
    pointer object (arguments built from temporaries stay valid for the call)."""
    if a is None:
        return None
    v = C.c_void_p(a.ctypes.data)
    v._keep = a
    return v


def _c(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class Dg:
    """Holds the matrices (FP64) and the C struct handed to the kernels."""

    def __init__(self, order, n_crs=1, mats=None, rfs=None):
        if mats is None:
            from edge_amd.setups import dg_matrices
            mats = dg_matrices(order)
        self.keep = {k: _c(v) for k, v in mats.items()}
        self.rfs = _c(rfs) if rfs is not None else None
        n_rms = 0 if rfs is None else len(rfs)
        self.s = OrcDg(order, n_crs, n_rms, *[_p(self.keep.get(k)) for k in
                                              ["stiffT", "stiffV", "fluxL", "fluxN", "fluxT"]], _p(self.rfs))
        self.order, self.n_crs, self.n_rms = order, n_crs, n_rms
        self.n = order * (order + 1) * (order + 2) // 6
        self.nf = order * (order + 1) // 2


# ---------------------------------------------------------------- kernels
def ck(dg, dt, star, dofs, star_a=None, src_a=None, dofs_a=None):
    """Returns (der [O][9][N][R], tInt [9][N][R], derA, tIntA)."""
    O, N, R, NR = dg.order, dg.n, dg.n_crs, dg.n_rms
    der = np.zeros((O, 9, N, R))
    tint = np.zeros((9, N, R))
    derA = np.zeros((max(NR, 1), O, 6, N, R))
    tintA = np.zeros((max(NR, 1), 6, N, R))
    star, dofs, star_a, src_a, dofs_a = map(_c, (star, dofs, star_a, src_a, dofs_a))
    lib().orc_ck(C.byref(dg.s), C.c_double(dt), _p(star), _p(star_a), _p(src_a), _p(dofs), _p(dofs_a), _p(der),
                 _p(derA), _p(tint), _p(tintA))
    return der, tint, derA[:NR], tintA[:NR]


def vol(dg, star, tint, dofs, star_a=None, src_a=None, tint_a=None, dofs_a=None):
    dofs = _c(dofs).copy()
    dofs_a = None if dofs_a is None else _c(dofs_a).copy()
    star, tint, star_a, src_a, tint_a = map(_c, (star, tint, star_a, src_a, tint_a))
    lib().orc_vol(C.byref(dg.s), _p(star), _p(star_a), _p(src_a), _p(tint), _p(tint_a), _p(dofs), _p(dofs_a))
    return dofs, dofs_a


def surf_local(dg, fs, tint, dofs, fs_a=None, dofs_a=None):
    dofs = _c(dofs).copy()
    dofs_a = None if dofs_a is None else _c(dofs_a).copy()
    fs, tint, fs_a = map(_c, (fs, tint, fs_a))
    lib().orc_surf_local(C.byref(dg.s), _p(fs), _p(fs_a), _p(tint), _p(dofs), _p(dofs_a))
    return dofs, dofs_a


def surf_neigh(dg, fa, v_id, f_id, fs, tint, dofs, fs_a=None):
    """Returns (dofs updated, unscaled anelastic update)."""
    dofs = _c(dofs).copy()
    up = np.zeros((6, dg.n, dg.n_crs))
    fs, tint, fs_a = map(_c, (fs, tint, fs_a))
    lib().orc_surf_neigh(C.byref(dg.s), C.c_int(fa), C.c_uint(v_id), C.c_uint(f_id), _p(fs), _p(fs_a), _p(tint),
                         _p(dofs), _p(up))
    return dofs, up


# ---------------------------------------------------------------- set-up
def jac(ve):
    """ve [4][3] -> (jac [3][3], det, jinv [3][3])."""
    j = np.zeros(9)
    ji = np.zeros(9)
    d = C.c_double()
    lib().orc_jac(_p(_c(ve)), _p(j), C.byref(d), _p(ji))
    return j.reshape(3, 3), d.value, ji.reshape(3, 3)


def star(rho, lam, mu, jinv):
    out = np.zeros(243)
    lib().orc_star(C.c_double(rho), C.c_double(lam), C.c_double(mu), _p(_c(jinv)), _p(out))
    return out.reshape(3, 81)


def trafo(n, s, t):
    T = np.zeros(81)
    Ti = np.zeros(81)
    lib().orc_trafo(_p(_c(n)), _p(_c(s)), _p(_c(t)), _p(T), _p(Ti))
    return T.reshape(9, 9), Ti.reshape(9, 9)


def fs_face(rL, rR, lL, lR, mL, mR, n, s, t, free_surface=False):
    a = np.zeros(81)
    b = np.zeros(81)
    d = C.c_double
    lib().orc_fs(d(rL), d(rR), d(lL), d(lR), d(mL), d(mR), _p(_c(n)), _p(_c(s)), _p(_c(t)),
                 C.c_int(1 if free_surface else 0), _p(a), _p(b))
    return a.reshape(9, 9), b.reshape(9, 9)


def setup_mesh(mesh, rho, lam, mu):
    """(star [nEl][3][81], fs [2][nEl][4][81], jdet [nEl]) built on the CPU."""
    ne = mesh.n_el
    ve = mesh.el_crds()
    d = ve[:, 1:, :] - ve[:, :1, :]
    jdet = np.linalg.det(d)
    st = np.zeros((ne, 3, 81))
    for e in range(ne):
        _, _, ji = jac(ve[e])
        st[e] = star(rho[e], lam[e], mu[e], ji)
    fs = np.full((2, ne, 4, 81), np.nan)
    lib().orc_init_fs(C.c_int64(mesh.n_fa), _p(np.ascontiguousarray(mesh.fa_el)),
                      _p(np.ascontiguousarray(mesh.el_fa)), _p(np.ascontiguousarray(mesh.fa_type)),
                      _p(_c(mesh.normals)), _p(_c(mesh.tangents)), _p(_c(mesh.areas)), _p(_c(jdet)),
                      _p(_c(rho)), _p(_c(lam)), _p(_c(mu)), _p(fs[0]), _p(fs[1]))
    return st, fs, jdet


# ---------------------------------------------------------------- viscoelastic set-up
_d = C.c_double


def visco_frequencies(n_mechs, f_cen, f_rat):
    out = np.zeros(2 * n_mechs - 1)
    lib().orc_visco_frequencies(C.c_int(n_mechs), _d(f_cen), _d(f_rat), _p(out))
    return out


def anelastic_coeff_ps(n_mechs, q, freqs):
    out = np.zeros(n_mechs)
    fr = _c(freqs)  # held for the call
    lib().orc_anelastic_coeff_ps(C.c_int(n_mechs), _d(q), _p(fr), _p(out))
    return out


def elastic_lame(n_mechs, f_ref, freqs, c_kappa, c_mu, lam, mu):
    le, me = _d(), _d()
    fr, ck, cm = _c(freqs), _c(c_kappa), _c(c_mu)  # held for the call
    lib().orc_elastic_lame(C.c_int(n_mechs), _d(f_ref), _p(fr), _p(ck), _p(cm), _d(lam), _d(mu), C.byref(le),
                           C.byref(me))
    return le.value, me.value


def anelastic_coeff_lame(n_mechs, f_cen, f_rat, qp, qs, lam, mu):
    le, me = _d(), _d()
    co = np.zeros(2 * n_mechs)
    lib().orc_anelastic_coeff_lame(C.c_int(n_mechs), _d(f_cen), _d(f_rat), _d(qp), _d(qs), _d(lam), _d(mu),
                                   C.byref(le), C.byref(me), _p(co))
    return le.value, me.value, co


def src_a(n_mechs, f_cen, f_rat, qp, qs, lam, mu):
    """(lam_elastic, mu_elastic, src [n][6][6])."""
    le, me = _d(), _d()
    out = np.zeros((n_mechs, 36))
    lib().orc_src_a(C.c_int(n_mechs), _d(f_cen), _d(f_rat), _d(qp), _d(qs), _d(lam), _d(mu), C.byref(le),
                    C.byref(me), _p(out))
    return le.value, me.value, out.reshape(n_mechs, 6, 6)


def star_a(jinv):
    out = np.zeros(54)
    ji = _c(np.asarray(jinv, dtype=np.float64).ravel())  # held for the call
    lib().orc_star_a(_p(ji), _p(out))
    return out.reshape(3, 6, 3)


def fs_mid_a(rL, rR, lL, lR, mL, mR):
    a, b = np.zeros(54), np.zeros(54)
    lib().orc_fs_mid_a(*[_d(x) for x in (rL, rR, lL, lR, mL, mR)], _p(a), _p(b))
    return a.reshape(6, 9), b.reshape(6, 9)


def fs_a(rL, rR, lL, lR, mL, mR, n, s, t, free_surface=False):
    a, b = np.zeros(54), np.zeros(54)
    n, s, t = _c(n), _c(s), _c(t)  # held for the call
    lib().orc_fs_a(*[_d(x) for x in (rL, rR, lL, lR, mL, mR)], _p(n), _p(s), _p(t), C.c_int(int(free_surface)),
                   _p(a), _p(b))
    return a.reshape(6, 9), b.reshape(6, 9)


def setup_mesh_visco(mesh, rho, lam, mu, qp, qs, n_rms, f_cen, f_rat):
    """Viscoelastic set-up of a mesh (AderDgInit.hpp:91-117, :129-174,
    :394-665 with AderDg.hpp:248-258): per element the anelastic source
    matrices and the unrelaxed Lame parameters, which then build the elastic
    star matrices and flux solvers; anelastic star matrices and flux solvers;
    relaxation frequencies (every second of the 2n-1 frequencies)."""
    ne = mesh.n_el
    lam_e, mu_e = np.zeros(ne), np.zeros(ne)
    src = np.zeros((ne, n_rms, 36))
    for e in range(ne):
        lam_e[e], mu_e[e], sm = src_a(n_rms, f_cen, f_rat, qp[e], qs[e], lam[e], mu[e])
        src[e] = sm.reshape(n_rms, 36)
    starE, fsE, jdet = setup_mesh(mesh, rho, lam_e, mu_e)
    ve = mesh.el_crds()
    starA = np.zeros((ne, 3, 18))
    for e in range(ne):
        _, _, ji = jac(ve[e])
        starA[e] = star_a(ji).reshape(3, 18)
    fsA = np.full((2, ne, 4, 54), np.nan)
    lib().orc_init_fs_a(C.c_int64(mesh.n_fa), _p(np.ascontiguousarray(mesh.fa_el)),
                        _p(np.ascontiguousarray(mesh.el_fa)), _p(np.ascontiguousarray(mesh.fa_type)),
                        _p(_c(mesh.normals)), _p(_c(mesh.tangents)), _p(_c(mesh.areas)), _p(_c(jdet)),
                        _p(_c(rho)), _p(_c(lam_e)), _p(_c(mu_e)), _p(fsA[0]), _p(fsA[1]))
    rfs = visco_frequencies(n_rms, f_cen, f_rat)[::2].copy()
    return {"starE": starE, "fsE": fsE, "starA": starA, "srcA": src, "fsA": fsA, "rfs": rfs,
            "lam_e": lam_e, "mu_e": mu_e, "jdet": jdet}


# ---------------------------------------------------------------- rupture
def ms_jump(lamL, lamR, muL, muR, rhoL, rhoR, side="l"):
    out = np.zeros(81)
    f = lib().orc_ms_jump_l if side == "l" else lib().orc_ms_jump_r
    f(*[_d(x) for x in (lamL, lamR, muL, muR, rhoL, rhoR)], _p(out))
    return out.reshape(9, 9)


def ms_flux(n, s, t, lam, mu, rho):
    out = np.zeros(81)
    lib().orc_ms_flux(_p(_c(n)), _p(_c(s)), _p(_c(t)), _d(lam), _d(mu), _d(rho), _p(out))
    return out.reshape(9, 9)


def rot_mat(v0, v1):
    out = np.zeros(9)
    lib().orc_rot_mat(_p(_c(v0)), _p(_c(v1)), _p(out))
    return out.reshape(3, 3)


def ib_solvers(normal, t0, t1, bnd, matL, matR, area, volL, volR, n_sf, n_sc):
    """(solvers [4][9][9]: T^-1, jump, flux left, flux right; preferred direction)."""
    lib().orc_ib_solvers.restype = C.c_int
    out = np.zeros(4 * 81)
    pref = lib().orc_ib_solvers(_p(_c(normal)), _p(_c(t0)), _p(_c(t1)),
                                None if bnd is None else _p(_c(np.asarray(bnd).ravel())), *[_d(x) for x in matL],
                                *[_d(x) for x in matR], _d(area), _d(volL), _d(volR), C.c_int(n_sf), C.c_int(n_sc),
                                _p(out))
    return out.reshape(4, 9, 9), bool(pref)


LSW_STATE = ["sn0", "co0", "ss0_0", "ss0_1", "ss0A", "muf", "st", "tr_0", "tr_1", "sr_0", "sr_1", "dd_0", "dd_1"]


def lsw_state(n, **vals):
    """Friction state [n][14] (columns: sn0, co0, ss0[2], ss0A, muf, st, tr[2], sr[2], dd[2])."""
    st = np.zeros((n, 14))
    cols = {"sn0": 0, "co0": 1, "ss0_0": 2, "ss0_1": 3, "ss0A": 4, "muf": 5, "st": 6, "tr_0": 7, "tr_1": 8,
            "sr_0": 9, "sr_1": 10, "dd_0": 11, "dd_1": 12}
    for k, v in vals.items():
        st[:, cols[k]] = v
    return st


def net_updates(solv, dofsL, dofsR, dt, l_eq_m, cs_dmu_m, cs_dmu_p, glob, state):
    """One rupture face: dofs [9][n_sf][R]; glob [3][R] (mus, mud, 1/Dc);
    state [n_sf*R][14] (updated in place).  Returns (netL, netR, per)."""
    n_sf, R = dofsL.shape[1], dofsL.shape[2]
    netL, netR = np.zeros((9, n_sf, R)), np.zeros((9, n_sf, R))
    per = np.zeros((n_sf, R), dtype=np.int32)
    solv, dofsL, dofsR, glob = _c(solv), _c(dofsL), _c(dofsR), _c(glob)
    assert state.dtype == np.float64 and state.flags.c_contiguous
    lib().orc_net_updates(C.c_int(n_sf), C.c_int(R), _p(solv), _p(dofsL), _p(dofsR), _d(dt), C.c_int(int(l_eq_m)),
                          _d(cs_dmu_m), _d(cs_dmu_p), _p(glob), _p(state), _p(netL), _p(netR), _p(per))
    return netL, netR, per.astype(bool)


# ---------------------------------------------------------------- sources
def series_integrate(dx, x0, vals, x1, x2, na=0.0):
    v = _c(vals)
    f = lib().orc_series_integrate
    f.restype = C.c_double
    return f(C.c_double(dx), C.c_double(x0), C.c_long(len(v)), _p(v), C.c_double(x1), C.c_double(x2),
             C.c_double(na))


def src_apply(dofs, el, run, t0, dts, bev, sca, series, t1, t2):
    """dofs [nEl][9][N][R] updated in place; series: list of sampled arrays."""
    assert dofs.dtype == np.float64 and dofs.flags.c_contiguous
    n = len(el)
    ptr = np.zeros(n + 1, dtype=np.int64)
    ptr[1:] = np.cumsum([len(x) for x in series])
    ts = np.concatenate([np.asarray(x, dtype=np.float64) for x in series])
    lib().orc_src_apply(C.c_int(dofs.shape[2]), C.c_int(dofs.shape[3]), C.c_long(n),
                        _p(np.asarray(el, dtype=np.int64)), _p(np.asarray(run, dtype=np.int32)), _p(_c(t0)),
                        _p(_c(dts)), _p(_c(bev)), _p(_c(sca)), _p(ptr), _p(ts), C.c_double(t1), C.c_double(t2),
                        _p(dofs))


def integrate(dg, dt, der):
    """Taylor integral over [0, dt] of der [O][9][N][R] (TimePred::integrate)."""
    out = np.zeros(der.shape[1:])
    lib().orc_integrate(C.byref(dg.s), C.c_double(dt), _p(_c(der)), _p(out))
    return out


def eval_time_pred(dg, t, der):
    """Taylor evaluation at relative time t of der [O][9][N][R] (evalTimePrediction)."""
    out = np.zeros(der.shape[1:])
    lib().orc_eval_time_pred(C.byref(dg.s), C.c_double(t), _p(_c(der)), _p(out))
    return out


def dt_cfl(rho, lam, mu, in_dia, order):
    return lib().orc_dt_cfl(rho, lam, mu, in_dia, order)


# ---------------------------------------------------------------- solver loops
class MeshOracle:
    """FP64 CPU time stepping of a whole mesh (AderDg::local + neigh, GTS)."""

    def __init__(self, mesh, order, n_crs=1, star=None, fs=None, rho=None, lam=None, mu=None, visco=None):
        """visco: the dict of setup_mesh_visco (anelastic terms, N_RMS > 0)."""
        self.mesh = mesh
        self.visco = visco
        if visco is not None:
            star, fs = visco["starE"], visco["fsE"]
            self.star_a, self.src_a, self.fs_a = _c(visco["starA"]), _c(visco["srcA"]), _c(visco["fsA"])
        else:
            self.star_a = self.src_a = None
            self.fs_a = [None, None]
        self.dg = Dg(order, n_crs, rfs=None if visco is None else visco["rfs"])
        if star is None:
            star, fs, _ = setup_mesh(mesh, rho, lam, mu)
        self.star = _c(star)
        self.fs = _c(fs)
        self.el_fa = np.ascontiguousarray(mesh.el_fa, dtype=np.int64)
        self.el_fa_el = np.ascontiguousarray(mesh.el_fa_el, dtype=np.int64)
        self.f_id = np.ascontiguousarray(mesh.f_id, dtype=np.uint16)
        self.v_id = np.ascontiguousarray(mesh.v_id, dtype=np.uint16)
        self.fa_type = np.ascontiguousarray(mesh.fa_type, dtype=np.int32)

    def local(self, dofs, tint, dt, first=0, n=None, dofs_a=None):
        """dofs [el][9][N][R], tint [el][9][N][R]; dofs_a [el][N_RMS][6][N][R]
        with anelastic terms."""
        n = self.mesh.n_el - first if n is None else n
        lib().orc_local(C.byref(self.dg.s), C.c_int64(first), C.c_int64(n), C.c_double(dt), _p(self.star),
                        _p(self.star_a), _p(self.src_a), _p(self.fs[0]), _p(self.fs_a[0]), _p(dofs), _p(dofs_a),
                        _p(tint))

    def neigh(self, dofs, tint, first=0, n=None, dofs_a=None):
        n = self.mesh.n_el - first if n is None else n
        lib().orc_neigh(C.byref(self.dg.s), C.c_int64(first), C.c_int64(n), _p(self.fa_type), _p(self.el_fa),
                        _p(self.el_fa_el), _p(self.f_id), _p(self.v_id), _p(self.fs[1]), _p(self.fs_a[1]),
                        _p(tint), _p(dofs), _p(dofs_a))

    def step(self, dofs, tint, dt, n_steps=1, dofs_a=None):
        for _ in range(n_steps):
            self.local(dofs, tint, dt, dofs_a=dofs_a)
            self.neigh(dofs, tint, dofs_a=dofs_a)

    # ---- local time stepping (rate 2, elastic)
    def set_lts(self, el_lts, fa_lts):
        self.el_lts = np.ascontiguousarray(el_lts, dtype=np.uint8)
        self.fa_lts = np.ascontiguousarray(fa_lts, dtype=np.uint8)

    def local_lts(self, dofs, t0, t1, t2, dt, first_ts, first=0, n=None, dofs_a=None):
        n = self.mesh.n_el - first if n is None else n
        lib().orc_local_lts(C.byref(self.dg.s), C.c_int64(first), C.c_int64(n), C.c_double(dt),
                            C.c_int(int(first_ts)), _p(self.star), _p(self.star_a), _p(self.src_a), _p(self.fs[0]),
                            _p(self.fs_a[0]), _p(self.el_lts), _p(dofs), _p(dofs_a), _p(t0), _p(t1), _p(t2))

    def neigh_lts(self, dofs, t0, t1, t2, first_ts, first=0, n=None, dofs_a=None):
        n = self.mesh.n_el - first if n is None else n
        lib().orc_neigh_lts(C.byref(self.dg.s), C.c_int64(first), C.c_int64(n), C.c_int(int(first_ts)),
                            _p(self.fa_type), _p(self.el_fa), _p(self.el_fa_el), _p(self.f_id), _p(self.v_id),
                            _p(self.fs[1]), _p(self.fs_a[1]), _p(self.fa_lts), _p(t0), _p(t1), _p(t2), _p(dofs),
                            _p(dofs_a))


def lts_groups(mesh, group):
    """Rate-2 LTS characteristics from a time group per element (group g
    steps with 2^g times the fundamental step): el_lts [nEl] (bit 0: an
    adjacent element has a larger step, bit 1: a smaller one) and fa_lts
    [nEl][4] (0 equal, 1 the element's step is the greater, 2 the smaller).
    Outflow faces and free surfaces count as equal."""
    group = np.asarray(group)
    ne = np.asarray(mesh.el_fa_el)
    ty = np.asarray(mesh.fa_type)[np.asarray(mesh.el_fa)]
    own = (ty == 101) | (ty == 105) | (ne < 0)
    gn = np.where(own, group[:, None], group[np.where(ne < 0, 0, ne)])
    g = group[:, None]
    fa = np.where(gn == g, 0, np.where(g > gn, 1, 2)).astype(np.uint8)
    el = ((fa == 2).any(axis=1) * 1 + (fa == 1).any(axis=1) * 2).astype(np.uint8)
    return el, fa


# ---------------------------------------------------------------- CPU baseline sample
def setup_mesh_sample(mesh, rho, lam, mu, n_sample):
    """Star matrices and flux solvers of the first n_sample elements of a
    mesh (flux solvers built over all faces, then sliced)."""
    n_s = min(n_sample, mesh.n_el)
    ve = mesh.el_crds()[:n_s]
    st = np.zeros((n_s, 3, 81))
    for e in range(n_s):
        _, _, ji = jac(ve[e])
        st[e] = star(rho[e], lam[e], mu[e], ji)
    d = mesh.el_crds()
    jdet = np.linalg.det(d[:, 1:, :] - d[:, :1, :])
    fs = np.zeros((2, mesh.n_el, 4, 81))
    lib().orc_init_fs(C.c_int64(mesh.n_fa), _p(np.ascontiguousarray(mesh.fa_el)),
                      _p(np.ascontiguousarray(mesh.el_fa)), _p(np.ascontiguousarray(mesh.fa_type)),
                      _p(_c(mesh.normals)), _p(_c(mesh.tangents)), _p(_c(mesh.areas)), _p(_c(jdet)),
                      _p(_c(rho)), _p(_c(lam)), _p(_c(mu)), _p(fs[0]), _p(fs[1]))
    sample = {"el_fa_el": np.where(mesh.el_fa_el[:n_s] >= 0, mesh.el_fa_el[:n_s] % n_s, -1),
              "fa_type": mesh.fa_type[mesh.el_fa[:n_s]].ravel().astype(np.int32),
              "f_id": mesh.f_id[:n_s], "v_id": mesh.v_id[:n_s]}
    fs_s = np.ascontiguousarray(fs[:, :n_s])
    return st, fs_s, sample


class SampleOracle(MeshOracle):
    """The oracle's loops over a sample of a mesh; neighbours outside the
    sample are folded back into it (same work per element)."""

    def __init__(self, order, n_crs, star, fs, sample):
        self.dg = Dg(order, n_crs)
        self.star = _c(star)
        self.fs = _c(fs)
        self.visco = self.star_a = self.src_a = None
        self.fs_a = [None, None]
        n_s = star.shape[0]

        class _M:
            n_el = n_s
        self.mesh = _M()
        self.el_fa = np.arange(4 * n_s, dtype=np.int64).reshape(n_s, 4)
        self.el_fa_el = np.ascontiguousarray(sample["el_fa_el"], dtype=np.int64)
        self.f_id = np.ascontiguousarray(sample["f_id"], dtype=np.uint16)
        self.v_id = np.ascontiguousarray(sample["v_id"], dtype=np.uint16)
        self.fa_type = np.ascontiguousarray(sample["fa_type"], dtype=np.int32)


# ---------------------------------------------------------------- a fault in the time step
# Restated conventions (tools/edge_pre/edge_pre/types/Tet.py:50-91): the face
# maps chi -> xi of the reference tet's faces (0: 0-2-1, 1: 0-1-3, 2: 0-3-2,
# 3: 1-2-3) and the neighbour face's coordinates as a function of the local
# face's, per position v of local face-vertex 0 in the neighbour's face.
def _fa_to_el(fa, c1, c2):
    z = np.zeros_like(c1)
    return np.stack([(c2, c1, z), (c1, z, c2), (z, c2, c1), (1 - c1 - c2, c1, c2)][fa], axis=-1)


def _fa_to_fa(v, c1, c2):
    return [(c2, c1), (1 - c1 - c2, c2), (c1, 1 - c1 - c2)][v]


def fault_sub_faces(order):
    """Vertices [n_sf][3][2] of the face's sub-triangles, n = 2 O - 1 per edge,
    ordered by rows along chi_1, up then down triangle per column
    (tools/edge_pre/edge_pre/sc/grid/Tet.py:750-777)."""
    n = 2 * order - 1
    out = []
    for e2 in range(n):
        for e1 in range(n):
            if e1 < n - e2:
                out.append([(e1, e2), (e1 + 1, e2), (e1, e2 + 1)])
            if e1 < n - e2 - 1:
                out.append([(e1 + 1, e2), (e1 + 1, e2 + 1), (e1, e2 + 1)])
    return np.array(out, dtype=np.float64) / n


def _tria_rule(npt):
    """Collapsed Gauss-Legendre rule on the unit triangle (weights sum to 1/2)."""
    x, w = np.polynomial.legendre.leggauss(npt)
    x, w = 0.5 * (x + 1), 0.5 * w
    a, b = np.meshgrid(x, x, indexing="ij")
    wa, wb = np.meshgrid(w, w, indexing="ij")
    return np.stack([a.ravel(), (b * (1 - a)).ravel()], axis=1), (wa * wb * (1 - a)).ravel()


def fault_tables(order):
    """(scat [16][N][n_sf], lift [16][n_sf][N]) by quadrature: the sub-face
    means of each table's trace (t < 4: face t; t = 4 + 4 v + f: neighbour
    face f seen with orientation v), and the lift of one sub-face's net
    update: the least-squares face polynomial (monomial basis) of the
    sub-face values, integrated against the trace, over the diagonal mass,
    times n_sf / (3 n_sc), n_sc = (2 O - 1)^3."""
    from edge_amd.setups import dg_matrices, eval_basis
    N = order * (order + 1) * (order + 2) // 6
    tris = fault_sub_faces(order)
    nsf = len(tris)
    nsc = (2 * order - 1) ** 3
    mass = dg_matrices(order)["mass"]
    qp, qw = _tria_rule(order + 2)
    mono = [(a, b) for a in range(order) for b in range(order - a)]
    # sub-face means of the monomials, least-squares inverse
    smono = np.zeros((nsf, len(mono)))
    sub_pts = []
    for sf, (a, b, c) in enumerate(tris):
        x = a + qp[:, :1] * (b - a) + qp[:, 1:] * (c - a)
        sub_pts.append(x)
        for j, (p, q) in enumerate(mono):
            smono[sf, j] = 2 * np.sum(qw * x[:, 0] ** p * x[:, 1] ** q)
    g = np.linalg.pinv(smono)
    scat = np.zeros((16, N, nsf))
    lift = np.zeros((16, nsf, N))
    for t in range(16):
        def to_el(c):
            if t < 4:
                return _fa_to_el(t, c[:, 0], c[:, 1])
            v, f = (t - 4) // 4, (t - 4) % 4
            return _fa_to_el(f, *_fa_to_fa(v, c[:, 0], c[:, 1]))
        for sf in range(nsf):
            scat[t, :, sf] = 2 * qw @ eval_basis(order, to_el(sub_pts[sf]))
        phi = eval_basis(order, to_el(qp))  # [q][N] on the whole face
        m = np.stack([qp[:, 0] ** p * qp[:, 1] ** q for p, q in mono], axis=1)
        inner = (phi * qw[:, None]).T @ m  # [N][mono]
        lift[t] = (inner @ g).T / mass[None, :] * (nsf / (3.0 * nsc))
    return scat, lift


class FaultOracle(MeshOracle):
    """The coupled step with a fault (faces of type 201): local, the fault
    phase, neighbour.  Fault faces' Godunov solvers are zeroed; the fault
    phase takes both sides' sub-face means of tInt / dt, the net updates of
    the pinned orc_net_updates, and lifts them back into both elements."""

    def __init__(self, mesh, order, n_crs, rho, lam, mu, fault_frame, glob, state0):
        super().__init__(mesh, order, n_crs, rho=rho, lam=lam, mu=mu)
        ft = np.asarray(mesh.fa_type)
        self.faces = np.nonzero((ft & 201) == 201)[0]
        self.scat, self.lift = fault_tables(order)
        self.n_sf = self.scat.shape[2]
        n_sc = (2 * order - 1) ** 3
        ve = mesh.el_crds()
        vol = np.abs(np.linalg.det(ve[:, 1:, :] - ve[:, :1, :])) / 6.0
        self.sides, self.solv, self.face = [], [], []
        bnd = np.asarray(fault_frame, dtype=np.float64).reshape(3, 3)
        mat = np.stack([rho, lam, mu], axis=1)
        for f in self.faces:
            eL, eR = int(mesh.fa_el[f][0]), int(mesh.fa_el[f][1])
            fl = int(np.nonzero(np.asarray(mesh.el_fa[eL]) == f)[0][0])
            tr = 4 + 4 * int(mesh.v_id[eL][fl]) + int(mesh.f_id[eL][fl])
            fr = int(np.nonzero(np.asarray(mesh.el_fa[eR]) == f)[0][0])
            for e, k in ((eL, fl), (eR, fr)):
                self.fs[0, e, k] = 0.0
                self.fs[1, e, k] = 0.0
            s, pref = ib_solvers(mesh.normals[f], mesh.tangents[f][0], mesh.tangents[f][1], bnd, mat[eL], mat[eR],
                                 mesh.areas[f], vol[eL], vol[eR], self.n_sf, n_sc)
            l_eq_m = not pref  # RuptureInit.hpp:198-206
            cdl, cdr = np.sqrt(1.0 / (mu[eL] * rho[eL])), np.sqrt(1.0 / (mu[eR] * rho[eR]))
            self.sides.append((eL, eR, fl, tr))
            self.solv.append(s)
            self.face.append((l_eq_m, cdl if l_eq_m else cdr, cdr if l_eq_m else cdl))
        self.glob = _c(glob)
        self.state = np.ascontiguousarray(np.asarray(state0, dtype=np.float64).reshape(len(self.faces), -1, 14))
        self.per = np.zeros((len(self.faces), self.n_sf, n_crs), dtype=bool)

    def fault(self, dofs, tint, dt):
        for i, (eL, eR, fl, tr) in enumerate(self.sides):
            qL = np.einsum("qnr,ns->qsr", tint[eL], self.scat[fl]) / dt
            qR = np.einsum("qnr,ns->qsr", tint[eR], self.scat[tr]) / dt
            l_eq_m, cm, cp = self.face[i]
            nl, nr, per = net_updates(self.solv[i], qL, qR, dt, l_eq_m, cm, cp, self.glob, self.state[i])
            self.per[i] = per
            dofs[eL] += np.einsum("qsr,sn->qnr", nl, self.lift[fl])
            dofs[eR] += np.einsum("qsr,sn->qnr", nr, self.lift[tr])

    def step(self, dofs, tint, dt, n_steps=1, dofs_a=None):
        for _ in range(n_steps):
            self.local(dofs, tint, dt)
            self.fault(dofs, tint, dt)
            self.neigh(dofs, tint)
