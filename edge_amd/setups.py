"""Set-ups around the element update: reference-element data, materials,
the plane-wave initial condition, L2 projection onto the modal basis, error
norms against analytic solutions and the CFL time step.

Follows src/setups/InitialDofs.hpp:172-370 (projection with (O+1)^3
collapsed Gauss points and error norms), src/impl/seismic/common.hpp:84-98
(CFL time step with scale 1/2) and the plane-wave set-ups of the reference's
CI (P wave q = r_P sin(k.x - w t)).
"""
import ctypes as C

import numpy as np

from . import _lib

N_QTS = 9


def n_modes(order):
    return order * (order + 1) * (order + 2) // 6


def n_face_modes(order):
    return order * (order + 1) // 2


def dg_matrices(order):
    """Reference-element matrices of order O as numpy arrays."""
    n, nf = n_modes(order), n_face_modes(order)
    out = dict(mass=np.zeros(n), stiffT=np.zeros((3, n, n)), stiffV=np.zeros((3, n, n)),
               fluxL=np.zeros((4, n, nf)), fluxN=np.zeros((12, n, nf)), fluxT=np.zeros((4, nf, n)))
    _lib.check(_lib.lib().edge_dg_matrices(order, *[out[k].ctypes.data for k in
                                                    ["mass", "stiffT", "stiffV", "fluxL", "fluxN", "fluxT"]]))
    return out


def quad_tet(n1d):
    """Collapsed Gauss-Legendre rule on the unit tet: pts [n^3][3], w [n^3]."""
    pts = np.zeros((n1d ** 3, 3))
    w = np.zeros(n1d ** 3)
    _lib.check(_lib.lib().edge_dg_quad_tet(n1d, pts.ctypes.data, w.ctypes.data))
    return pts, w


def eval_basis(order, pts):
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    val = np.zeros((pts.shape[0], n_modes(order)))
    _lib.check(_lib.lib().edge_dg_eval_basis(order, pts.shape[0], pts.ctypes.data, val.ctypes.data))
    return val


# --------------------------------------------------------------------------
# materials
# --------------------------------------------------------------------------

def lame_from_velocities(rho, vp, vs):
    mu = rho * vs * vs
    lam = rho * vp * vp - 2.0 * mu
    return lam, mu


def random_materials(n_el, seed=3343, rho=(2.0, 3.0), vp=(4.0, 6.0)):
    """Per-element rho ~ U[rho], vp ~ U[vp], vs = vp / sqrt(3)."""
    rng = np.random.default_rng(seed)
    r = rng.uniform(rho[0], rho[1], n_el)
    p = rng.uniform(vp[0], vp[1], n_el)
    lam, mu = lame_from_velocities(r, p, p / np.sqrt(3.0))
    return r, lam, mu


def dt_cfl(mesh, rho, lam, mu, order):
    """Global time step: min over elements of inDia * 1/2 / c_p / (2 O - 1)."""
    cp = np.sqrt((lam + 2 * mu) / rho)
    return float(np.min(mesh.in_dia * 0.5 / cp / (2.0 * order - 1.0)))


# --------------------------------------------------------------------------
# plane waves
# --------------------------------------------------------------------------

class PlaneWave:
    """P plane wave q(x, t) = amp * r_P sin(k.x - w t) in a homogeneous
    material (rho, lam, mu); quantities (sxx, syy, szz, sxy, syz, sxz, u, v, w).

    r_P is the eigenvector of the normal Jacobian A_n (n = k / |k|) for the
    eigenvalue c_p: velocity n, stress -(lam I + 2 mu n n^T) / c_p."""

    def __init__(self, rho=1.0, lam=2.0, mu=1.0, k=(2 * np.pi, 2 * np.pi, 2 * np.pi), amp=1.0):
        k = np.asarray(k, dtype=np.float64)
        n = k / np.linalg.norm(k)
        cp = np.sqrt((lam + 2 * mu) / rho)
        s = -(lam * np.eye(3) + 2 * mu * np.outer(n, n)) / cp
        self.r = amp * np.array([s[0, 0], s[1, 1], s[2, 2], s[0, 1], s[1, 2], s[0, 2], n[0], n[1], n[2]])
        self.k = k
        self.omega = cp * np.linalg.norm(k)

    def __call__(self, x, t=0.0):
        """x [..., 3] -> q [..., 9]."""
        ph = np.sin(x @ self.k - self.omega * t)
        return ph[..., None] * self.r

    def polarisations(self, rho, lam, mu):
        """r_P of this wave's direction in each of the materials (rho, lam,
        mu) [nEl], scaled like self.r (velocity amplitude amp): [nEl][9]."""
        n = self.k / np.linalg.norm(self.k)
        amp = np.linalg.norm(self.r[6:])
        lam, mu = np.asarray(lam, np.float64), np.asarray(mu, np.float64)
        cp = np.sqrt((lam + 2 * mu) / np.asarray(rho, np.float64))
        s = -(lam[:, None, None] * np.eye(3) + 2 * mu[:, None, None] * np.outer(n, n)) / cp[:, None, None]
        r = np.stack([s[:, 0, 0], s[:, 1, 1], s[:, 2, 2], s[:, 0, 1], s[:, 1, 2], s[:, 0, 2]] +
                     [np.full(len(cp), n[i]) for i in range(3)], axis=1)
        return amp * r


def map_points(mesh, ref_pts, els=None):
    """Physical coordinates of reference points per element: [nEl][nPts][3]."""
    ve = mesh.el_crds() if els is None else mesh.ve_crds[mesh.el_ve[els]]
    d = ve[:, 1:, :] - ve[:, :1, :]          # [e][i][j] = (v_{i+1} - v0)_j
    return ve[:, None, 0, :] + np.matmul(ref_pts[None], d)


def project(mesh, order, fun, n_runs=1, chunk=8192, materials=None):
    """L2 projection of fun(x [..., 3]) -> [..., 9] (or [..., 9, R]) onto the
    modal basis with (O+1)^3 collapsed Gauss points: dofs [nEl][9][N][R]
    (src/setups/InitialDofs.hpp:172-251, src/dg/Basis.cpp:136-161).
    materials = (rho, lam, mu) per element, for a PlaneWave only: each
    element takes the wave's P polarisation in its own material (a P wave
    of the element's medium) instead of the wave's own."""
    pts, w = quad_tet(order + 1)
    b = eval_basis(order, pts)                 # [qp][N]
    mass = dg_matrices(order)["mass"]
    proj = (b * w[:, None]) / mass[None, :]    # [qp][N]
    n = n_modes(order)
    out = np.zeros((mesh.n_el, N_QTS, n, n_runs))
    wave = isinstance(fun, PlaneWave)
    if materials is not None and not wave:
        raise ValueError("per-element polarisations need a PlaneWave")
    r_el = fun.polarisations(*materials) if materials is not None else None
    for e0 in range(0, mesh.n_el, chunk):
        els = np.arange(e0, min(mesh.n_el, e0 + chunk))
        x = map_points(mesh, pts, els)
        if wave:  # q = r sin(k.x): project the scalar phase once, one matrix product
            coef = np.sin(x @ fun.k) @ proj  # [e][N]
            r = fun.r[None, :] if r_el is None else r_el[els]
            out[els] = coef[:, None, :, None] * r[:, :, None, None]
            continue
        f = fun(x)
        if f.ndim == 3:
            f = np.repeat(f[..., None], n_runs, axis=-1)
        out[els] = np.einsum("epqr,pm->eqmr", f, proj, optimize=True)
    return out


def error_norms(mesh, order, dofs, fun, n1d=None, chunk=8192):
    """L1, squared L2 and Linf errors per quantity and run of DOFs
    [nEl][9][N][R] against fun at quadrature points
    (src/setups/InitialDofs.hpp:277-370)."""
    n1d = n1d or order + 1
    pts, w = quad_tet(n1d)
    b = eval_basis(order, pts)
    n_runs = dofs.shape[-1]
    l1 = np.zeros((N_QTS, n_runs))
    l2 = np.zeros((N_QTS, n_runs))
    li = np.zeros((N_QTS, n_runs))
    ve = mesh.el_crds()
    for e0 in range(0, mesh.n_el, chunk):
        els = np.arange(e0, min(mesh.n_el, e0 + chunk))
        x = map_points(mesh, pts, els)
        ref = fun(x)
        if ref.ndim == 3:
            ref = np.repeat(ref[..., None], n_runs, axis=-1)
        num = np.einsum("eqmr,pm->epqr", dofs[els].astype(np.float64), b, optimize=True)
        d = ve[els]
        jd = np.abs(np.linalg.det(d[:, 1:, :] - d[:, :1, :]))
        wp = w[None, :] * jd[:, None]
        diff = np.abs(num - ref)
        l1 += np.einsum("epqr,ep->qr", diff, wp)
        l2 += np.einsum("epqr,ep->qr", diff * diff, wp)
        li = np.maximum(li, diff.max(axis=(0, 1)))
    return l1, l2, li
