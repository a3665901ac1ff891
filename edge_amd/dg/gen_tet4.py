"""Exact-rational generator of the tet4 reference-element DG matrices.

This is the build step that produces ``edge_amd/csrc/generated/dg_tet4.inc``:
the constant matrices that the element update multiplies with (mass, the two
stiffness forms, and the local / neighbouring / transposed flux matrices),
plus the monomial form of the hierarchical modal basis.

Everything is computed with ``fractions.Fraction`` polynomials and rounded to
double only at the very end (``float(Fraction)`` rounds correctly), so every
entry is the double nearest to the exact rational value.

Reference behaviour restated here (no code taken from it):

* Tet basis: collapsed-coordinate Dubiner basis, hierarchical ordering by
  total degree, then p1, then p2 (tools/edge_pre/edge_pre/dg/basis/Tet.py:31-76).
* Triangle (face) basis: same construction in 2D
  (tools/edge_pre/edge_pre/dg/basis/Tria.py:31-62).
* Reference tet with vertices (0,0,0),(1,0,0),(0,1,0),(0,0,1); face maps
  ``faToEl`` and vertex-orientation maps ``faToFa``
  (tools/edge_pre/edge_pre/types/Tet.py:50-91).
* Matrix definitions (tools/edge_pre/edge_pre/int/Matrices.py:92-249,
  tools/edge_pre/edge_pre.py:508-538):
    mass[i][j]      = int phi_i phi_j
    K_d[i][j]       = int phi_i d/dxi_d phi_j
    stiffT_d        = K_d^T M^-1
    stiffV_d        = K_d   M^-1
    fluxL_fa[k][f]  = int_face phi_k(faToEl_fa(chi)) psi_f(chi)
    fluxN_{v*4+fa}  = (fluxL_fa M_F^-1) Ori_v,
                      Ori_v[g][f] = int_face psi_g(faToFa_v(chi)) psi_f(chi)
    fluxT_fa        = (fluxL_fa M_F^-1)^T M^-1
* Fault sub-faces (the internal boundary's sub-face grid, src/sc/cexprs.inc:
  51-64, ordered as tools/edge_pre/edge_pre/sc/grid/Tet.py:750-777 orders a
  face's sub-triangles):
    sfMean[sf][f]   = mean of psi_f over sub-triangle sf
    sfGather        = (sfMean^T sfMean)^-1 sfMean^T  (sfGather sfMean = I)

Run: ``python -m edge_amd.dg.gen_tet4 [max_degree]``.
"""

from fractions import Fraction
from math import comb, factorial
import os
import sys

# --------------------------------------------------------------------------
# sparse multivariate polynomials over Fraction: {exponent tuple: coeff}
# --------------------------------------------------------------------------


def p_const(c, nvar):
    return {(0,) * nvar: Fraction(c)} if c != 0 else {}


def p_lin(coeffs, const):
    """Affine polynomial const + sum_i coeffs[i] x_i."""
    nvar = len(coeffs)
    out = {}
    if const != 0:
        out[(0,) * nvar] = Fraction(const)
    for i, c in enumerate(coeffs):
        if c != 0:
            e = [0] * nvar
            e[i] = 1
            out[tuple(e)] = Fraction(c)
    return out


def p_add(a, b, sb=1):
    out = dict(a)
    for e, c in b.items():
        v = out.get(e, 0) + sb * c
        if v == 0:
            out.pop(e, None)
        else:
            out[e] = v
    return out


def p_scale(a, s):
    s = Fraction(s)
    return {e: c * s for e, c in a.items()} if s != 0 else {}


def p_mul(a, b):
    out = {}
    for ea, ca in a.items():
        for eb, cb in b.items():
            e = tuple(x + y for x, y in zip(ea, eb))
            v = out.get(e, 0) + ca * cb
            if v == 0:
                out.pop(e, None)
            else:
                out[e] = v
    return out


def p_pow(a, n, nvar):
    out = p_const(1, nvar)
    for _ in range(n):
        out = p_mul(out, a)
    return out


def p_diff(a, var):
    out = {}
    for e, c in a.items():
        if e[var] > 0:
            ne = list(e)
            ne[var] -= 1
            out[tuple(ne)] = c * e[var]
    return out


def p_compose(a, subs):
    """Substitute every variable x_i of ``a`` by the polynomial subs[i]."""
    nvar_out = len(next(iter(subs[0]))) if subs[0] else None
    if nvar_out is None:
        for s in subs:
            if s:
                nvar_out = len(next(iter(s)))
                break
    pows = [[p_const(1, nvar_out)] for _ in subs]
    out = {}
    for e, c in a.items():
        term = p_const(c, nvar_out)
        for i, k in enumerate(e):
            while len(pows[i]) <= k:
                pows[i].append(p_mul(pows[i][-1], subs[i]))
            term = p_mul(term, pows[i][k])
        out = p_add(out, term)
    return out


# --------------------------------------------------------------------------
# Jacobi polynomials in homogeneous form
# --------------------------------------------------------------------------


def jacobi_homog(n, alpha, a, b, nvar):
    """b^n * P_n^{(alpha,0)}(a/b) for polynomials a, b.

    Uses the explicit sum
    P_n^{(al,be)}(x) = sum_s C(n+al, n-s) C(n+be, s) ((x-1)/2)^s ((x+1)/2)^(n-s).
    """
    am = p_scale(p_add(a, b, -1), Fraction(1, 2))  # (a-b)/2
    ap = p_scale(p_add(a, b, +1), Fraction(1, 2))  # (a+b)/2
    out = {}
    for s in range(n + 1):
        c = comb(n + alpha, n - s) * comb(n, s)
        term = p_mul(p_pow(am, s, nvar), p_pow(ap, n - s, nvar))
        out = p_add(out, p_scale(term, c))
    return out


# --------------------------------------------------------------------------
# bases
# --------------------------------------------------------------------------


def tet_basis(deg):
    """Hierarchical tet basis in xi in [0,1]^3-simplex, ordered as the
    reference's generator orders it (degree, then p1, then p2)."""
    nv = 3
    a1 = p_lin([4, 2, 2], -2)   # 2+2x1+x2+x3 with x = 2 xi - 1
    b1 = p_lin([0, -2, -2], 2)  # -x2-x3
    a2 = p_lin([0, 4, 2], -2)   # 1+2x2+x3
    b2 = p_lin([0, 0, -2], 2)   # 1-x3
    x3 = p_lin([0, 0, 2], -1)
    one = p_const(1, nv)
    basis = []
    for de in range(deg + 1):
        for p1 in range(de + 1):
            for p2 in range(de + 1):
                p3 = de - p1 - p2
                if p3 < 0:
                    continue
                fa = jacobi_homog(p1, 0, a1, b1, nv)
                fb = jacobi_homog(p2, 2 * p1 + 1, a2, b2, nv)
                fc = jacobi_homog(p3, 2 * p1 + 2 * p2 + 2, x3, one, nv)
                f = p_mul(p_mul(fa, fb), fc)
                basis.append(p_scale(f, Fraction(1, 2 ** (p1 + p2))))
    return basis


def tria_basis(deg):
    """Hierarchical triangle basis in chi on the unit triangle."""
    nv = 2
    a = p_lin([4, 2], -2)   # 1+2x1+x2
    b = p_lin([0, -2], 2)   # 1-x2
    x2 = p_lin([0, 2], -1)
    one = p_const(1, nv)
    basis = []
    for de in range(deg + 1):
        for p in range(de + 1):
            q = de - p
            fa = jacobi_homog(p, 0, a, b, nv)
            fb = jacobi_homog(q, 2 * p + 1, x2, one, nv)
            basis.append(p_scale(p_mul(fa, fb), Fraction(1, 2 ** p)))
    return basis


# --------------------------------------------------------------------------
# exact integration over the reference simplices
# --------------------------------------------------------------------------

_fact = [factorial(i) for i in range(64)]


def int_tet(poly):
    s = Fraction(0)
    for (a, b, c), v in poly.items():
        s += v * Fraction(_fact[a] * _fact[b] * _fact[c], _fact[a + b + c + 3])
    return s


def int_tria(poly):
    s = Fraction(0)
    for (a, b), v in poly.items():
        s += v * Fraction(_fact[a] * _fact[b], _fact[a + b + 2])
    return s


# face maps chi -> xi, one per face (faces 0: 0-2-1, 1: 0-1-3, 2: 0-3-2, 3: 1-2-3)
def fa_to_el(fa):
    c1 = p_lin([1, 0], 0)
    c2 = p_lin([0, 1], 0)
    zero = {}
    rest = p_lin([-1, -1], 1)
    return [[c2, c1, zero], [c1, zero, c2], [zero, c2, c1], [rest, c1, c2]][fa]


# neighbour face coordinates as a function of the local face coordinates,
# one per position of the local face-vertex 0 in the neighbour's face
def fa_to_fa(v):
    c1 = p_lin([1, 0], 0)
    c2 = p_lin([0, 1], 0)
    rest = p_lin([-1, -1], 1)
    return [[c2, c1], [rest, c2], [c1, rest]][v]


def sub_faces(deg):
    """The uniform sub-triangles of the unit face triangle that the fault's
    friction is evaluated on: n = 2 deg + 1 divisions per edge, n^2
    sub-faces (CE_N_SUB_FACES, src/sc/cexprs.inc:51-64), ordered by rows
    along chi_1 and, in a row, upward- then downward-pointing triangle per
    column (the default face ordering of the sub-grid,
    tools/edge_pre/edge_pre/sc/grid/Tet.py:750-777).  Vertices [3][2] each."""
    n = 2 * deg + 1
    out = []
    for e2 in range(n):
        for e1 in range(n):
            if e1 < n - e2:
                out.append([(Fraction(e1, n), Fraction(e2, n)), (Fraction(e1 + 1, n), Fraction(e2, n)),
                            (Fraction(e1, n), Fraction(e2 + 1, n))])
            if e1 < n - e2 - 1:
                out.append([(Fraction(e1 + 1, n), Fraction(e2, n)), (Fraction(e1 + 1, n), Fraction(e2 + 1, n)),
                            (Fraction(e1, n), Fraction(e2 + 1, n))])
    assert len(out) == n * n
    return out


def sub_face_means(psi, tris):
    """mean[sf][j]: the mean of psi_j over sub-triangle sf (exact)."""
    out = []
    for a, b, c in tris:
        amap = [p_lin([b[0] - a[0], c[0] - a[0]], a[0]), p_lin([b[1] - a[1], c[1] - a[1]], a[1])]
        out.append([2 * int_tria(p_compose(pj, amap)) for pj in psi])
    return out


def solve_exact(a, b):
    """x = a^-1 b for a square Fraction matrix a and a matrix b (Gauss-Jordan)."""
    n = len(a)
    m = [list(a[i]) + list(b[i]) for i in range(n)]
    for c in range(n):
        piv = next(r for r in range(c, n) if m[r][c] != 0)
        m[c], m[piv] = m[piv], m[c]
        inv = 1 / m[c][c]
        m[c] = [v * inv for v in m[c]]
        for r in range(n):
            if r != c and m[r][c] != 0:
                f = m[r][c]
                m[r] = [vr - f * vc for vr, vc in zip(m[r], m[c])]
    return [row[n:] for row in m]


def sub_face_gather(mean):
    """Least-squares inverse of the sub-face means on the face polynomials:
    gather[j][sf] = ((S^T S)^-1 S^T)[j][sf] with S = mean, so that
    gather . mean = I (a face polynomial of degree <= deg is recovered from
    its sub-face means)."""
    nsf, nf = len(mean), len(mean[0])
    sts = [[sum(mean[k][i] * mean[k][j] for k in range(nsf)) for j in range(nf)] for i in range(nf)]
    st = [[mean[k][i] for k in range(nsf)] for i in range(nf)]
    return solve_exact(sts, st)


def generate(deg):
    """Returns a dict of exact Fraction matrices for polynomial degree ``deg``."""
    phi = tet_basis(deg)
    psi = tria_basis(deg)
    n, nf = len(phi), len(psi)

    mass = [[int_tet(p_mul(phi[i], phi[j])) for j in range(n)] for i in range(n)]
    for i in range(n):
        for j in range(n):
            assert (i == j) or mass[i][j] == 0, "tet basis is not orthogonal"
    minv = [Fraction(1) / mass[i][i] for i in range(n)]

    dphi = [[p_diff(phi[j], d) for j in range(n)] for d in range(3)]
    kmat = [[[int_tet(p_mul(phi[i], dphi[d][j])) for j in range(n)] for i in range(n)]
            for d in range(3)]
    stiff_t = [[[kmat[d][j][i] * minv[j] for j in range(n)] for i in range(n)]
               for d in range(3)]
    stiff_v = [[[kmat[d][i][j] * minv[j] for j in range(n)] for i in range(n)]
               for d in range(3)]

    mass_f = [[int_tria(p_mul(psi[i], psi[j])) for j in range(nf)] for i in range(nf)]
    for i in range(nf):
        for j in range(nf):
            assert (i == j) or mass_f[i][j] == 0, "tria basis is not orthogonal"
    mfinv = [Fraction(1) / mass_f[i][i] for i in range(nf)]

    # traces of the element basis on each face in face coordinates
    trace = [[p_compose(phi[k], fa_to_el(fa)) for k in range(n)] for fa in range(4)]
    flux_l = [[[int_tria(p_mul(trace[fa][k], psi[f])) for f in range(nf)]
               for k in range(n)] for fa in range(4)]
    ori = []
    for v in range(3):
        m = fa_to_fa(v)
        rot = [p_compose(psi[g], m) for g in range(nf)]
        ori.append([[int_tria(p_mul(rot[g], psi[f])) for f in range(nf)]
                    for g in range(nf)])
    flux_n = []
    for v in range(3):
        for fa in range(4):
            flux_n.append([[sum(flux_l[fa][k][g] * mfinv[g] * ori[v][g][f]
                                for g in range(nf)) for f in range(nf)]
                           for k in range(n)])
    flux_t = [[[flux_l[fa][k][f] * mfinv[f] * minv[k] for k in range(n)]
               for f in range(nf)] for fa in range(4)]

    sf_mean = sub_face_means(psi, sub_faces(deg))
    sf_gather = sub_face_gather(sf_mean)

    return {
        "deg": deg, "n": n, "nf": nf,
        "sfMean": sf_mean, "sfGather": sf_gather,
        "mass": [mass[i][i] for i in range(n)],
        "mass_face": [mass_f[i][i] for i in range(nf)],
        "stiffT": stiff_t, "stiffV": stiff_v,
        "fluxL": flux_l, "fluxN": flux_n, "fluxT": flux_t,
        "basis": phi, "basis_face": psi,
    }


# --------------------------------------------------------------------------
# C emission
# --------------------------------------------------------------------------


def _flat(x):
    if isinstance(x, list):
        for y in x:
            yield from _flat(y)
    else:
        yield x


def _emit_array(name, data, per_line=4):
    vals = [repr(float(v)) for v in _flat(data)]
    lines = []
    for i in range(0, len(vals), per_line):
        lines.append("  " + ", ".join(vals[i:i + per_line]) + ",")
    return f"static constexpr double {name}[{len(vals)}] = {{\n" + "\n".join(lines) + "\n};\n"


def _emit_basis(name, polys, nvar):
    """Basis as (function id, exponents, coefficient) triples."""
    rows = []
    for b, p in enumerate(polys):
        for e in sorted(p):
            rows.append((b, e, p[e]))
    ent = []
    for b, e, c in rows:
        ent.append("  {%d, {%s}, %r}," % (b, ", ".join(str(x) for x in e) + (", 0" if nvar == 2 else ""),
                                          float(c)))
    return (f"static const edge_dg_term {name}[{len(rows)}] = {{\n" + "\n".join(ent) + "\n};\n",
            len(rows))


def emit(mats_by_deg, path):
    out = ["// Generated by edge_amd/dg/gen_tet4.py -- do not edit.\n",
           "// Exact rational tet4 DG matrices rounded to the nearest double.\n",
           "#pragma once\n",
           "struct edge_dg_term { int fn; int ex[3]; double c; };\n"]
    for m in mats_by_deg:
        d = m["deg"]
        out.append(f"\n// ---- degree {d}: {m['n']} modes, {m['nf']} face modes ----\n")
        out.append(_emit_array(f"g_tet4_d{d}_mass", m["mass"]))
        out.append(_emit_array(f"g_tet4_d{d}_massF", m["mass_face"]))
        out.append(_emit_array(f"g_tet4_d{d}_stiffT", m["stiffT"]))
        out.append(_emit_array(f"g_tet4_d{d}_stiffV", m["stiffV"]))
        out.append(_emit_array(f"g_tet4_d{d}_fluxL", m["fluxL"]))
        out.append(_emit_array(f"g_tet4_d{d}_fluxN", m["fluxN"]))
        out.append(_emit_array(f"g_tet4_d{d}_fluxT", m["fluxT"]))
        out.append(_emit_array(f"g_tet4_d{d}_sfMean", m["sfMean"]))
        out.append(_emit_array(f"g_tet4_d{d}_sfGather", m["sfGather"]))
        s, nt = _emit_basis(f"g_tet4_d{d}_basis", m["basis"], 3)
        out.append(s)
        out.append(f"static const int g_tet4_d{d}_basis_nterms = {nt};\n")
        s, nt = _emit_basis(f"g_tet4_d{d}_basisF", m["basis_face"], 2)
        out.append(s)
        out.append(f"static const int g_tet4_d{d}_basisF_nterms = {nt};\n")
    degs = [m["deg"] for m in mats_by_deg]
    out.append(f"\nstatic const int g_tet4_min_deg = {min(degs)};\n")
    out.append(f"static const int g_tet4_max_deg = {max(degs)};\n")
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        f.write("".join(out))
    os.replace(tmp, path)


DEFAULT_OUT = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                           "csrc", "generated", "dg_tet4.inc")


def main(argv):
    max_deg = int(argv[1]) if len(argv) > 1 else 4
    mats = [generate(d) for d in range(1, max_deg + 1)]
    emit(mats, DEFAULT_OUT)
    print(f"wrote {DEFAULT_OUT} (degrees 1..{max_deg})")


if __name__ == "__main__":
    main(sys.argv)
