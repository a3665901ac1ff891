"""Extracts the reference's own golden vectors into small JSON fixtures.

Run in a container that has the reference tree mounted (it reads
``/root/reference`` and writes next to this script).  Only numbers are
extracted -- inputs and expected outputs of the reference's unit tests -- no
source text is kept.

Sources (all read as text):

* src/impl/seismic/kernels/TimePred.test.inc  -- degree-3 transposed stiffness
  (python block, lines 29-90) and the C tables of the CK test (lines 245-400).
* src/impl/seismic/kernels/VolInt.test.inc    -- degree-3 volume stiffness and
  the volume-integral test tables.
* src/impl/seismic/kernels/SurfInt.test.inc   -- degree-2 fluxL / fluxN[6] /
  fluxT and the surface-integral test tables.
* tools/edge_pre/edge_pre/int/test_Matrices.py -- exact degree-2 stiffness
  matrices (test_stiffTet) and degree-3 flux products fm_fa / fp_{f1 f2 v}
  (test_fluxTet).
* src/dg/generated/basis_tet.inc, basis_tria.inc -- the reference's generated
  basis polynomials, evaluated here at seeded points (values only are stored).
* src/impl/seismic/solvers/InternalBoundary.test.cpp -- the 3D internal-boundary
  test: left / right quantities and the expected perturbed middle states after
  linear slip weakening (lines 339-628).
* src/linalg/Series.test.cpp -- sampled values, arguments and expected
  integrals of every ``Series::integrate`` call.
* src/impl/advection/kernels/TimePred.test.inc -- time derivatives, time step
  and the [0, dt] and [0, dt/2] Taylor integrals of the advection CK test.
* src/impl/seismic/setups/ViscoElasticity.test.cpp -- expected values of the
  viscoelastic set-up tests (relaxation frequencies, anelastic coefficients,
  elastic Lame parameters, source and star matrices, 2D middle-state flux
  solvers), stored as (array, index, value).
"""

from fractions import Fraction
import json
import math
import os
import re
import sys

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
KER = os.path.join(REF, "src/impl/seismic/kernels")


def _read(path):
    with open(path) as f:
        return f.read()


def py_list(text, name):
    """Values of a python list literal ``name = [ ... ]`` inside a comment."""
    m = re.search(r"^%s\s*=\s*\[(.*?)\]" % re.escape(name), text, re.S | re.M)
    if m is None:
        raise KeyError(name)
    return [float(t) for t in re.findall(r"[-+]?[0-9]*\.?[0-9]+(?:[eE][-+]?[0-9]+)?", m.group(1))]


def c_array(text, name):
    """Values and shape of ``float name[a][b].. = { ... };``."""
    m = re.search(r"float\s+%s((?:\[\d+\])+)\s*=\s*\{(.*?)\};" % re.escape(name), text, re.S)
    if m is None:
        raise KeyError(name)
    shape = [int(x) for x in re.findall(r"\[(\d+)\]", m.group(1))]
    vals = [float(t) for t in re.findall(r"[-+]?[0-9]*\.?[0-9]+(?:[eE][-+]?[0-9]+)?", m.group(2))]
    assert len(vals) == int(np.prod(shape)), (name, len(vals), shape)
    return np.array(vals).reshape(shape)


def _frac_tok(tok):
    tok = tok.strip()
    if "/" in tok:
        a, b = tok.split("/")
        return Fraction(int(a), int(b))
    return Fraction(tok)


def sympy_matrix_block(text, key):
    """Rational entries of ``'key': sympy.Matrix([ [..], .. ])``."""
    m = re.search(r"'%s'\s*:\s*sympy\.Matrix\(\s*\[(.*?)\]\s*\)" % re.escape(key), text, re.S)
    if m is None:
        raise KeyError(key)
    rows = re.findall(r"\[([^\[\]]*)\]", m.group(1))
    return [[_frac_tok(t) for t in r.split(",") if t.strip()] for r in rows]


def stiff_ref_deg2(text):
    """The three exact 10x10 stiffness matrices of test_stiffTet."""
    start = text.index("def test_stiffTet")
    block = text[start:text.index("# 3D coords", start)]
    mats = re.findall(r"sympy\.Matrix\(\[(.*?)\]\)", block, re.S)
    out = []
    for mt in mats:
        rows = re.findall(r"\[([^\[\]]*)\]", mt)
        out.append([[float(_frac_tok(t)) for t in r.split(",") if t.strip()] for r in rows])
    assert len(out) == 3
    return out


def c_expr_basis(path, var_name, nvar, max_fn, pts):
    """Evaluates the reference's generated basis (C expressions of the form
    ``if( b == i ) { val = <expr> ; }``) at the given points."""
    txt = _read(path)
    exprs = {}
    for m in re.finditer(r"if\(\s*b\s*==\s*(\d+)\s*\)\s*\{\s*%s\s*=\s*(.*?);\s*\}" % var_name, txt, re.S):
        b = int(m.group(1))
        if b < max_fn:
            exprs[b] = re.sub(r"(\d)L\b", r"\1", m.group(2).replace("std::pow", "pow").replace("std::sqrt", "sqrt"))
    assert len(exprs) == max_fn, (path, len(exprs))
    out = np.zeros((max_fn, len(pts)))
    for b in range(max_fn):
        code = compile(exprs[b], "<basis>", "eval")
        for p, pt in enumerate(pts):
            env = {"pow": pow, "sqrt": math.sqrt}
            for d in range(nvar):
                env["xi_%d" % (d + 1)] = float(pt[d])
            out[b, p] = eval(code, {"__builtins__": {}}, env)
    return out


VISCO_CASES = {  # reference TEST_CASE title -> fixture key
    "Derivation of relaxation frequencies.": "frequencies",
    "Derivation of anelastic coefficients, p/s-version": "coeff_ps",
    "Derivation of elastic Lame parameters.": "elastic_lame",
    "Derivation of anelastic coefficients, lame-version": "coeff_lame",
    "Derivation of the source matrix in 3D": "src3d",
    "Derivation of the viscoelastic part of the star matrix in 3D": "star3d",
    "Derivation of the viscoelastic part of the flux solvers in 2D": "fs_mid2d",
}


def _num_expr(tok):
    """Value of a literal or a product of literals such as '-0.5 * 4'."""
    v = 1.0
    for f in tok.split("*"):
        v *= float(f)
    return v


def visco_golden(text):
    """(array name, indices, expected value) of every REQUIRE of the
    viscoelastic set-up test cases, in file order."""
    out = {}
    parts = re.split(r'TEST_CASE\(\s*"', text)[1:]
    for part in parts:
        title = part.split('"')[0]
        key = VISCO_CASES.get(title)
        if key is None:
            continue
        rows = []
        for name, idx, val in re.findall(
                r"REQUIRE\(\s*l_(\w+)((?:\[[^\]]+\])*)\s*==\s*(?:Approx\(\s*)?([-+0-9.eE* ]+?)\s*\)?\s*\);", part):
            try:
                ids = [eval_index(i) for i in re.findall(r"\[([^\]]+)\]", idx)]
            except ValueError:  # indexed by a loop variable (checks of zero rows)
                continue
            rows.append([name, ids, _num_expr(val)])
        out[key] = rows
    return out


def eval_index(tok):
    """Value of an index expression made of integer literals, + - and *."""
    tok = tok.replace(" ", "")
    total = 0
    for term in re.findall(r"[+-]?[^+-]+", tok):
        sign = -1 if term.startswith("-") else 1
        v = 1
        for f in term.lstrip("+-").split("*"):
            v *= int(f)
        total += sign * v
    return total


def main():
    fx = {}

    # ---------------- CK (TimePred) ----------------
    t = _read(os.path.join(KER, "TimePred.test.inc"))
    fx["timepred"] = {
        "stiffT": [py_list(t, "l_stiffT%d" % d) for d in range(3)],
        **{k: c_array(t, "l_" + k).tolist() for k in
           ["rfs", "starE", "starA", "srcA", "dofsE", "dofsA", "refEtDofs", "refVtDofsE", "refVtDofsA"]},
        "dt": 0.017,
    }

    # ---------------- volume integral ----------------
    t = _read(os.path.join(KER, "VolInt.test.inc"))
    fx["volint"] = {
        "stiffV": [py_list(t, "l_stiff%d" % d) for d in range(3)],
        **{k: c_array(t, "l_" + k).tolist() for k in
           ["rfs", "starE", "starA", "srcA", "dofsE", "dofsA", "tDofsE", "tDofsA",
            "refEdofs", "refVdofsE", "refVdofsA"]},
    }

    # ---------------- surface integral ----------------
    t = _read(os.path.join(KER, "SurfInt.test.inc"))
    fx["surfint"] = {
        "fluxL": py_list(t, "l_fIntL"),
        "fluxN6": py_list(t, "l_fIntN6"),
        "fluxT": py_list(t, "l_fIntT"),
        **{k: c_array(t, "l_" + k).tolist() for k in
           ["rfs", "fSolvE", "fSolvA", "dofsE", "dofsA", "tDofsE", "refEdofs", "refEneighDofs",
            "refEneighFdofs", "refVdofsA", "refVneighDofsA", "refVneighFdofsA"]},
    }
    with open(os.path.join(HERE, "ref_kernel_fixtures.json"), "w") as f:
        json.dump(fx, f)

    # ---------------- exact matrices from the generator's tests ----------------
    t = _read(os.path.join(REF, "tools/edge_pre/edge_pre/int/test_Matrices.py"))
    start = t.index("def test_fluxTet")
    tt = t[start:]
    mats = {"stiff_deg2": stiff_ref_deg2(t)}
    mats["fm_deg3"] = [[[float(v) for v in r] for r in sympy_matrix_block(tt, "fm%d" % fa)] for fa in range(4)]
    fp = {}
    for f1 in range(4):
        for f2 in range(4):
            for v in range(3):
                key = "fp%d%d%d" % (f1, f2, v)
                fp[key] = [[float(x) for x in r] for r in sympy_matrix_block(tt, key)]
    mats["fp_deg3"] = fp

    # ---------------- generated basis polynomials ----------------
    rng = np.random.default_rng(3343)
    pts3 = []
    while len(pts3) < 24:
        p = rng.random(3)
        if p.sum() < 1.0:
            pts3.append(p)
    pts2 = []
    while len(pts2) < 16:
        p = rng.random(2)
        if p.sum() < 1.0:
            pts2.append(p)
    gen = os.path.join(REF, "src/dg/generated")
    nb3 = 56  # degree <= 5
    nb2 = 21
    mats["basis_tet_pts"] = [list(map(float, p)) for p in pts3]
    mats["basis_tet_val"] = c_expr_basis(os.path.join(gen, "basis_tet.inc"), "val", 3, nb3, pts3).tolist()
    for d in range(3):
        mats["basis_tet_der%d" % d] = c_expr_basis(os.path.join(gen, "basis_tet_der_xi%d.inc" % (d + 1)),
                                                   "valDxi%d" % (d + 1), 3, nb3, pts3).tolist()
    mats["basis_tria_pts"] = [list(map(float, p)) for p in pts2]
    mats["basis_tria_val"] = c_expr_basis(os.path.join(gen, "basis_tria.inc"), "val", 2, nb2, pts2).tolist()
    with open(os.path.join(HERE, "ref_matrix_fixtures.json"), "w") as f:
        json.dump(mats, f)

    # ---------------- internal boundary (rupture), 3D test ----------------
    t = _read(os.path.join(REF, "src/impl/seismic/solvers/InternalBoundary.test.cpp"))
    t3 = t[t.index('"Internal Boundary: 3D solvers with middle state perturbation."'):]
    ib = {}
    for name in ["qL", "qR", "msLp", "msRp"]:
        m = re.search(r"double\s+l_%s\[9\]\s*=\s*\{(.*?)\};" % name, t3, re.S)
        ib[name] = [float(x) for x in re.findall(r"[-+]?[0-9]*\.?[0-9]+(?:[eE][-+]?[0-9]+)?", m.group(1))]
    with open(os.path.join(HERE, "ref_rupture_fixtures.json"), "w") as f:
        json.dump(ib, f, indent=0)

    # ---------------- time series integration and Taylor integration ----------------
    # src/linalg/Series.test.cpp: the 5 sample values (first series) and every
    # integrate(dx, x0, n, vals, x1, x2, out[, na]) call with its expected value
    t = _read(os.path.join(REF, "src/linalg/Series.test.cpp"))
    m = re.search(r"double\s+l_vals\[5\]\[8\]\s*=\s*\{(.*?)\};", t, re.S)
    vals = [float(x) for x in re.findall(r"[-+]?[0-9]*\.?[0-9]+", m.group(1))]
    num = r"([-+]?[0-9]*\.?[0-9]+)"
    calls = re.findall(r"integrate\(\s*" + num + r",\s*" + num + r",\s*(\d+),\s*l_vals,\s*" + num + r",\s*" + num +
                       r",\s*l_re\s*(?:,\s*" + num + r")?\s*\);.*?Approx\((.*?)\)\s*\);", t, re.S)
    series = {"vals": vals[::8], "cases": []}
    for dx, x0, npt, x1, x2, na, exp in calls:
        series["cases"].append({"dx": float(dx), "x0": float(x0), "n": int(npt), "x1": float(x1), "x2": float(x2),
                                "na": float(na) if na else 0.0, "expected": float(eval(exp, {"__builtins__": {}}))})
    # src/impl/advection/kernels/TimePred.test.inc: derivatives, dt and the
    # [0, dt] and [0, dt/2] Taylor integrals of the advection CK test
    t = _read(os.path.join(REF, "src/impl/advection/kernels/TimePred.test.inc"))
    adv = {"dt": float(re.search(r"float\s+l_utDt\s*=\s*([-+0-9.eE]+);", t).group(1)),
           "ders": c_array(t, "l_utDers").tolist(), "tint": c_array(t, "l_utTint").tolist(),
           "tint12": c_array(t, "l_utTint12").tolist()}
    with open(os.path.join(HERE, "ref_series_fixtures.json"), "w") as f:
        json.dump({"series": series, "advection_time_pred": adv}, f, indent=0)

    # ---------------- viscoelastic set-up ----------------
    t = _read(os.path.join(REF, "src/impl/seismic/setups/ViscoElasticity.test.cpp"))
    with open(os.path.join(HERE, "ref_visco_fixtures.json"), "w") as f:
        json.dump(visco_golden(t), f, indent=0)
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    sys.exit(main())
