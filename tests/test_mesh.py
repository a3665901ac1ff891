"""Conventions of the generated box meshes: orientation, face numbering,
normals and tangents, neighbour face / vertex ids and periodic matching."""
import numpy as np
import pytest

from edge_amd import mesh as M
from edge_amd import setups as S
from edge_amd._lib import FA_FREE_SURFACE, FA_OUTFLOW, FA_PERIODIC

FA_VE = [(0, 2, 1), (0, 1, 3), (0, 3, 2), (1, 2, 3)]


@pytest.fixture(scope="module", params=[(FA_PERIODIC,) * 3, (FA_OUTFLOW, FA_PERIODIC, FA_FREE_SURFACE)])
def mesh(request):
    return M.box((4, 3, 5), extent=(1.0, 0.75, 1.25), bc=request.param, jitter=0.1, seed=5, block=2)


def test_counts_and_orientation(mesh):
    assert mesh.n_el == 6 * 4 * 3 * 5
    ve = mesh.el_crds()
    det = np.linalg.det(ve[:, 1:, :] - ve[:, :1, :])
    assert np.all(det > 0)
    np.testing.assert_allclose(mesh.volume, det / 6)
    np.testing.assert_allclose(mesh.volume.sum(), 1.0 * 0.75 * 1.25)
    # every element face appears exactly once in el_fa
    f = mesh.el_fa.ravel()
    counts = np.bincount(f, minlength=mesh.n_fa)
    inner = mesh.fa_el[:, 1] >= 0
    assert np.all(counts[inner] == 2) and np.all(counts[~inner] == 1)


def test_normals_tangents(mesh):
    n, t = mesh.normals, mesh.tangents
    np.testing.assert_allclose(np.linalg.norm(n, axis=1), 1, atol=1e-14)
    np.testing.assert_allclose(np.einsum("fi,fi->f", t[:, 0], t[:, 1]), 0, atol=1e-14)
    np.testing.assert_allclose(np.cross(t[:, 0], t[:, 1]), n, atol=1e-14)
    ve = mesh.el_crds()
    for fa in range(mesh.n_fa):
        e = mesh.fa_el[fa, 0]
        lf = list(mesh.el_fa[e]).index(fa)
        fc = ve[e, list(FA_VE[lf])].mean(0)
        assert np.dot(n[fa], fc - ve[e].mean(0)) > 0  # outward for the first element
        a = 0.5 * np.linalg.norm(np.cross(ve[e, FA_VE[lf][1]] - ve[e, FA_VE[lf][0]],
                                          ve[e, FA_VE[lf][2]] - ve[e, FA_VE[lf][0]]))
        assert abs(a - mesh.areas[fa]) < 1e-14


def test_boundary_types(mesh):
    bound = mesh.fa_el[:, 1] < 0
    assert set(np.unique(mesh.fa_type[bound])) <= {FA_FREE_SURFACE, FA_OUTFLOW}
    per = mesh.fa_type == FA_PERIODIC
    assert np.all(mesh.fa_el[per, 1] >= 0)


def test_neighbour_face_and_vertex_ids(mesh):
    """fIdElFaEl points back through the neighbour, and vIdElFaEl is the
    position of this face's vertex 0 in the neighbour's face (coordinates
    compared modulo the box on periodic sides)."""
    ext = np.array([1.0, 0.75, 1.25])
    ve = mesh.el_crds()
    for e in range(mesh.n_el):
        for lf in range(4):
            ne = mesh.el_fa_el[e, lf]
            if ne < 0:
                assert mesh.f_id[e, lf] == 0xFFFF and mesh.v_id[e, lf] == 0xFFFF
                continue
            nf = mesh.f_id[e, lf]
            assert mesh.el_fa[ne, nf] == mesh.el_fa[e, lf]
            x0 = ve[e, 0 if lf <= 2 else 1]
            cand = ve[ne, list(FA_VE[nf])]
            d = cand - x0
            d -= np.round(d / ext) * ext  # periodic images
            pos = int(np.argmin(np.linalg.norm(d, axis=1)))
            assert np.linalg.norm(d[pos]) < 1e-12
            assert mesh.v_id[e, lf] == pos


@pytest.mark.parametrize("order", [2, 3, 4, 5])
def test_traces_match_through_flux_matrices(order):
    """A global polynomial of degree p is represented exactly; its trace seen
    from the left (fluxL_fa) and from the neighbour (fluxN_{v*4+f}) must give
    the same face moments on every interior face: this ties the mesh's
    vIdElFaEl / fIdElFaEl to the matrices' face orientations."""
    m = M.box(3, bc=(FA_OUTFLOW,) * 3, jitter=0.15, seed=9)
    rng = np.random.default_rng(order)
    deg = order - 1
    exps = [(a, b, c) for a in range(deg + 1) for b in range(deg + 1) for c in range(deg + 1) if a + b + c <= deg]
    coef = rng.normal(size=(len(exps), 9))

    def poly(x):
        out = np.zeros(x.shape[:-1] + (9,))
        for (a, b, c), k in zip(exps, coef):
            out += (x[..., 0] ** a * x[..., 1] ** b * x[..., 2] ** c)[..., None] * k
        return out

    dofs = S.project(m, order, poly)[..., 0]
    mats = S.dg_matrices(order)
    checked = 0
    for e in range(m.n_el):
        for lf in range(4):
            ne = m.el_fa_el[e, lf]
            if ne < 0:
                continue
            own = dofs[e] @ mats["fluxL"][lf]
            nb = dofs[ne] @ mats["fluxN"][m.v_id[e, lf] * 4 + m.f_id[e, lf]]
            np.testing.assert_allclose(nb, own, rtol=1e-9, atol=1e-9)
            checked += 1
    assert checked > 0


def test_union_keeps_each_member():
    """The disjoint union holds every member's entities with shifted ids;
    each member's elements see only neighbours of the same member."""
    parts = [M.box(3, jitter=0.1, seed=1), M.box((2, 3, 2), bc=(FA_OUTFLOW, FA_PERIODIC, FA_FREE_SURFACE), seed=2),
             M.box(1, bc=(FA_FREE_SURFACE,) * 3)]
    u, ranges = M.union(parts)
    assert u.n_el == sum(p.n_el for p in parts) and u.n_fa == sum(p.n_fa for p in parts)
    assert ranges == [(0, 162), (162, 72), (234, 6)]
    ve = fa = 0
    for p, (first, n) in zip(parts, ranges):
        sl = slice(first, first + n)
        np.testing.assert_array_equal(u.el_ve[sl], p.el_ve + ve)
        np.testing.assert_array_equal(u.el_fa[sl], p.el_fa + fa)
        ne = u.el_fa_el[sl]
        assert np.all((ne == -1) | ((ne >= first) & (ne < first + n)))
        np.testing.assert_array_equal(np.where(ne >= 0, ne - first, -1), np.where(p.el_fa_el >= 0, p.el_fa_el, -1))
        np.testing.assert_array_equal(u.el_crds()[sl], p.el_crds())
        np.testing.assert_array_equal(u.fa_type[fa:fa + p.n_fa], p.fa_type)
        ve, fa = ve + p.n_ve, fa + p.n_fa
    fe = u.fa_el
    assert (u.el_fa[fe[:, 0]] == np.arange(u.n_fa)[:, None]).any(axis=1).all()
