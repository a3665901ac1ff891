"""Any-tet meshes (the entity layout EDGE-V produces, src/mesh/EdgeV.cpp:182-466)
built from vertices and elements, and Gmsh files (MSH 2.2 / 4.1 ASCII, read
as EDGE-V does through gmsh, tools/edge_v/src/io/Gmsh.cpp:100-240): a
non-periodic box mesh written to a file and read back gives the box
generator's connectivity, face types and geometry; so does a periodic one,
whose sides pair up by their coordinates as EDGE-V pairs them
(tools/edge_v/src/mesh/Mesh.cpp:157-288); vertex orders are made positive;
faces are matched whatever the vertex order of the input."""
import numpy as np
import pytest

from edge_amd import gmsh
from edge_amd import mesh as M
from edge_amd._lib import EdgeError

KEYS = ["el_ve", "el_fa", "fa_el", "el_fa_el", "f_id", "v_id", "fa_type", "normals", "tangents", "areas", "in_dia",
        "volume"]


def _same(a, b):
    for k in KEYS:
        x, y = getattr(a, k), getattr(b, k)
        if x.dtype.kind == "f":
            np.testing.assert_allclose(x, y, rtol=1e-12, atol=1e-13, err_msg=k)
        else:
            np.testing.assert_array_equal(x, y, err_msg=k)


BC = (M.FA_FREE_SURFACE, M.FA_OUTFLOW, M.FA_FREE_SURFACE)


def test_from_tets_reproduces_the_box_generator():
    box = M.box((3, 4, 2), bc=BC, jitter=0.2, seed=9)
    ty = box.fa_type
    typed = np.nonzero(ty != 0)[0]
    kf = [[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]]
    bf = np.array([[box.el_ve[box.fa_el[f][0]][s] for s in kf[list(box.el_fa[box.fa_el[f][0]]).index(f)]]
                   for f in typed])
    m = M.from_tets(box.ve_crds, box.el_ve, bf, ty[typed])
    _same(m, box)


def test_vertex_order_and_triangle_order_do_not_matter():
    box = M.box(2, bc=(M.FA_OUTFLOW,) * 3, jitter=0.1, seed=3)
    el = box.el_ve.copy()
    el[::2] = el[::2][:, [0, 1, 3, 2]]  # negative orientation on every second tet
    m = M.from_tets(box.ve_crds, el, default_bc=M.FA_OUTFLOW)
    _same(m, box)


@pytest.mark.parametrize("version", [2.2, 4.1])
def test_gmsh_round_trip(tmp_path, version):
    box = M.box((3, 2, 2), bc=BC, jitter=0.15, seed=5)
    p = tmp_path / f"box_{version}.msh"
    gmsh.write(str(p), box, version)
    _same(gmsh.read(str(p)), box)


def test_bad_meshes_are_rejected():
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [2, 2, 2.0]])
    with pytest.raises(EdgeError, match="degenerate"):
        M.from_tets(x, [[0, 1, 2, 2]])
    with pytest.raises(EdgeError, match="out of range"):
        M.from_tets(x, [[0, 1, 2, 7]])
    with pytest.raises(EdgeError, match="boundary type"):
        M.from_tets(x, [[0, 1, 2, 3]], default_bc=0)


@pytest.mark.parametrize("version", [2.2, 4.1])
@pytest.mark.parametrize("bc", [(M.FA_PERIODIC,) * 3, (M.FA_PERIODIC, M.FA_OUTFLOW, M.FA_PERIODIC),
                                (M.FA_FREE_SURFACE, M.FA_FREE_SURFACE, M.FA_PERIODIC)])
def test_periodic_gmsh_round_trip(tmp_path, version, bc):
    box = M.box((4, 3, 5), bc=bc, jitter=0.15, seed=8)
    p = tmp_path / f"periodic_{version}.msh"
    gmsh.write(str(p), box, version)
    m = gmsh.read(str(p))
    _same(m, box)
    assert (m.fa_type == M.FA_PERIODIC).sum() == (box.fa_type == M.FA_PERIODIC).sum() > 0


def test_periodic_pairs_found_whatever_the_element_order():
    box = M.box(3, jitter=0.1, seed=4)
    kf = [[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]]
    per = np.nonzero(box.fa_type == M.FA_PERIODIC)[0]
    bf = [[box.el_ve[e][s] for s in kf[list(box.el_fa[e]).index(f)]] for f in per for e in box.fa_el[f]]
    perm = np.random.default_rng(1).permutation(box.n_el)
    m = M.from_tets(box.ve_crds, box.el_ve[perm], np.array(bf), np.full(len(bf), M.FA_PERIODIC))
    assert m.n_fa == box.n_fa
    assert (m.fa_type == M.FA_PERIODIC).sum() == len(per)
    # every element keeps its neighbours (up to the renumbering)
    for e in range(m.n_el):
        assert sorted(perm[m.el_fa_el[e]]) == sorted(box.el_fa_el[perm[e]])
    np.testing.assert_allclose(m.in_dia, box.in_dia[perm], rtol=1e-12)


def test_unpaired_periodic_triangles_are_rejected():
    box = M.box(3, jitter=0.1, seed=4)
    kf = [[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]]
    per = np.nonzero(box.fa_type == M.FA_PERIODIC)[0]
    bf = [[box.el_ve[e][s] for s in kf[list(box.el_fa[e]).index(f)]] for f in per for e in box.fa_el[f]]
    with pytest.raises(EdgeError, match="partner"):  # one side of one pair missing
        M.from_tets(box.ve_crds, box.el_ve, np.array(bf[1:]), np.full(len(bf) - 1, M.FA_PERIODIC))
    x = box.ve_crds.copy()
    tri = np.array(bf[:1])
    x[tri[0, 0]] += 0.01  # off its plane
    with pytest.raises(EdgeError):
        M.from_tets(x, box.el_ve, tri, np.array([M.FA_PERIODIC]))
