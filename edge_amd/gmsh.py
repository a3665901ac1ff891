"""Gmsh mesh files (MSH 2.2 and 4.1, ASCII) to tet4 meshes, as EDGE-V
reads them through gmsh (tools/edge_v/src/io/Gmsh.cpp:100-240): the
tetrahedra are the elements, and the triangles of a physical group of
surfaces carry the group's tag as their face type (101 free surface,
105 outflow, or an internal-boundary type).

Periodic boundaries are triangles of a physical group tagged 106 on both
sides; the library pairs them by their coordinates as EDGE-V does
(tools/edge_v/src/mesh/Mesh.cpp:157-288; edge_mesh_tets).

Materials and time groups, which EDGE-V keeps in HDF5 annotations, are the
caller's (no HDF5 library here)."""
import numpy as np

from . import mesh as M


def _sections(text):
    out, cur, buf = {}, None, []
    for line in text.splitlines():
        t = line.strip()
        if t.startswith("$End"):
            out.setdefault(cur, []).append(buf)
            cur, buf = None, []
        elif t.startswith("$") and cur is None:
            cur = t[1:]
        elif cur is not None:
            buf.append(t)
    return out


def _read_v2(sec):
    nodes = sec["Nodes"][0]
    n = int(nodes[0])
    ids = np.zeros(n, dtype=np.int64)
    xyz = np.zeros((n, 3))
    for i, line in enumerate(nodes[1:n + 1]):
        f = line.split()
        ids[i] = int(f[0])
        xyz[i] = [float(x) for x in f[1:4]]
    tets, tris, tri_tags = [], [], []
    els = sec["Elements"][0]
    for line in els[1:int(els[0]) + 1]:
        f = [int(x) for x in line.split()]
        ty, nt = f[1], f[2]
        nodes_ = f[3 + nt:]
        if ty == 4:
            tets.append(nodes_[:4])
        elif ty == 2:
            tris.append(nodes_[:3])
            tri_tags.append(f[3] if nt > 0 else 0)
    return ids, xyz, tets, tris, tri_tags


def _read_v4(sec):
    # surface entity -> physical tag
    phys = {}
    if "Entities" in sec:
        ent = sec["Entities"][0]
        npt, ncu, nsu, _ = (int(x) for x in ent[0].split()[:4])
        for line in ent[1 + npt + ncu:1 + npt + ncu + nsu]:
            f = line.split()
            nph = int(f[7])
            phys[int(f[0])] = int(f[8]) if nph > 0 else 0
    nodes = sec["Nodes"][0]
    nb, nn = (int(x) for x in nodes[0].split()[:2])
    ids = np.zeros(nn, dtype=np.int64)
    xyz = np.zeros((nn, 3))
    pos, k = 1, 0
    for _ in range(nb):
        f = nodes[pos].split()
        cnt = int(f[3])
        tags = [int(x) for x in nodes[pos + 1:pos + 1 + cnt]]
        crd = [[float(x) for x in line.split()[:3]] for line in nodes[pos + 1 + cnt:pos + 1 + 2 * cnt]]
        ids[k:k + cnt] = tags
        xyz[k:k + cnt] = crd
        k += cnt
        pos += 1 + 2 * cnt
    tets, tris, tri_tags = [], [], []
    els = sec["Elements"][0]
    nb = int(els[0].split()[0])
    pos = 1
    for _ in range(nb):
        dim, etag, ty, cnt = (int(x) for x in els[pos].split()[:4])
        for line in els[pos + 1:pos + 1 + cnt]:
            f = [int(x) for x in line.split()]
            if ty == 4:
                tets.append(f[1:5])
            elif ty == 2:
                tris.append(f[1:4])
                tri_tags.append(phys.get(etag, 0) if dim == 2 else 0)
        pos += 1 + cnt
    return ids, xyz, tets, tris, tri_tags


def read(path, default_bc=M.FA_OUTFLOW):
    """Mesh of the tetrahedra in a Gmsh ASCII file; triangles in physical
    groups become tagged faces (untagged boundary faces: default_bc)."""
    with open(path) as f:
        sec = _sections(f.read())
    version = sec["MeshFormat"][0][0].split()
    if version[1] != "0":
        raise ValueError("only ASCII Gmsh files are read")
    v = float(version[0])
    ids, xyz, tets, tris, tri_tags = _read_v4(sec) if v >= 4 else _read_v2(sec)
    if not tets:
        raise ValueError("no tetrahedra in the mesh file")
    index = {int(t): i for i, t in enumerate(ids)}
    el_ve = np.array([[index[t] for t in e] for e in tets], dtype=np.int64)
    sel = [i for i, t in enumerate(tri_tags) if t != 0]
    bf_ve = np.array([[index[t] for t in tris[i]] for i in sel], dtype=np.int64).reshape(-1, 3)
    bf_type = np.array([tri_tags[i] for i in sel], dtype=np.int32)
    return M.from_tets(xyz, el_ve, bf_ve, bf_type, default_bc)


def write(path, mesh, version=4.1):
    """Writes a mesh's tetrahedra and its typed faces (free surface, outflow,
    periodic and other non-zero types) as one physical group per type; a
    periodic face becomes its two triangles, one on each side of the box."""
    ve = mesh.ve_crds
    ty = np.asarray(mesh.fa_type)
    faces = [f for f in range(mesh.n_fa) if ty[f] != 0]
    kf = [[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]]
    tri = {}
    for f in faces:
        for side in ((0, 1) if ty[f] == M.FA_PERIODIC else (0,)):
            e = mesh.fa_el[f][side]
            lf = list(mesh.el_fa[e]).index(f)
            tri.setdefault(int(ty[f]), []).append([int(mesh.el_ve[e][s]) for s in kf[lf]])
    types = sorted(tri)
    lines = []
    if version >= 4:
        lines += ["$MeshFormat", "4.1 0 8", "$EndMeshFormat", "$Entities", f"0 0 {len(types)} 1"]
        lines += [f"{i + 1} 0 0 0 1 1 1 1 {t} 0" for i, t in enumerate(types)]
        lines += ["1 0 0 0 1 1 1 0 0", "$EndEntities", "$Nodes", f"1 {len(ve)} 1 {len(ve)}",
                  f"3 1 0 {len(ve)}"]
        lines += [str(i + 1) for i in range(len(ve))]
        lines += [" ".join(repr(float(x)) for x in p) for p in ve]
        lines += ["$EndNodes", "$Elements"]
        n_tri = sum(len(v) for v in tri.values())
        lines += [f"{len(types) + 1} {n_tri + mesh.n_el} 1 {n_tri + mesh.n_el}"]
        k = 1
        for i, t in enumerate(types):
            lines.append(f"2 {i + 1} 2 {len(tri[t])}")
            for tr in tri[t]:
                lines.append(f"{k} " + " ".join(str(v + 1) for v in tr))
                k += 1
        lines.append(f"3 1 4 {mesh.n_el}")
        for e in mesh.el_ve:
            lines.append(f"{k} " + " ".join(str(int(v) + 1) for v in e))
            k += 1
        lines.append("$EndElements")
    else:
        lines += ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$Nodes", str(len(ve))]
        lines += [f"{i + 1} " + " ".join(repr(float(x)) for x in p) for i, p in enumerate(ve)]
        lines += ["$EndNodes", "$Elements"]
        n_tri = sum(len(v) for v in tri.values())
        lines.append(str(n_tri + mesh.n_el))
        k = 1
        for t in types:
            for tr in tri[t]:
                lines.append(f"{k} 2 2 {t} {t} " + " ".join(str(v + 1) for v in tr))
                k += 1
        for e in mesh.el_ve:
            lines.append(f"{k} 4 2 1 1 " + " ".join(str(int(v) + 1) for v in e))
            k += 1
        lines.append("$EndElements")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
