// Structured tet4 meshes of a box: every hex is cut into the 6 tets of the
// Kuhn (Freudenthal) split, which is conforming across hexes and across
// periodic sides.
//
// The output follows the conventions the element update depends on:
//  * positively oriented vertex order (tools/edge_v/src/geom/Tet4.cpp:80-112),
//  * local faces 0: 0-2-1, 1: 0-1-3, 2: 0-3-2, 3: 1-2-3 (src/constants.hpp:527-553),
//  * fIdElFaEl = local id of the shared face in the neighbour and
//    vIdElFaEl = position of this face's vertex 0 within the neighbour's face
//    (tools/edge_v/src/geom/Tet4.cpp:114-228),
//  * per-face outward normal of the first adjacent element and tangents
//    (tools/edge_v/src/geom/Tria3.cpp:48-115, mesh/Mesh.cpp:763-841),
//  * in-diameter 12 V / sum |2 A_face| (tools/edge_v/src/geom/Tet4.cpp:49-78).
// Periodic faces are stored once and shared by both elements, so a periodic
// face looks like an interior face to the solver; its type is PERIODIC.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "common.hpp"

namespace {

// local face -> local vertices
const int kFaVe[4][3] = {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {1, 2, 3}};

struct Rng {  // splitmix64, documented so a caller can reproduce the jitter
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

void cross(const double a[3], const double b[3], double c[3]) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
double dot(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
double norm(const double a[3]) { return std::sqrt(dot(a, a)); }

struct TriKey {
  int64_t v[3];
  bool operator==(const TriKey& o) const { return v[0] == o.v[0] && v[1] == o.v[1] && v[2] == o.v[2]; }
};
struct TriHash {
  size_t operator()(const TriKey& k) const {
    uint64_t h = 1469598103934665603ull;
    for (int i = 0; i < 3; i++) h = (h ^ (uint64_t)k.v[i]) * 1099511628211ull;
    return (size_t)(h ^ (h >> 29));
  }
};
TriKey tri_key(int64_t a, int64_t b, int64_t c) {
  if (a > b) std::swap(a, b);
  if (b > c) std::swap(b, c);
  if (a > b) std::swap(a, b);
  return TriKey{{a, b, c}};
}

// Faces of positively oriented tets (el_ve, volume set), matched through
// canonical vertex triples (canon: periodic identification), with their
// types (face_type(eL, local face, eR or -1, &type)), geometry and the
// element-face-element relations.  Writes everything of m but ve_crds,
// el_ve and volume (at most max_fa faces); *n_fa receives the face count.
template <typename Canon, typename FaceType>
int connect(edge_mesh_out* m, int64_t nel, Canon canon, FaceType face_type, int64_t max_fa, int64_t* n_fa) {
  std::unordered_map<TriKey, int64_t, TriHash> faces;
  faces.reserve((size_t)nel * 3);
  int64_t fa_count = 0;
  for (int64_t e = 0; e < nel; e++) {
    for (int lf = 0; lf < 4; lf++) {
      const TriKey key = tri_key(canon(m->el_ve[e * 4 + kFaVe[lf][0]]), canon(m->el_ve[e * 4 + kFaVe[lf][1]]),
                                 canon(m->el_ve[e * 4 + kFaVe[lf][2]]));
      auto it = faces.find(key);
      if (it == faces.end()) {
        if (fa_count >= max_fa) return edge::fail(EDGE_ERR_MESH, "more faces than the output arrays hold");
        faces.emplace(key, fa_count);
        m->fa_el[fa_count * 2] = e;
        m->fa_el[fa_count * 2 + 1] = -1;
        m->el_fa[e * 4 + lf] = fa_count;
        fa_count++;
      } else {
        const int64_t f = it->second;
        if (m->fa_el[f * 2 + 1] != -1) return edge::fail(EDGE_ERR_MESH, "face shared by more than two tets");
        m->fa_el[f * 2 + 1] = e;
        m->el_fa[e * 4 + lf] = f;
      }
    }
  }
  *n_fa = fa_count;
  // face types, neighbours, geometry
  for (int64_t f = 0; f < fa_count; f++) {
    const int64_t eL = m->fa_el[f * 2], eR = m->fa_el[f * 2 + 1];
    int lf = -1;
    for (int s = 0; s < 4; s++)
      if (m->el_fa[eL * 4 + s] == f) lf = s;
    const int st = face_type(eL, lf, eR, &m->fa_type[f]);
    if (st != EDGE_OK) return st;

    // geometry of the left element's face
    double x[3][3], np[3];
    for (int s = 0; s < 3; s++)
      for (int d = 0; d < 3; d++) x[s][d] = m->ve_crds[3 * m->el_ve[eL * 4 + kFaVe[lf][s]] + d];
    {
      int opp = 0;
      for (int v = 0; v < 4; v++) {
        bool in = false;
        for (int s = 0; s < 3; s++) in = in || kFaVe[lf][s] == v;
        if (!in) opp = v;
      }
      for (int d = 0; d < 3; d++) np[d] = m->ve_crds[3 * m->el_ve[eL * 4 + opp] + d];
    }
    double d0[3], d1[3], c[3];
    for (int d = 0; d < 3; d++) {
      d0[d] = x[1][d] - x[0][d];
      d1[d] = x[2][d] - x[0][d];
    }
    cross(d0, d1, c);
    m->areas[f] = 0.5 * norm(c);
    // tangents: d0 and its in-plane normal n0; swapped if d0 x n0 points inwards
    const double alpha = -dot(d0, d1) / dot(d0, d0);
    double n0[3];
    for (int d = 0; d < 3; d++) n0[d] = d1[d] + alpha * d0[d];
    const double l0 = norm(d0), l1 = norm(n0);
    for (int d = 0; d < 3; d++) {
      d0[d] /= l0;
      n0[d] /= l1;
    }
    double n1[3], dn[3];
    cross(d0, n0, n1);
    for (int d = 0; d < 3; d++) dn[d] = np[d] - x[0][d];
    if (dot(n1, dn) > 0)
      for (int d = 0; d < 3; d++) std::swap(d0[d], n0[d]);
    double nrm[3];
    cross(d0, n0, nrm);
    for (int d = 0; d < 3; d++) {
      m->tangents[f * 6 + d] = d0[d];
      m->tangents[f * 6 + 3 + d] = n0[d];
      m->normals[f * 3 + d] = nrm[d];
    }
  }

  // element-face-element relations
  for (int64_t e = 0; e < nel; e++) {
    double asum = 0;
    for (int lf = 0; lf < 4; lf++) {
      const int64_t f = m->el_fa[e * 4 + lf];
      asum += 2.0 * m->areas[f];
      const int64_t ne = (m->fa_el[f * 2] == e) ? m->fa_el[f * 2 + 1] : m->fa_el[f * 2];
      m->el_fa_el[e * 4 + lf] = ne;
      if (ne < 0) {
        m->f_id[e * 4 + lf] = EDGE_NO_ID16;
        m->v_id[e * 4 + lf] = EDGE_NO_ID16;
        continue;
      }
      int nf = -1;
      for (int s = 0; s < 4; s++)
        if (m->el_fa[ne * 4 + s] == f) nf = s;
      if (nf < 0) return edge::fail(EDGE_ERR_MESH, "neighbour misses the shared face");
      const int64_t c0 = canon(m->el_ve[e * 4 + (lf <= 2 ? 0 : 1)]);
      int pos = -1;
      for (int s = 0; s < 3; s++)
        if (canon(m->el_ve[ne * 4 + kFaVe[nf][s]]) == c0) pos = s;
      if (pos < 0) return edge::fail(EDGE_ERR_MESH, "shared face vertex not found in the neighbour");
      m->f_id[e * 4 + lf] = (uint16_t)nf;
      m->v_id[e * 4 + lf] = (uint16_t)pos;
    }
    m->in_dia[e] = 12.0 * m->volume[e] / asum;
  }
  return EDGE_OK;
}

}  // namespace

extern "C" int edge_mesh_box_sizes(const edge_box_params* p, int64_t* n_ve, int64_t* n_el, int64_t* n_fa) {
  if (!p || !n_ve || !n_el || !n_fa) return edge::fail(EDGE_ERR_ARG, "null argument");
  const int64_t n[3] = {p->n[0], p->n[1], p->n[2]};
  for (int d = 0; d < 3; d++) {
    if (n[d] < 1) return edge::fail(EDGE_ERR_ARG, "box needs at least one hex per axis");
    if (p->bc[d] == EDGE_FA_PERIODIC && n[d] < 3)
      return edge::fail(EDGE_ERR_ARG, "periodic axes need at least three hexes");
  }
  const int64_t nh = n[0] * n[1] * n[2];
  *n_ve = (n[0] + 1) * (n[1] + 1) * (n[2] + 1);
  *n_el = 6 * nh;
  int64_t nb = 0;  // boundary (non-periodic) triangles
  for (int d = 0; d < 3; d++)
    if (p->bc[d] != EDGE_FA_PERIODIC) nb += 2 * 2 * (nh / n[d]);
  *n_fa = (4 * (*n_el) - nb) / 2 + nb;
  return EDGE_OK;
}

extern "C" int edge_mesh_box(const edge_box_params* p, edge_mesh_out* m) {
  int64_t nve, nel, nfa;
  int st = edge_mesh_box_sizes(p, &nve, &nel, &nfa);
  if (st != EDGE_OK) return st;
  if (!m || !m->ve_crds || !m->el_ve || !m->el_fa || !m->fa_el || !m->el_fa_el || !m->f_id || !m->v_id ||
      !m->fa_type || !m->normals || !m->tangents || !m->areas || !m->in_dia || !m->volume)
    return edge::fail(EDGE_ERR_ARG, "mesh output arrays missing");
  if (!(p->jitter >= 0.0 && p->jitter < 0.5))
    return edge::fail(EDGE_ERR_ARG, "jitter must be in [0, 0.5) of the hex edge (vertices may not cross)");
  for (int d = 0; d < 3; d++) {
    const int t = p->bc[d];
    if (t != EDGE_FA_PERIODIC && t != EDGE_FA_FREE_SURFACE && t != EDGE_FA_OUTFLOW)
      return edge::fail(EDGE_ERR_ARG, "boundary type must be periodic, free surface or outflow");
  }
  const int64_t n[3] = {p->n[0], p->n[1], p->n[2]};
  const double h[3] = {p->extent[0] / n[0], p->extent[1] / n[1], p->extent[2] / n[2]};
  const int64_t nv[3] = {n[0] + 1, n[1] + 1, n[2] + 1};
  auto vid = [&](int64_t i, int64_t j, int64_t k) { return (k * nv[1] + j) * nv[0] + i; };

  // vertices, interior ones jittered
  Rng rng{p->seed};
  for (int64_t k = 0; k < nv[2]; k++)
    for (int64_t j = 0; j < nv[1]; j++)
      for (int64_t i = 0; i < nv[0]; i++) {
        const int64_t ijk[3] = {i, j, k};
        bool interior = true;
        for (int d = 0; d < 3; d++) interior = interior && ijk[d] > 0 && ijk[d] < n[d];
        double* x = m->ve_crds + 3 * vid(i, j, k);
        for (int d = 0; d < 3; d++) {
          // every vertex draws three numbers so the sequence does not depend on the box shape
          double u = rng.uniform();
          x[d] = p->origin[d] + ijk[d] * h[d] + (interior ? (2.0 * u - 1.0) * p->jitter * h[d] : 0.0);
        }
      }

  // canonical (periodic) vertex id
  auto canon = [&](int64_t v) -> int64_t {
    int64_t i = v % nv[0], j = (v / nv[0]) % nv[1], k = v / (nv[0] * nv[1]);
    if (p->bc[0] == EDGE_FA_PERIODIC) i %= n[0];
    if (p->bc[1] == EDGE_FA_PERIODIC) j %= n[1];
    if (p->bc[2] == EDGE_FA_PERIODIC) k %= n[2];
    return vid(i, j, k);
  };

  // hex order: blocks of b^3 hexes, lexicographic inside and across blocks
  const int64_t b = p->block > 0 ? p->block : 1;
  std::vector<int64_t> hexes;
  hexes.reserve(n[0] * n[1] * n[2]);
  for (int64_t bk = 0; bk < n[2]; bk += b)
    for (int64_t bj = 0; bj < n[1]; bj += b)
      for (int64_t bi = 0; bi < n[0]; bi += b)
        for (int64_t k = bk; k < std::min(bk + b, n[2]); k++)
          for (int64_t j = bj; j < std::min(bj + b, n[1]); j++)
            for (int64_t i = bi; i < std::min(bi + b, n[0]); i++) hexes.push_back((k * n[1] + j) * n[0] + i);

  // elements: Kuhn split, vertex order v0 -> +e_p0 -> +e_p1 -> (1,1,1)
  const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  int64_t el = 0;
  for (int64_t hx : hexes) {
    const int64_t i0 = hx % n[0], j0 = (hx / n[0]) % n[1], k0 = hx / (n[0] * n[1]);
    for (int t = 0; t < 6; t++) {
      int64_t c[3] = {0, 0, 0};
      int64_t ve[4];
      ve[0] = vid(i0, j0, k0);
      for (int s = 0; s < 2; s++) {
        c[perm[t][s]] = 1;
        ve[s + 1] = vid(i0 + c[0], j0 + c[1], k0 + c[2]);
      }
      ve[3] = vid(i0 + 1, j0 + 1, k0 + 1);
      // positive orientation: swap vertices 2 and 3 on a negative determinant
      double a[3][3];
      for (int r = 0; r < 3; r++)
        for (int d = 0; d < 3; d++) a[r][d] = m->ve_crds[3 * ve[r + 1] + d] - m->ve_crds[3 * ve[0] + d];
      double cr[3];
      cross(a[1], a[2], cr);
      const double det = dot(a[0], cr);
      if (std::fabs(det) < 1e-14 * h[0] * h[1] * h[2])
        return edge::fail(EDGE_ERR_MESH, "degenerate tet (jitter too large?)");
      if (det < 0) std::swap(ve[2], ve[3]);
      for (int v = 0; v < 4; v++) m->el_ve[el * 4 + v] = ve[v];
      m->volume[el] = std::fabs(det) / 6.0;
      el++;
    }
  }

  // which box side (if any) a face of element e (local face lf) lies on
  auto side_of = [&](int64_t e, int lf) {
    for (int d = 0; d < 3; d++)
      for (int hi = 0; hi < 2; hi++) {
        bool on = true;
        for (int s = 0; s < 3; s++) {
          const int64_t v = m->el_ve[e * 4 + kFaVe[lf][s]];
          const int64_t ijk[3] = {v % nv[0], (v / nv[0]) % nv[1], v / (nv[0] * nv[1])};
          on = on && ijk[d] == (hi ? n[d] : 0);
        }
        if (on) return 2 * d + hi;
      }
    return -1;
  };
  auto face_type = [&](int64_t eL, int lf, int64_t eR, int32_t* ty) {
    const int side = side_of(eL, lf);
    if (eR < 0) {
      if (side < 0) return edge::fail(EDGE_ERR_MESH, "unmatched interior face");
      *ty = p->bc[side / 2];
    } else {
      *ty = (side >= 0) ? EDGE_FA_PERIODIC : 0;
    }
    return EDGE_OK;
  };
  int64_t nfa_built = 0;
  st = connect(m, nel, canon, face_type, nfa, &nfa_built);
  if (st != EDGE_OK) return st;
  if (nfa_built != nfa) return edge::fail(EDGE_ERR_MESH, "face count mismatch");
  return EDGE_OK;
}

// ---------------------------------------------------------------- any tet mesh
namespace {

int tets_check(const edge_tets_in* in) {
  if (!in || !in->ve_crds || !in->el_ve || (in->n_bf > 0 && (!in->bf_ve || !in->bf_type)))
    return edge::fail(EDGE_ERR_ARG, "tet mesh: argument missing");
  if (in->n_ve < 4 || in->n_el < 1 || in->n_bf < 0) return edge::fail(EDGE_ERR_ARG, "tet mesh: sizes");
  for (int64_t i = 0; i < 4 * in->n_el; i++)
    if (in->el_ve[i] < 0 || in->el_ve[i] >= in->n_ve) return edge::fail(EDGE_ERR_MESH, "tet mesh: vertex id out of range");
  for (int64_t i = 0; i < 3 * in->n_bf; i++)
    if (in->bf_ve[i] < 0 || in->bf_ve[i] >= in->n_ve) return edge::fail(EDGE_ERR_MESH, "tet mesh: vertex id out of range");
  return EDGE_OK;
}

// Periodic boundaries (EDGE-V's setPeriodicBnds, tools/edge_v/src/mesh/
// Mesh.cpp:157-288): every triangle tagged EDGE_FA_PERIODIC lies on a plane
// of one constant coordinate, and pairs with the periodic triangle on another
// plane of the same axis whose vertices agree in the other two coordinates
// (absolute tolerance 1e-6, the reference's).  The paired vertices are
// identified (union-find over vertex ids), so the two triangles' tets match
// through canonical vertex ids as the box generator's periodic faces do.
// canon[v] = the smallest vertex id identified with v.
constexpr double kPeriodicTol = 1e-6;

int periodic_canon(const edge_tets_in* in, std::vector<int64_t>* canon) {
  const int64_t nv = in->n_ve;
  canon->resize(nv);
  for (int64_t v = 0; v < nv; v++) (*canon)[v] = v;
  const double* x = in->ve_crds;
  std::vector<int64_t> tri;  // periodic triangles (indices into bf_ve)
  for (int64_t i = 0; i < in->n_bf; i++)
    if (in->bf_type[i] == EDGE_FA_PERIODIC) tri.push_back(i);
  if (tri.empty()) return EDGE_OK;
  // constant axis and in-plane centroid of each triangle
  std::vector<int> axis(tri.size());
  std::vector<double> cen(2 * tri.size()), pln(tri.size());
  for (size_t t = 0; t < tri.size(); t++) {
    const int64_t* v = in->bf_ve + 3 * tri[t];
    int ax = -1;
    for (int d = 0; d < 3; d++) {
      const double a = x[3 * v[0] + d], b = x[3 * v[1] + d], c = x[3 * v[2] + d];
      if (std::fabs(a - b) < kPeriodicTol && std::fabs(a - c) < kPeriodicTol && std::fabs(b - c) < kPeriodicTol) {
        if (ax >= 0) return edge::fail(EDGE_ERR_MESH, "periodic triangle is degenerate");
        ax = d;
      }
    }
    if (ax < 0) return edge::fail(EDGE_ERR_MESH, "periodic triangle not on a plane of constant coordinate");
    axis[t] = ax;
    pln[t] = x[3 * v[0] + ax];
    for (int k = 0, d = 0; d < 3; d++) {
      if (d == ax) continue;
      cen[2 * t + k++] = (x[3 * v[0] + d] + x[3 * v[1] + d] + x[3 * v[2] + d]) / 3.0;
    }
  }
  // candidates through a hash of the centroid on a grid of cell 4 tol
  // (partners' centroids differ by less than tol: the 3 x 3 neighbourhood)
  const double cell = 4 * kPeriodicTol;
  struct CellKey {
    int ax;
    int64_t i, j;
    bool operator==(const CellKey& o) const { return ax == o.ax && i == o.i && j == o.j; }
  };
  struct CellHash {
    size_t operator()(const CellKey& k) const {
      return (size_t)((uint64_t)k.i * 0x9E3779B97F4A7C15ull ^ (uint64_t)k.j * 0xC2B2AE3D27D4EB4Full ^ (uint64_t)k.ax);
    }
  };
  std::unordered_map<CellKey, std::vector<size_t>, CellHash> grid;
  auto key_of = [&](size_t t, int di, int dj) {
    return CellKey{axis[t], (int64_t)std::floor(cen[2 * t] / cell) + di, (int64_t)std::floor(cen[2 * t + 1] / cell) + dj};
  };
  for (size_t t = 0; t < tri.size(); t++) grid[key_of(t, 0, 0)].push_back(t);
  // vertex b of triangle u matching vertex a (same coordinates off the axis)
  auto match = [&](int64_t a, const int64_t* u, int ax) -> int64_t {
    for (int s = 0; s < 3; s++) {
      bool eq = true;
      for (int d = 0; d < 3; d++)
        if (d != ax && std::fabs(x[3 * a + d] - x[3 * u[s] + d]) >= kPeriodicTol) eq = false;
      if (eq) return u[s];
    }
    return -1;
  };
  std::vector<int64_t> partner(tri.size(), -1);
  for (size_t t = 0; t < tri.size(); t++) {
    const int64_t* v = in->bf_ve + 3 * tri[t];
    for (int di = -1; di <= 1; di++)
      for (int dj = -1; dj <= 1; dj++) {
        const auto it = grid.find(key_of(t, di, dj));
        if (it == grid.end()) continue;
        for (size_t u : it->second) {
          if (u == t || std::fabs(pln[u] - pln[t]) < kPeriodicTol) continue;  // same plane
          const int64_t* w = in->bf_ve + 3 * tri[u];
          if (match(v[0], w, axis[t]) < 0 || match(v[1], w, axis[t]) < 0 || match(v[2], w, axis[t]) < 0) continue;
          if (partner[t] >= 0 && partner[t] != (int64_t)u)
            return edge::fail(EDGE_ERR_MESH, "periodic triangle with more than one partner");
          partner[t] = (int64_t)u;
        }
      }
    if (partner[t] < 0) return edge::fail(EDGE_ERR_MESH, "periodic triangle without a partner on the opposite side");
  }
  // union-find; the root of a set is its smallest vertex id
  std::vector<int64_t>& par = *canon;
  auto find = [&](int64_t v) {
    while (par[v] != v) {
      par[v] = par[par[v]];
      v = par[v];
    }
    return v;
  };
  for (size_t t = 0; t < tri.size(); t++) {
    if ((size_t)partner[partner[t]] != t) return edge::fail(EDGE_ERR_MESH, "periodic triangles do not pair up");
    const int64_t* v = in->bf_ve + 3 * tri[t];
    const int64_t* w = in->bf_ve + 3 * tri[partner[t]];
    for (int s = 0; s < 3; s++) {
      const int64_t a = find(v[s]), b = find(match(v[s], w, axis[t]));
      if (a < b) par[b] = a;
      if (b < a) par[a] = b;
    }
  }
  for (int64_t v = 0; v < nv; v++) par[v] = find(v);
  return EDGE_OK;
}

}  // namespace

extern "C" int edge_mesh_tets_sizes(const edge_tets_in* in, int64_t* n_fa) {
  int st = tets_check(in);
  if (st != EDGE_OK) return st;
  if (!n_fa) return edge::fail(EDGE_ERR_ARG, "null argument");
  std::vector<int64_t> canon;
  if ((st = periodic_canon(in, &canon)) != EDGE_OK) return st;
  std::unordered_map<TriKey, int, TriHash> faces;
  faces.reserve((size_t)in->n_el * 3);
  for (int64_t e = 0; e < in->n_el; e++)
    for (int lf = 0; lf < 4; lf++) {
      const int64_t* v = in->el_ve + e * 4;
      const int c = ++faces[tri_key(canon[v[kFaVe[lf][0]]], canon[v[kFaVe[lf][1]]], canon[v[kFaVe[lf][2]]])];
      if (c > 2) return edge::fail(EDGE_ERR_MESH, "face shared by more than two tets");
    }
  *n_fa = (int64_t)faces.size();
  return EDGE_OK;
}

extern "C" int edge_mesh_tets(const edge_tets_in* in, edge_mesh_out* m) {
  int64_t nfa = 0;
  int st = edge_mesh_tets_sizes(in, &nfa);
  if (st != EDGE_OK) return st;
  if (!m || !m->ve_crds || !m->el_ve || !m->el_fa || !m->fa_el || !m->el_fa_el || !m->f_id || !m->v_id ||
      !m->fa_type || !m->normals || !m->tangents || !m->areas || !m->in_dia || !m->volume)
    return edge::fail(EDGE_ERR_ARG, "mesh output arrays missing");
  std::memcpy(m->ve_crds, in->ve_crds, sizeof(double) * 3 * in->n_ve);
  // positively oriented vertex order (vertices 2 and 3 swapped on a negative determinant)
  for (int64_t e = 0; e < in->n_el; e++) {
    int64_t ve[4];
    for (int v = 0; v < 4; v++) ve[v] = in->el_ve[e * 4 + v];
    double a[3][3], cr[3];
    for (int r = 0; r < 3; r++)
      for (int d = 0; d < 3; d++) a[r][d] = in->ve_crds[3 * ve[r + 1] + d] - in->ve_crds[3 * ve[0] + d];
    cross(a[1], a[2], cr);
    const double det = dot(a[0], cr);
    const double scale = norm(a[0]) * norm(a[1]) * norm(a[2]);
    if (!(std::fabs(det) > 1e-12 * scale)) return edge::fail(EDGE_ERR_MESH, "degenerate tet");
    if (det < 0) std::swap(ve[2], ve[3]);
    for (int v = 0; v < 4; v++) m->el_ve[e * 4 + v] = ve[v];
    m->volume[e] = std::fabs(det) / 6.0;
  }
  std::unordered_map<TriKey, int32_t, TriHash> tags;
  for (int64_t i = 0; i < in->n_bf; i++)
    tags[tri_key(in->bf_ve[3 * i], in->bf_ve[3 * i + 1], in->bf_ve[3 * i + 2])] = in->bf_type[i];
  std::vector<int64_t> canon;
  if ((st = periodic_canon(in, &canon)) != EDGE_OK) return st;
  auto face_type = [&](int64_t eL, int lf, int64_t eR, int32_t* ty) {
    const int64_t* v = m->el_ve + eL * 4;
    const auto it = tags.find(tri_key(v[kFaVe[lf][0]], v[kFaVe[lf][1]], v[kFaVe[lf][2]]));
    *ty = it != tags.end() ? it->second : (eR < 0 ? in->default_bc : 0);
    if (eR < 0 && (*ty == 0 || *ty == EDGE_FA_PERIODIC))
      return edge::fail(EDGE_ERR_MESH, "a boundary face needs a boundary type (free surface or outflow)");
    if (eR >= 0 && *ty == EDGE_FA_PERIODIC) {  // a periodic pair: both triangles tagged
      const int64_t* w = m->el_ve + eR * 4;
      int nf = -1;
      for (int s = 0; s < 4; s++)
        if (m->el_fa[eR * 4 + s] == m->el_fa[eL * 4 + lf]) nf = s;
      if (nf < 0 || !tags.count(tri_key(w[kFaVe[nf][0]], w[kFaVe[nf][1]], w[kFaVe[nf][2]])))
        return edge::fail(EDGE_ERR_MESH, "periodic face tagged on one side only");
    }
    return EDGE_OK;
  };
  int64_t built = 0;
  st = connect(m, in->n_el, [&](int64_t v) { return canon[v]; }, face_type, nfa, &built);
  if (st != EDGE_OK) return st;
  return built == nfa ? EDGE_OK : edge::fail(EDGE_ERR_MESH, "face count mismatch");
}
