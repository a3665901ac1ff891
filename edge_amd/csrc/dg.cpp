// Reference-element data on the host: the generated exact-rational matrices,
// basis evaluation and the collapsed Gauss-Legendre rule, plus error state.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "generated/dg_tet4.inc"

namespace {
thread_local std::string g_last_error;

struct Basis {
  const edge_dg_term* terms;
  int n_terms;
};

bool basis_of(int order, Basis* b) {
  switch (order) {
#define EDGE_CASE(D)                        \
  case D + 1:                               \
    b->terms = g_tet4_d##D##_basis;         \
    b->n_terms = g_tet4_d##D##_basis_nterms; \
    return true;
    EDGE_CASE(1)
    EDGE_CASE(2)
    EDGE_CASE(3)
    EDGE_CASE(4)
#undef EDGE_CASE
    default:
      return false;
  }
}

// Gauss-Legendre nodes / weights on [-1, 1] by Newton iteration on P_n.
void gauss_legendre(int n, double* x, double* w) {
  for (int i = 0; i < n; i++) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; it++) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; k++) {
        double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      const double dp = n * (z * p1 - p0) / (z * z - 1.0);
      const double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    double p0 = 1.0, p1 = z;
    for (int k = 2; k <= n; k++) {
      double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    const double dp = n * (z * p1 - p0) / (z * z - 1.0);
    x[n - 1 - i] = z;
    w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}
}  // namespace

namespace edge {
int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

bool dg_tables(int order, DgTables* t) {
  switch (order) {
#define EDGE_CASE(D)                        \
  case D + 1:                               \
    t->mass = g_tet4_d##D##_mass;           \
    t->stiffT = g_tet4_d##D##_stiffT;       \
    t->stiffV = g_tet4_d##D##_stiffV;       \
    t->fluxL = g_tet4_d##D##_fluxL;         \
    t->fluxN = g_tet4_d##D##_fluxN;         \
    t->fluxT = g_tet4_d##D##_fluxT;         \
    t->massF = g_tet4_d##D##_massF;         \
    t->sfMean = g_tet4_d##D##_sfMean;       \
    t->sfGather = g_tet4_d##D##_sfGather;   \
    break;
    EDGE_CASE(1)
    EDGE_CASE(2)
    EDGE_CASE(3)
    EDGE_CASE(4)
#undef EDGE_CASE
    default:
      return false;
  }
  t->order = order;
  t->n = n_modes(order);
  t->nf = n_face_modes(order);
  t->nsf = n_sub_faces(order);
  return true;
}
}  // namespace edge

extern "C" const char* edge_last_error(void) { return g_last_error.c_str(); }

extern "C" const char* edge_version(void) { return "edge_amd 0.1 (gfx950)"; }

extern "C" int edge_dg_sizes(int order, int* n_modes, int* n_face_modes) {
  if (order < edge::kMinOrder || order > edge::kMaxOrder) return edge::fail(EDGE_ERR_UNSUPPORTED, "order");
  if (n_modes) *n_modes = edge::n_modes(order);
  if (n_face_modes) *n_face_modes = edge::n_face_modes(order);
  return EDGE_OK;
}

extern "C" int edge_dg_matrices(int order, double* mass, double* stiff_t, double* stiff_v, double* flux_l,
                                double* flux_n, double* flux_t) {
  edge::DgTables t;
  if (!edge::dg_tables(order, &t)) return edge::fail(EDGE_ERR_UNSUPPORTED, "order not generated");
  const size_t n = t.n, nf = t.nf;
  if (mass) std::memcpy(mass, t.mass, sizeof(double) * n);
  if (stiff_t) std::memcpy(stiff_t, t.stiffT, sizeof(double) * 3 * n * n);
  if (stiff_v) std::memcpy(stiff_v, t.stiffV, sizeof(double) * 3 * n * n);
  if (flux_l) std::memcpy(flux_l, t.fluxL, sizeof(double) * 4 * n * nf);
  if (flux_n) std::memcpy(flux_n, t.fluxN, sizeof(double) * 12 * n * nf);
  if (flux_t) std::memcpy(flux_t, t.fluxT, sizeof(double) * 4 * n * nf);
  return EDGE_OK;
}

extern "C" int edge_dg_eval_basis(int order, int64_t n_pts, const double* pts, double* val) {
  Basis b;
  if (!basis_of(order, &b)) return edge::fail(EDGE_ERR_UNSUPPORTED, "order not generated");
  if (n_pts < 0 || (n_pts > 0 && (!pts || !val))) return edge::fail(EDGE_ERR_ARG, "points");
  const int n = edge::n_modes(order);
  for (int64_t p = 0; p < n_pts; p++) {
    double* v = val + p * n;
    for (int i = 0; i < n; i++) v[i] = 0;
    const double x = pts[3 * p], y = pts[3 * p + 1], z = pts[3 * p + 2];
    for (int t = 0; t < b.n_terms; t++) {
      const edge_dg_term& tm = b.terms[t];
      v[tm.fn] += tm.c * std::pow(x, tm.ex[0]) * std::pow(y, tm.ex[1]) * std::pow(z, tm.ex[2]);
    }
  }
  return EDGE_OK;
}

extern "C" int edge_dg_quad_tet(int n1d, double* pts, double* w) {
  if (n1d < 1 || n1d > 32 || !pts || !w) return edge::fail(EDGE_ERR_ARG, "quadrature size");
  std::vector<double> x(n1d), wt(n1d);
  gauss_legendre(n1d, x.data(), wt.data());
  // collapsed coordinates, unit tet; weights include the collapse Jacobian
  for (int i = 0; i < n1d; i++)
    for (int j = 0; j < n1d; j++)
      for (int k = 0; k < n1d; k++) {
        const int q = (i * n1d + j) * n1d + k;
        const double a = x[i], b = x[j], c = x[k];
        pts[3 * q + 0] = (a + 1) * (b - 1) * (c - 1) / 8.0;
        pts[3 * q + 1] = (b + 1) * (c - 1) / -4.0;
        pts[3 * q + 2] = (c + 1) / 2.0;
        w[q] = wt[i] * wt[j] * wt[k] * (b - 1) * (c - 1) * (c - 1) / -64.0;
      }
  return EDGE_OK;
}
