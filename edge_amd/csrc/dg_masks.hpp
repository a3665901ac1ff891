// Which (output tile, k-step) blocks of the MFMA products with the
// reference-element matrices touch a non-zero entry, computed at compile
// time from the generated matrices.  In the hierarchical orthogonal basis the
// stiffness matrices are sparse (a derivative lowers the degree, orthogonality
// removes lower degrees: at order 5 18 % of the entries are non-zero), so
// whole 16 x 4 A blocks are zero; the kernels skip their MFMAs (and loads).
// A skipped block would only have added exact zeros.
#pragma once

#include <cstdint>

#include "common.hpp"
#include "generated/dg_tet4.inc"

namespace edge {
namespace masks {

// The generated matrices of order O (degree O - 1), layouts as DgTables.
template <int O>
struct Mats;
#define EDGE_MATS(O, D)                                         \
  template <>                                                   \
  struct Mats<O> {                                              \
    static constexpr const double* stiffT = g_tet4_d##D##_stiffT; \
    static constexpr const double* stiffV = g_tet4_d##D##_stiffV; \
    static constexpr const double* fluxL = g_tet4_d##D##_fluxL;   \
    static constexpr const double* fluxT = g_tet4_d##D##_fluxT;   \
  };
EDGE_MATS(2, 1)
EDGE_MATS(3, 2)
EDGE_MATS(4, 3)
EDGE_MATS(5, 4)
#undef EDGE_MATS

// Row of k-step s in lane group g: row order (4 s + g), or the FP32 D-tile
// order of chained products (16 (s / 4) + 4 g + s % 4).
constexpr int krow(int s, int g, bool perm) { return perm ? 16 * (s / 4) + 4 * g + s % 4 : 4 * s + g; }

// Bit s set: k-step s of output tile t of Out[m] = sum_{k < nk} M[k][m] X[k]
// (m < nn; M row-major with leading dimension ld) has a non-zero A block.
constexpr uint32_t ksteps_nz(const double* m, int ld, int nk, int nn, int t, int nsteps, bool perm) {
  uint32_t bits = 0;
  for (int s = 0; s < nsteps; s++)
    for (int g = 0; g < 4; g++) {
      const int k = krow(s, g, perm);
      if (k >= nk) continue;
      for (int c = 0; c < 16; c++) {
        const int col = 16 * t + c;
        if (col < nn && m[k * ld + col] != 0.0) bits |= 1u << s;
      }
    }
  return bits;
}

// The masks of the element kernels' CK and volume products of order O, for
// the FP32 (perm = the chained products' FP32 row order) or FP64 kernels:
// ck[de][d][t] (stage de = 1 .. O - 1, direction d, output tile t) and
// vol[d][t], bit s = k-step s.
struct Tab {
  uint32_t ck[5][3][3];
  uint32_t vol[3][3];
};

template <int O, bool F32>
constexpr Tab make_tab() {
  Tab x{};
  constexpr int N = n_modes(O), NV = n_modes(O - 1);
  for (int de = 1; de < O; de++) {
    const int nk = n_modes(O - de + 1), nn = n_modes(O - de);
    // k-steps: LDS rows for stage 1, D-tile rows after it
    const int tp = (nk + 15) / 16, rem = nk - 16 * (tp - 1);
    const int ns = de == 1 ? (nk + 3) / 4 : (F32 ? 4 * (tp - 1) + (rem < 4 ? rem : 4) : (nk + 3) / 4);
    for (int d = 0; d < 3; d++)
      for (int t = 0; t < (nn + 15) / 16; t++)
        x.ck[de][d][t] = ksteps_nz(Mats<O>::stiffT + d * N * N, N, nk, nn, t, ns, de >= 2 && F32);
  }
  for (int d = 0; d < 3; d++)
    for (int t = 0; t < (N + 15) / 16; t++)
      x.vol[d][t] = ksteps_nz(Mats<O>::stiffV + d * N * N, N, NV, N, t, (NV + 3) / 4, false);
  return x;
}

template <int O, bool F32>
struct Nz {
  static constexpr Tab tab = make_tab<O, F32>();
};

}  // namespace masks
}  // namespace edge
