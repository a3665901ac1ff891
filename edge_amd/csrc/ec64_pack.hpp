// Operands of the element-batched single-run kernels (kernels_ec64.hip,
// FP64 and FP32): the reference-element matrices as A operands of the 16x16x4
// MFMAs with 16 elements on the N dimension, and the face-orientation maps
// between the face coordinates of two neighbours.
#pragma once

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace edge {
namespace ec {

constexpr int tiles(int rows) { return (rows + 15) / 16; }
constexpr int ksteps(int rows) { return (rows + 3) / 4; }
constexpr int pairs(int steps) { return (steps + 1) / 2; }  // k-steps in pairs (one load each)

// A product Out[m][e] = sum_k M[k][m] X_e[k] over k < nk, m < nn takes, for
// output tile t and k-step s, the A operand of lane l = 16 g + c:
// M[k(s, g)][16 t + c] (zero outside k < nk, m < nn).  The row map k(s, g):
//   B operands from an LDS tile:        k = 4 s + g (row order);
//   B operands from a previous D tile   (register s % 4 of tile s / 4):
//     FP64 (D rows g + 4 v):            k = 4 s + g;
//     FP32 (D rows 4 g + v):            k = 16 (s / 4) + 4 g + s % 4.
// Packed per product as [t][s / 2][lane][s % 2] (two k-steps per load),
// products one after the other:
//   CK stage 1, direction d:   stiffTn_d, k < N (LDS),                 m < n_modes(O-1)
//   CK stage de >= 2, dir. d:  stiffTn_d, k < n_modes(O-de+1) (D tiles), m < n_modes(O-de)
//   volume, direction d:       stiffV_d,  k < n_modes(O-1) (LDS),        m < N
//   projection, face f:        fluxL_f,   k < N (LDS),                   m < NF
//   lift from D tiles, face f: fluxT_f,   k < NF (D tiles),              m < N
//   lift from LDS, face f:     fluxT_f,   k < NF (LDS),                  m < N
struct Layout {
  int order;
  bool f32;
  // k-steps of a product over nk rows of D tiles
  constexpr int chain(int nk) const {
    return f32 ? 4 * (tiles(nk) - 1) + (nk - 16 * (tiles(nk) - 1) < 4 ? nk - 16 * (tiles(nk) - 1) : 4) : ksteps(nk);
  }
  static constexpr int64_t block(int ks, int nn) { return (int64_t)tiles(nn) * pairs(ks) * 128; }
  constexpr int64_t stage(int de) const {
    return de == 1 ? block(ksteps(n_modes(order)), n_modes(order - 1))
                   : block(chain(n_modes(order - de + 1)), n_modes(order - de));
  }
  constexpr int64_t ck(int de, int d) const {
    int64_t o = 0;
    for (int e = 1; e < de; e++) o += 3 * stage(e);
    return o + d * stage(de);
  }
  constexpr int64_t vol(int d) const { return ck(order, 0) + d * block(ksteps(n_modes(order - 1)), n_modes(order)); }
  constexpr int64_t prj(int f) const { return vol(3) + f * block(ksteps(n_modes(order)), n_face_modes(order)); }
  constexpr int64_t lfd(int f) const {
    return prj(4) + f * block(chain(n_face_modes(order)), n_modes(order));
  }
  constexpr int64_t lfn(int f) const {
    return lfd(4) + f * block(ksteps(n_face_modes(order)), n_modes(order));
  }
  constexpr int64_t size() const { return lfn(4); }
};

// Face-orientation maps, [4][16][16] (row k, column k'): 0 the identity, 1 + v
// the map R_v with fluxN[4 v + f] = fluxL[f] R_v for every face f (the trace
// of a neighbour in its own face coordinates, taken to the receiving
// element's).  Block-diagonal by face-basis degree.
constexpr int kRotSize = 4 * 16 * 16;

}  // namespace ec

// Host: the packed operands (ec::Layout{t.order, f32} order), FP64 values.
std::vector<double> ec_pack_operands(const DgTables& t, bool f32);
// Host: the face-orientation maps; false if some fluxN[4 v + f] is not
// fluxL[f] R_v with one block-diagonal R_v for all four faces (relative
// residual 1e-12).
bool ec_rotations(const DgTables& t, std::vector<double>* rot);

}  // namespace edge
