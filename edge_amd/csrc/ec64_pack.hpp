// Operands of the FP64 element-batched kernels (kernels_ec64.hip): the
// reference-element matrices as A operands of v_mfma_f64_16x16x4f64 with 16
// elements on the N dimension, and the face-orientation maps between the
// face coordinates of two neighbours.
#pragma once

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace edge {
namespace ec {

constexpr int tiles(int rows) { return (rows + 15) / 16; }
constexpr int ksteps(int rows) { return (rows + 3) / 4; }
constexpr int pairs(int steps) { return (steps + 1) / 2; }  // k-steps in 16-byte pairs

// A product Out[m][e] = sum_k M[k][m] X_e[k] over k < nk, m < nn takes, for
// output tile t and k-step s, the A operand of lane l = 16 g + c:
// M[4 s + g][16 t + c] (zero outside k < nk, m < nn).  Packed per product as
// [t][s / 2][lane][s % 2] (two k-steps per 16-byte load), products one after
// the other:
//   CK stage de (1 .. O-1), direction d: stiffTn_d, k < n_modes(O-de+1), m < n_modes(O-de)
//   volume, direction d:                 stiffV_d,  k < n_modes(O-1),    m < N
//   projection, face f:                  fluxL_f,   k < N,               m < NF
//   lift, face f:                        fluxT_f,   k < NF,              m < N
struct Layout {
  int order;
  static constexpr int64_t block(int nk, int nn) { return (int64_t)tiles(nn) * pairs(ksteps(nk)) * 128; }
  constexpr int64_t ck(int de, int d) const {
    int64_t o = 0;
    for (int e = 1; e < de; e++) o += 3 * block(n_modes(order - e + 1), n_modes(order - e));
    return o + d * block(n_modes(order - de + 1), n_modes(order - de));
  }
  constexpr int64_t vol(int d) const { return ck(order, 0) + d * block(n_modes(order - 1), n_modes(order)); }
  constexpr int64_t prj(int f) const { return vol(3) + f * block(n_modes(order), n_face_modes(order)); }
  constexpr int64_t lft(int f) const { return prj(4) + f * block(n_face_modes(order), n_modes(order)); }
  constexpr int64_t size() const { return lft(4); }
};

// Face-orientation maps, [4][16][16] (row k, column k'): 0 the identity, 1 + v
// the map R_v with fluxN[4 v + f] = fluxL[f] R_v for every face f (the trace
// of a neighbour in its own face coordinates, taken to the receiving
// element's).  Block-diagonal by face-basis degree.
constexpr int kRotSize = 4 * 16 * 16;

}  // namespace ec

// Host: the packed operands (ec::Layout order) of the order t.order.
std::vector<double> ec64_pack_operands(const DgTables& t);
// Host: the face-orientation maps; false if some fluxN[4 v + f] is not
// fluxL[f] R_v with one R_v for all four faces (relative residual 1e-12).
bool ec64_rotations(const DgTables& t, std::vector<double>* rot);

}  // namespace edge
