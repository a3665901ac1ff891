// Reference-element matrices as A operands of v_mfma_f32_16x16x4f32 for the
// MFMA-tiled kernels: zero-padded and packed in lane order, so that a lane
// reads the operands of four consecutive k-steps as one 16-byte vector.
//
// For a product with k-steps s = 0, 1, ..., lane l = 16 g + c of k-step s
// holds A[k(s, g)][16 t + c]; the packed block of an (operand, tile) pair is
// [s / 4][l][s % 4].  Row maps k(s, g):
//   CK stage 1, volume, projections:  k = 4 s + g        (row order)
//   CK stages >= 2:                   k = 16 u + 4 g + v (s = 4 u + v, D tiles as B)
//   lift:                             k = 4 g + s        (s = v < 4)
#pragma once

#include <cstdint>
#include <vector>

#include "common.hpp"

namespace edge {
namespace pk {

constexpr int tiles(int rows) { return (rows + 15) / 16; }
constexpr int ksteps(int rows) { return (rows + 3) / 4; }
constexpr int groups(int steps) { return (steps + 3) / 4; }

// CK stage de >= 2: k-steps per direction over the previous derivative's
// rows (permuted order, rows 16 u + 4 g + v below n_modes(order - de + 1))
constexpr int next_steps(int order, int de) {
  const int nk = n_modes(order - de + 1), tp = tiles(nk), rem = nk - 16 * (tp - 1);
  return (tp - 1) * 4 + (rem < 4 ? rem : 4);
}

// Offsets (floats) of the packed operands of one order.
struct Layout {
  int order;
  // stage 1: [d][t][group][lane][4], t < tiles(n_modes(order - 1))
  constexpr int64_t ck1() const { return 0; }
  // stage de >= 2: [d][group][lane][4] (one output tile)
  constexpr int64_t ck(int de) const {
    int64_t o = 3LL * tiles(n_modes(order - 1)) * groups(ksteps(n_modes(order))) * 256;
    for (int e = 2; e < de; e++) o += 3LL * groups(next_steps(order, e)) * 256;
    return o;
  }
  // volume: [d][t][group][lane][4], t < tiles(n_modes(order))
  constexpr int64_t vol() const { return ck(order); }
  // projections with fluxL (0..3) / fluxN (4 + 4 v + f): [m][group][lane][4]
  constexpr int64_t prj() const {
    return vol() + 3LL * tiles(n_modes(order)) * groups(ksteps(n_modes(order - 1))) * 256;
  }
  // lifts with fluxT: [fa][t][lane][4]
  constexpr int64_t lft() const { return prj() + 16LL * groups(ksteps(n_modes(order))) * 256; }
  constexpr int64_t size() const { return lft() + 4LL * tiles(n_modes(order)) * 256; }
};

}  // namespace pk

// Host: the packed operands of order t.order in FP32 (Layout order).
std::vector<float> mfma_pack_operands(const DgTables& t);

}  // namespace edge
