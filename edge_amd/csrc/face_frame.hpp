// Rotation of the 9 elastic quantities between a face frame (n, s, t) and
// the physical frame: T maps face-aligned to physical quantities, Ti back
// (Voigt order xx yy zz xy yz xz, then velocities), as
// src/impl/seismic/common.hpp:314-344 sets up Trafo3D / TrafoInv3D.
#pragma once

#include <hip/hip_runtime.h>

namespace edge {

// quantities face-aligned (n, s, t) -> xyz and back, Voigt order xx yy zz xy yz xz
__host__ __device__ inline void face_trafo(const double n[3], const double s[3], const double t[3], double T[81], double Ti[81]) {
  const double* f[3] = {n, s, t};
  const int vi[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {1, 2}, {0, 2}};
  for (int i = 0; i < 81; i++) T[i] = Ti[i] = 0;
  for (int r = 0; r < 6; r++)
    for (int c = 0; c < 6; c++) {
      const int i = vi[r][0], j = vi[r][1], a = vi[c][0], b = vi[c][1];
      double v = f[a][i] * f[b][j];
      if (a != b) v += f[b][i] * f[a][j];
      T[r * 9 + c] = v;
      // inverse: face-aligned (a', b') = vi[r] from xyz (i', j') = vi[c]
      const int ap = vi[r][0], bp = vi[r][1], ip = vi[c][0], jp = vi[c][1];
      double w = f[ap][ip] * f[bp][jp];
      if (ip != jp) w += f[ap][jp] * f[bp][ip];
      Ti[r * 9 + c] = w;
    }
  for (int i = 0; i < 3; i++)
    for (int a = 0; a < 3; a++) {
      T[(6 + i) * 9 + 6 + a] = f[a][i];
      Ti[(6 + a) * 9 + 6 + i] = f[a][i];
    }
}

}  // namespace edge
