// FP64 GPU set-up of star matrices and flux solvers.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace edge {

struct SetupInput {  // host arrays
  int64_t n_el, n_fa, n_ve;
  const double* ve_crds;
  const int64_t* el_ve;
  const int64_t* el_fa;
  const int64_t* fa_el;
  const int32_t* fa_type;
  const double* normals;
  const double* tangents;
  const double* areas;
  const double* rho;
  const double* lam;
  const double* mu;
};

// Writes star [nEl][3][81], flux solvers [2][nEl][4][81] (device, type T)
// and the Jacobian determinants d_jdet [nEl] (device, FP64).
template <typename T>
int build_setup(const SetupInput& in, T* d_star, T* d_fs, double* d_jdet, hipStream_t st);

}  // namespace edge
