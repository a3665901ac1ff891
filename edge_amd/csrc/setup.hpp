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
// and the Jacobian determinants d_jdet [nEl] (device, FP64).  With d_fsf
// (may be null) also the rank-3 factors [2][nEl][4][54] of every flux solver
// (U [9][3], then V^T [3][9], from FP64); *rank3 tells whether every solver
// is rank 3 (then fs = U V^T to rounding).
// With d_star_a / d_fs_a (may be null) also the anelastic star matrices
// [nEl][3][6][3] and anelastic flux solvers [2][nEl][4][6][9]; with d_fsf
// and d_fsfa also the anelastic solvers' factors U_A [2][nEl][4][6][3] over
// the elastic factors' V^T (fs_a = U_A V^T to rounding, counted in *rank3).
template <typename T>
int build_setup(const SetupInput& in, T* d_star, T* d_fs, double* d_jdet, T* d_fsf, bool* rank3, T* d_star_a,
                T* d_fs_a, T* d_fsfa, hipStream_t st);

// Rank-3 factors of n_mats host FP64 flux solvers [n_mats][81] into d_fsf
// [n_mats][54] (device); *rank3 is false when any solver is not rank 3.
template <typename T>
int factor_flux_solvers(const double* h_fs, int64_t n_mats, T* d_fsf, bool* rank3);

}  // namespace edge
