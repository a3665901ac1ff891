// MFMA-tiled element kernels (see kernels_mfma.hip).
#pragma once

#include "kernels.hpp"

namespace edge {

bool mfma_supported(int order, int n_runs, int precision);
template <typename T>
int launch_local_mfma(const SolverDev<T>& s, int64_t first, int64_t n, T dt, T* dofs, T* t_int, hipStream_t st);
template <typename T>
int launch_neigh_mfma(const SolverDev<T>& s, int64_t first, int64_t n, const T* t_int, T* dofs, hipStream_t st);
// Neighbour step k - 1 (projections proj_in) and local step k in one launch
// over all elements; this step's projections go to s.proj.
int launch_fused_mfma(const SolverDev<float>& s, float dt, const float* proj_in, float* dofs, float* t_int,
                      hipStream_t st);

}  // namespace edge
