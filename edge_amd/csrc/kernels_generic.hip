// Generic element kernels: one workgroup per element, DOF tiles staged in LDS,
// threads spread over the (quantity, mode, run) outputs of every small GEMM.
// These are the correctness-first form of the update (any order, precision
// and run count that fits LDS); the tuned kernels live in kernels_mfma.hip.
//
// Local step (src/impl/seismic/solvers/AderDg.hpp:362-533):
//   CK:    der_0 = Q, der_de = sum_d S_d (der_{de-1} (-K_d^T M^-1)),
//          tInt = sum_de dt^{de+1}/(de+1)! der_de, modes truncated per
//          derivative (kernels/TimePredVanilla.hpp:200-306)
//   vol:   Q += sum_d S_d (tInt K_d M^-1)  (kernels/VolIntVanilla.hpp:189-241)
//   surf:  Q += sum_fa ((fs0_fa (tInt F^L_fa)) F^T_fa)
//          (kernels/SurfIntVanilla.hpp:212-262)
// Neighbour step (AderDg.hpp:553-722, SurfIntVanilla.hpp:274-357):
//   Q += sum_fa ((fs1_fa (tInt_ne F^N_{v*4+f})) F^T_fa)
#include <hip/hip_runtime.h>

#include "gemm_lds.hpp"
#include "kernels.hpp"

namespace edge {
namespace {


template <typename T, int O, int R>
struct Sizes {
  static constexpr int N = n_modes(O);
  static constexpr int NF = n_face_modes(O);
  static constexpr int SZ = 9 * N * R;
  static constexpr int SZF = 9 * NF * R;
  static constexpr size_t kLocalLds = sizeof(T) * (5 * SZ + 2 * SZF + 4 * 81 + 3 * 81);
  static constexpr size_t kNeighLds = sizeof(T) * (2 * SZ + 2 * SZF + 81);
};

template <typename T, int O, int R>
__global__ __launch_bounds__(kThreads) void k_local(SolverDev<T> s, int64_t first, T dt, T* __restrict__ dofs,
                                                    T* __restrict__ t_int) {
  using S = Sizes<T, O, R>;
  constexpr int N = S::N, NF = S::NF, SZ = S::SZ;
  __shared__ T l_dof[SZ], l_der[2][SZ], l_scr[SZ], l_tint[SZ], l_f0[S::SZF], l_f1[S::SZF];
  __shared__ T l_star[3 * 81], l_fs[4 * 81];
  const int64_t el = first + blockIdx.x;
  const T* g_dof = dofs + el * SZ;
  for (int i = threadIdx.x; i < SZ; i += kThreads) {
    const T v = g_dof[i];
    l_dof[i] = v;
    l_der[0][i] = v;
    l_tint[i] = dt * v;
  }
  for (int i = threadIdx.x; i < 3 * 81; i += kThreads) l_star[i] = s.star[el * 243 + i];
  for (int i = threadIdx.x; i < 4 * 81; i += kThreads) l_fs[i] = s.fs[el * 324 + i];
  __syncthreads();

  // Cauchy-Kowalevski
  T sca = dt;
  for (int de = 1; de < O; de++) {
    const int nk = n_modes(O - de + 1), nn = n_modes(O - de);
    T* prev = l_der[(de - 1) & 1];
    T* cur = l_der[de & 1];
    for (int i = threadIdx.x; i < SZ; i += kThreads) cur[i] = T(0);
    __syncthreads();
    for (int d = 0; d < 3; d++) {
      gemm_ac<T, R>(9, nn, nk, N, N, N, false, prev, s.dg.stiffTn + d * N * N, l_scr);
      __syncthreads();
      gemm_bc<T, R>(9, nn, 9, 9, N, N, true, l_star + d * 81, l_scr, cur);
      __syncthreads();
    }
    sca *= dt / T(de + 1);
    for (int i = threadIdx.x; i < 9 * nn * R; i += kThreads) {
      const int q = i / (nn * R), mr = i % (nn * R);
      l_tint[q * N * R + mr] += sca * cur[q * N * R + mr];
    }
    __syncthreads();
  }
  T* g_tint = t_int + el * SZ;
  for (int i = threadIdx.x; i < SZ; i += kThreads) g_tint[i] = l_tint[i];

  // volume integral
  for (int d = 0; d < 3; d++) {
    gemm_ac<T, R>(9, N, n_modes(O - 1), N, N, N, false, l_tint, s.dg.stiffV + d * N * N, l_scr);
    __syncthreads();
    gemm_bc<T, R>(9, N, 9, 9, N, N, true, l_star + d * 81, l_scr, l_dof);
    __syncthreads();
  }
  // local surface integral
  for (int fa = 0; fa < 4; fa++) {
    gemm_ac<T, R>(9, NF, N, N, NF, NF, false, l_tint, s.dg.fluxLN + fa * N * NF, l_f0);
    __syncthreads();
    gemm_bc<T, R>(9, NF, 9, 9, NF, NF, false, l_fs + fa * 81, l_f0, l_f1);
    __syncthreads();
    gemm_ac<T, R>(9, N, NF, NF, N, N, true, l_f1, s.dg.fluxT + fa * NF * N, l_dof);
    __syncthreads();
  }
  T* g_out = dofs + el * SZ;
  for (int i = threadIdx.x; i < SZ; i += kThreads) g_out[i] = l_dof[i];
}

template <typename T, int O, int R>
__global__ __launch_bounds__(kThreads) void k_neigh(SolverDev<T> s, int64_t first, const T* __restrict__ t_int,
                                                    T* __restrict__ dofs) {
  using S = Sizes<T, O, R>;
  constexpr int N = S::N, NF = S::NF, SZ = S::SZ;
  __shared__ T l_dof[SZ], l_tn[SZ], l_f0[S::SZF], l_f1[S::SZF], l_fs[81];
  const int64_t el = first + blockIdx.x;
  for (int i = threadIdx.x; i < SZ; i += kThreads) l_dof[i] = dofs[el * SZ + i];
  for (int fa = 0; fa < 4; fa++) {
    const int32_t ne = s.nb.ne[el * 4 + fa];
    if (ne < 0) continue;  // outflow: no neighbouring contribution (uniform per block)
    const int fm = s.nb.fmat[el * 4 + fa];
    __syncthreads();
    for (int i = threadIdx.x; i < SZ; i += kThreads) l_tn[i] = t_int[(int64_t)ne * SZ + i];
    for (int i = threadIdx.x; i < 81; i += kThreads) l_fs[i] = s.fs[((s.n_el + el) * 4 + fa) * 81 + i];
    __syncthreads();
    gemm_ac<T, R>(9, NF, N, N, NF, NF, false, l_tn, s.dg.fluxLN + fm * N * NF, l_f0);
    __syncthreads();
    gemm_bc<T, R>(9, NF, 9, 9, NF, NF, false, l_fs, l_f0, l_f1);
    __syncthreads();
    gemm_ac<T, R>(9, N, NF, NF, N, N, true, l_f1, s.dg.fluxT + fa * NF * N, l_dof);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < SZ; i += kThreads) dofs[el * SZ + i] = l_dof[i];
}

template <typename T, int O, int R>
constexpr bool fits() {
  return Sizes<T, O, R>::kLocalLds <= 160 * 1024;
}

template <typename T, int O, int R>
int run_local(const SolverDev<T>& s, int64_t first, int64_t n, T dt, T* dofs, T* t_int, hipStream_t st) {
  if constexpr (fits<T, O, R>()) {
    hipLaunchKernelGGL((k_local<T, O, R>), dim3((unsigned)n), dim3(kThreads), 0, st, s, first, dt, dofs, t_int);
    EDGE_HIP_TRY(hipGetLastError());
    return EDGE_OK;
  } else {
    return fail(EDGE_ERR_UNSUPPORTED, "generic kernel: tile does not fit LDS");
  }
}

template <typename T, int O, int R>
int run_neigh(const SolverDev<T>& s, int64_t first, int64_t n, const T* t_int, T* dofs, hipStream_t st) {
  if constexpr (fits<T, O, R>()) {
    hipLaunchKernelGGL((k_neigh<T, O, R>), dim3((unsigned)n), dim3(kThreads), 0, st, s, first, t_int, dofs);
    EDGE_HIP_TRY(hipGetLastError());
    return EDGE_OK;
  } else {
    return fail(EDGE_ERR_UNSUPPORTED, "generic kernel: tile does not fit LDS");
  }
}

template <typename T, int R>
int dispatch_local(int order, const SolverDev<T>& s, int64_t first, int64_t n, T dt, T* dofs, T* t_int,
                   hipStream_t st) {
  switch (order) {
    case 2: return run_local<T, 2, R>(s, first, n, dt, dofs, t_int, st);
    case 3: return run_local<T, 3, R>(s, first, n, dt, dofs, t_int, st);
    case 4: return run_local<T, 4, R>(s, first, n, dt, dofs, t_int, st);
    case 5: return run_local<T, 5, R>(s, first, n, dt, dofs, t_int, st);
  }
  return fail(EDGE_ERR_UNSUPPORTED, "order");
}

template <typename T, int R>
int dispatch_neigh(int order, const SolverDev<T>& s, int64_t first, int64_t n, const T* t_int, T* dofs,
                   hipStream_t st) {
  switch (order) {
    case 2: return run_neigh<T, 2, R>(s, first, n, t_int, dofs, st);
    case 3: return run_neigh<T, 3, R>(s, first, n, t_int, dofs, st);
    case 4: return run_neigh<T, 4, R>(s, first, n, t_int, dofs, st);
    case 5: return run_neigh<T, 5, R>(s, first, n, t_int, dofs, st);
  }
  return fail(EDGE_ERR_UNSUPPORTED, "order");
}

}  // namespace

bool generic_supported(int order, int n_runs, int prec_bytes) {
  if (order < 2 || order > 5 || (n_runs != 1 && n_runs != 8 && n_runs != 16)) return false;
  const size_t sz = 9 * n_modes(order) * n_runs, szf = 9 * n_face_modes(order) * n_runs;
  return prec_bytes * (5 * sz + 2 * szf + 7 * 81) <= 160 * 1024;
}

template <typename T>
int launch_local_generic(const SolverDev<T>& s, int64_t first, int64_t n, T dt, T* dofs, T* t_int, hipStream_t st) {
  if (n <= 0) return EDGE_OK;
  if (s.n_runs == 1) return dispatch_local<T, 1>(s.order, s, first, n, dt, dofs, t_int, st);
  if (s.n_runs == 8) return dispatch_local<T, 8>(s.order, s, first, n, dt, dofs, t_int, st);
  if (s.n_runs == 16) return dispatch_local<T, 16>(s.order, s, first, n, dt, dofs, t_int, st);
  return fail(EDGE_ERR_UNSUPPORTED, "runs");
}

template <typename T>
int launch_neigh_generic(const SolverDev<T>& s, int64_t first, int64_t n, const T* t_int, T* dofs, hipStream_t st) {
  if (n <= 0) return EDGE_OK;
  if (s.n_runs == 1) return dispatch_neigh<T, 1>(s.order, s, first, n, t_int, dofs, st);
  if (s.n_runs == 8) return dispatch_neigh<T, 8>(s.order, s, first, n, t_int, dofs, st);
  if (s.n_runs == 16) return dispatch_neigh<T, 16>(s.order, s, first, n, t_int, dofs, st);
  return fail(EDGE_ERR_UNSUPPORTED, "runs");
}

template int launch_local_generic<double>(const SolverDev<double>&, int64_t, int64_t, double, double*, double*,
                                          hipStream_t);
template int launch_local_generic<float>(const SolverDev<float>&, int64_t, int64_t, float, float*, float*,
                                         hipStream_t);
template int launch_neigh_generic<double>(const SolverDev<double>&, int64_t, int64_t, const double*, double*,
                                          hipStream_t);
template int launch_neigh_generic<float>(const SolverDev<float>&, int64_t, int64_t, const float*, float*,
                                         hipStream_t);

}  // namespace edge
