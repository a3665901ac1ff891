// Rate-2 local time stepping: the extra time-integrated DOFs of elements at
// time-group interfaces and the neighbour data their faces take.
//
//   tDofs[1] (AderDg.hpp:384-411): elements with an adjacent element of a
//     larger time step sum tDofs[0] over the two steps of a pair;
//   tDofs[2] (AderDg.hpp:413-417): elements with an adjacent element of a
//     smaller time step integrate their time prediction over [0, dt/2]
//     (TimePred::integrate, kernels/TimePred.hpp:165-202);
//   neighbour faces (AderDg.hpp:620-674): a face to a smaller-step element
//     takes its tDofs[1], a face to a larger-step element its tDofs[2] on
//     the first step of a pair and tDofs[0] - tDofs[2] on the second.
// The faces between equal steps go through the element kernels as in global
// time stepping; the interface faces are added here, one workgroup per
// interface element.
#include <hip/hip_runtime.h>

#include "gemm_lds.hpp"
#include "kernels.hpp"

namespace edge {
namespace {

// CK derivatives as the local step computes them (one workgroup per listed
// element, from the DOFs at the start of the step) and their Taylor integral
// over [0, h]: t2 = sum_de h^{de+1} / (de+1)! der_de.
template <typename T, int O, int R>
struct HalfLds {
  static constexpr int N = n_modes(O), N1 = n_modes(O > 1 ? O - 1 : 1), N2 = n_modes(O > 2 ? O - 2 : 1);
  static constexpr size_t kQ = 0, kOdd = kQ + sizeof(T) * 9 * N * R, kEven = kOdd + sizeof(T) * 9 * N1 * R,
                          kScr = kEven + sizeof(T) * 9 * N2 * R, kStar = kScr + sizeof(T) * 9 * N1 * R,
                          kSize = kStar + sizeof(T) * 243;
};

template <typename T, int O, int R>
__global__ __launch_bounds__(kThreads) void k_lts_half(SolverDev<T> s, const int64_t* __restrict__ list, T h,
                                                       const T* __restrict__ dofs, T* __restrict__ t2) {
  using L = HalfLds<T, O, R>;
  constexpr int N = L::N, N1 = L::N1, N2 = L::N2, SZ = 9 * N * R;
  extern __shared__ __align__(16) unsigned char lds[];
  T* l_q = reinterpret_cast<T*>(lds + L::kQ);
  T* l_odd = reinterpret_cast<T*>(lds + L::kOdd);
  T* l_even = reinterpret_cast<T*>(lds + L::kEven);
  T* l_scr = reinterpret_cast<T*>(lds + L::kScr);
  T* l_star = reinterpret_cast<T*>(lds + L::kStar);
  const int64_t el = list[blockIdx.x];
  T* out = t2 + el * SZ;
  for (int i = threadIdx.x; i < SZ; i += kThreads) {
    const T v = dofs[el * SZ + i];
    l_q[i] = v;
    out[i] = h * v;
  }
  for (int i = threadIdx.x; i < 243; i += kThreads) l_star[i] = s.star[el * 243 + i];
  T sca = h;
  for (int de = 1; de < O; de++) {
    const int nk = n_modes(O - de + 1), nn = n_modes(O - de);
    const T* prev = de == 1 ? l_q : (de % 2 == 0 ? l_odd : l_even);
    const int ldp = de == 1 ? N : (de % 2 == 0 ? N1 : N2);
    T* cur = de % 2 == 1 ? l_odd : l_even;
    const int ldc = de % 2 == 1 ? N1 : N2;
    for (int d = 0; d < 3; d++) {
      __syncthreads();
      gemm_ac<T, R>(9, nn, nk, ldp, N, N1, false, prev, s.dg.stiffTn + d * N * N, l_scr);
      __syncthreads();
      gemm_bc<T, R>(9, nn, 9, 9, N1, ldc, d > 0, l_star + d * 81, l_scr, cur);
    }
    __syncthreads();
    sca *= h / T(de + 1);
    for (int i = threadIdx.x; i < 9 * nn * R; i += kThreads) {
      const int q = i / (nn * R), mr = i % (nn * R);
      out[q * N * R + mr] += sca * cur[q * ldc * R + mr];
    }
  }
}

// t1 = (first_ts ? 0 : t1) + t0 for the listed elements
template <typename T>
__global__ void k_lts_sum(const int64_t* __restrict__ list, int64_t n, int sz, int first_ts, const T* __restrict__ t0,
                          T* __restrict__ t1) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * sz) return;
  const int64_t el = list[i / sz], k = el * sz + i % sz;
  t1[k] = (first_ts ? T(0) : t1[k]) + t0[k];
}

// The interface faces of the listed elements (one workgroup each): the
// neighbouring surface integral (SurfIntVanilla.hpp:274-357) with the
// neighbour's time-integrated DOFs chosen per face.
template <typename T, int O, int R>
__global__ __launch_bounds__(kThreads) void k_lts_faces(SolverDev<T> s, const uint8_t* __restrict__ fa_lts,
                                                        const int64_t* __restrict__ list, int first_ts,
                                                        const T* __restrict__ t0, const T* __restrict__ t1,
                                                        const T* __restrict__ t2, T* __restrict__ dofs) {
  constexpr int N = n_modes(O), NF = n_face_modes(O), SZ = 9 * N * R, SZF = 9 * NF * R;
  __shared__ T l_dof[SZ], l_tn[SZ], l_f0[SZF], l_f1[SZF], l_fs[81];
  const int64_t el = list[blockIdx.x];
  for (int i = threadIdx.x; i < SZ; i += kThreads) l_dof[i] = dofs[el * SZ + i];
  for (int fa = 0; fa < 4; fa++) {
    const int mode = fa_lts[el * 4 + fa];
    if (mode == 0) continue;  // equal steps: done by the element kernel (uniform per block)
    const int64_t ne = s.nb.ne[el * 4 + fa];
    const int fm = s.nb.fmat[el * 4 + fa];
    __syncthreads();
    for (int i = threadIdx.x; i < SZ; i += kThreads) {
      const int64_t k = ne * SZ + i;
      l_tn[i] = mode == 1 ? t1[k] : (first_ts ? t2[k] : t0[k] - t2[k]);
    }
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

template <typename T, int R>
int half_r(const SolverDev<T>& s, const int64_t* list, int64_t n, T h, const T* dofs, T* t2, hipStream_t st) {
  auto go = [&](auto kern, size_t lds) {
    hipLaunchKernelGGL(kern, dim3((unsigned)n), dim3(kThreads), lds, st, s, list, h, dofs, t2);
    EDGE_HIP_TRY(hipGetLastError());
    return EDGE_OK;
  };
  switch (s.order) {
    case 2: return go(k_lts_half<T, 2, R>, HalfLds<T, 2, R>::kSize);
    case 3: return go(k_lts_half<T, 3, R>, HalfLds<T, 3, R>::kSize);
    case 4: return go(k_lts_half<T, 4, R>, HalfLds<T, 4, R>::kSize);
    case 5: return go(k_lts_half<T, 5, R>, HalfLds<T, 5, R>::kSize);
  }
  return fail(EDGE_ERR_UNSUPPORTED, "local time stepping: order");
}

template <typename T, int O, int R>
constexpr bool faces_fit() {
  return sizeof(T) * (2 * 9 * n_modes(O) * R + 2 * 9 * n_face_modes(O) * R + 81) <= 64 * 1024;
}

template <typename T, int O, int R>
int faces_o(const SolverDev<T>& s, const uint8_t* fa_lts, const int64_t* list, int64_t n, int first_ts, const T* t0,
            const T* t1, const T* t2, T* dofs, hipStream_t st) {
  if constexpr (faces_fit<T, O, R>()) {
    hipLaunchKernelGGL((k_lts_faces<T, O, R>), dim3((unsigned)n), dim3(kThreads), 0, st, s, fa_lts, list, first_ts,
                       t0, t1, t2, dofs);
    EDGE_HIP_TRY(hipGetLastError());
    return EDGE_OK;
  } else {
    return fail(EDGE_ERR_UNSUPPORTED, "local time stepping: tiles do not fit LDS");
  }
}

template <typename T, int R>
int faces_r(const SolverDev<T>& s, const uint8_t* fa_lts, const int64_t* list, int64_t n, int first_ts, const T* t0,
            const T* t1, const T* t2, T* dofs, hipStream_t st) {
  switch (s.order) {
    case 2: return faces_o<T, 2, R>(s, fa_lts, list, n, first_ts, t0, t1, t2, dofs, st);
    case 3: return faces_o<T, 3, R>(s, fa_lts, list, n, first_ts, t0, t1, t2, dofs, st);
    case 4: return faces_o<T, 4, R>(s, fa_lts, list, n, first_ts, t0, t1, t2, dofs, st);
    case 5: return faces_o<T, 5, R>(s, fa_lts, list, n, first_ts, t0, t1, t2, dofs, st);
  }
  return fail(EDGE_ERR_UNSUPPORTED, "local time stepping: order");
}

}  // namespace

bool lts_supported(int order, int n_runs, int prec_bytes) {
  if (order < 2 || order > 5 || (n_runs != 1 && n_runs != 8 && n_runs != 16)) return false;
  const int N = n_modes(order), NF = n_face_modes(order), N1 = n_modes(order - 1),
            N2 = n_modes(order > 2 ? order - 2 : 1);
  const size_t half = (size_t)prec_bytes * (9 * (N + 2 * N1 + N2) * n_runs + 243);
  const size_t faces = (size_t)prec_bytes * (2 * 9 * N * n_runs + 2 * 9 * NF * n_runs + 81);
  return half <= 64 * 1024 && faces <= 64 * 1024;
}

template <typename T>
int launch_lts_half(const SolverDev<T>& s, const int64_t* list, int64_t n, T h, const T* dofs, T* t2,
                    hipStream_t st) {
  if (n <= 0) return EDGE_OK;
  if (s.n_runs == 1) return half_r<T, 1>(s, list, n, h, dofs, t2, st);
  if (s.n_runs == 8) return half_r<T, 8>(s, list, n, h, dofs, t2, st);
  return half_r<T, 16>(s, list, n, h, dofs, t2, st);
}

template <typename T>
int launch_lts_sum(const int64_t* list, int64_t n, int sz, int first_ts, const T* t0, T* t1, hipStream_t st) {
  if (n <= 0) return EDGE_OK;
  const int64_t tot = n * sz;
  hipLaunchKernelGGL(k_lts_sum<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, list, n, sz, first_ts, t0,
                     t1);
  EDGE_HIP_TRY(hipGetLastError());
  return EDGE_OK;
}

template <typename T>
int launch_lts_faces(const SolverDev<T>& s, const uint8_t* fa_lts, const int64_t* list, int64_t n, int first_ts,
                     const T* t0, const T* t1, const T* t2, T* dofs, hipStream_t st) {
  if (n <= 0) return EDGE_OK;
  if (s.n_runs == 1) return faces_r<T, 1>(s, fa_lts, list, n, first_ts, t0, t1, t2, dofs, st);
  if (s.n_runs == 8) return faces_r<T, 8>(s, fa_lts, list, n, first_ts, t0, t1, t2, dofs, st);
  return faces_r<T, 16>(s, fa_lts, list, n, first_ts, t0, t1, t2, dofs, st);
}

template int launch_lts_half<float>(const SolverDev<float>&, const int64_t*, int64_t, float, const float*, float*,
                                    hipStream_t);
template int launch_lts_half<double>(const SolverDev<double>&, const int64_t*, int64_t, double, const double*,
                                     double*, hipStream_t);
template int launch_lts_sum<float>(const int64_t*, int64_t, int, int, const float*, float*, hipStream_t);
template int launch_lts_sum<double>(const int64_t*, int64_t, int, int, const double*, double*, hipStream_t);
template int launch_lts_faces<float>(const SolverDev<float>&, const uint8_t*, const int64_t*, int64_t, int,
                                     const float*, const float*, const float*, float*, hipStream_t);
template int launch_lts_faces<double>(const SolverDev<double>&, const uint8_t*, const int64_t*, int64_t, int,
                                      const double*, const double*, const double*, double*, hipStream_t);

}  // namespace edge
