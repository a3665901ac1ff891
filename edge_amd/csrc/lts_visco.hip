// Rate-2 local time stepping of viscoelastic solvers: the time-group
// interface data with the anelastic terms.
//
//   tDofs[2] (AderDg.hpp:413-417): the Taylor integral over [0, dt/2] of the
//     elastic derivatives of the CK with the anelastic terms
//     (TimePredVanilla.hpp:200-306 with TL_N_RMS > 0: every derivative keeps
//     all modes, der_de += sum_rm E_rm derA_{rm,de-1},
//     derA_{rm,de} = w_rm (S^A der_{de-1}K + derA_{rm,de-1}));
//   interface faces (AderDg.hpp:620-674, :715-720): the neighbouring surface
//     integral of the chosen neighbour tInt, elastic part into the DOFs and
//     the anelastic part (fs_A) scattered into every mechanism scaled by w_rm
//     (SurfInt.hpp:139-156).
// tDofs[0] and tDofs[1] are the elastic time-integrated DOFs as without
// attenuation; the anelastic tInt stays inside the local step.
#include <hip/hip_runtime.h>

#include "gemm_lds.hpp"
#include "kernels.hpp"

namespace edge {
namespace {

struct HalfViscoLds {
  size_t q, prev, cur, scr, scra, da, star, star_a, src, size;
  __host__ __device__ HalfViscoLds(int order, int runs, int nr, int tb) {
    const size_t N = n_modes(order), sz = 9 * N * runs, sza = 6 * N * runs;
    size_t o = 0;
    auto take = [&](size_t bytes) {
      const size_t at = o;
      o += (bytes + 15) / 16 * 16;
      return at;
    };
    q = take(tb * sz);
    prev = take(tb * sz);
    cur = take(tb * sz);
    scr = take(tb * sz);
    scra = take(tb * sza);
    da = take(tb * sza * nr);
    star = take(tb * 243);
    star_a = take(tb * 54);
    src = take(tb * 36 * nr);
    size = o;
  }
};

// One workgroup per listed element: t2 = sum_de h^{de+1} / (de+1)! der_de.
template <typename T, int O, int R>
__global__ __launch_bounds__(kThreads) void k_lts_half_visco(SolverDev<T> s, const int64_t* __restrict__ list, T h,
                                                             const T* __restrict__ dofs, const T* __restrict__ dofs_a,
                                                             T* __restrict__ t2) {
  constexpr int N = n_modes(O), NV = n_modes(O - 1), SZ = 9 * N * R, SZA = 6 * N * R;
  const int nr = s.n_rms;
  const HalfViscoLds L(O, R, nr, sizeof(T));
  extern __shared__ __align__(16) unsigned char lds[];
  T* l_prev = reinterpret_cast<T*>(lds + L.prev);
  T* l_cur = reinterpret_cast<T*>(lds + L.cur);
  T* l_scr = reinterpret_cast<T*>(lds + L.scr);
  T* l_scra = reinterpret_cast<T*>(lds + L.scra);
  T* l_da = reinterpret_cast<T*>(lds + L.da);
  T* l_star = reinterpret_cast<T*>(lds + L.star);
  T* l_star_a = reinterpret_cast<T*>(lds + L.star_a);
  T* l_src = reinterpret_cast<T*>(lds + L.src);
  T* l_out = reinterpret_cast<T*>(lds + L.q);
  const int64_t el = list[blockIdx.x];
  for (int i = threadIdx.x; i < SZ; i += kThreads) {
    const T v = dofs[el * SZ + i];
    l_prev[i] = v;
    l_out[i] = h * v;
  }
  for (int i = threadIdx.x; i < nr * SZA; i += kThreads) l_da[i] = dofs_a[el * nr * SZA + i];
  for (int i = threadIdx.x; i < 243; i += kThreads) l_star[i] = s.star[el * 243 + i];
  for (int i = threadIdx.x; i < 54; i += kThreads) l_star_a[i] = s.star_a[el * 54 + i];
  for (int i = threadIdx.x; i < 36 * nr; i += kThreads) l_src[i] = s.src_a[el * 36 * nr + i];
  T sca = h;
  for (int de = 1; de < O; de++) {
    __syncthreads();
    for (int i = threadIdx.x; i < SZ; i += kThreads) l_cur[i] = T(0);
    for (int i = threadIdx.x; i < SZA; i += kThreads) l_scra[i] = T(0);
    for (int d = 0; d < 3; d++) {
      __syncthreads();
      gemm_ac<T, R>(9, NV, N, N, N, N, false, l_prev, s.dg.stiffTn + d * N * N, l_scr);
      __syncthreads();
      gemm_bc<T, R>(9, NV, 9, 9, N, N, true, l_star + d * 81, l_scr, l_cur);
      gemm_bc<T, R>(6, NV, 3, 3, N, N, true, l_star_a + d * 18, l_scr + 6 * N * R, l_scra);
    }
    __syncthreads();
    for (int rm = 0; rm < nr; rm++) {
      gemm_bc<T, R>(6, N, 6, 6, N, N, true, l_src + rm * 36, l_da + rm * SZA, l_cur);
      __syncthreads();
    }
    for (int i = threadIdx.x; i < nr * SZA; i += kThreads) {
      const int rm = i / SZA;
      l_da[i] = s.rfs[rm] * (l_scra[i % SZA] + l_da[i]);
    }
    sca *= h / T(de + 1);
    for (int i = threadIdx.x; i < SZ; i += kThreads) l_out[i] += sca * l_cur[i];
    T* t = l_prev;
    l_prev = l_cur;
    l_cur = t;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < SZ; i += kThreads) t2[el * SZ + i] = l_out[i];
}

struct FacesViscoLds {
  size_t dof, tn, f0, f1, f1a, upa, fs, fsa, size;
  __host__ __device__ FacesViscoLds(int order, int runs, int tb) {
    const size_t N = n_modes(order), NF = n_face_modes(order);
    size_t o = 0;
    auto take = [&](size_t bytes) {
      const size_t at = o;
      o += (bytes + 15) / 16 * 16;
      return at;
    };
    dof = take(tb * 9 * N * runs);
    tn = take(tb * 9 * N * runs);
    f0 = take(tb * 9 * NF * runs);
    f1 = take(tb * 9 * NF * runs);
    f1a = take(tb * 6 * NF * runs);
    upa = take(tb * 6 * N * runs);
    fs = take(tb * 81);
    fsa = take(tb * 54);
    size = o;
  }
};

// The interface faces of the listed elements (one workgroup each).
template <typename T, int O, int R>
__global__ __launch_bounds__(kThreads) void k_lts_faces_visco(SolverDev<T> s, const uint8_t* __restrict__ fa_lts,
                                                              const int64_t* __restrict__ list, int first_ts,
                                                              const T* __restrict__ t0, const T* __restrict__ t1,
                                                              const T* __restrict__ t2, T* __restrict__ dofs,
                                                              T* __restrict__ dofs_a) {
  constexpr int N = n_modes(O), NF = n_face_modes(O), SZ = 9 * N * R, SZA = 6 * N * R;
  const FacesViscoLds L(O, R, sizeof(T));
  extern __shared__ __align__(16) unsigned char lds[];
  T* l_dof = reinterpret_cast<T*>(lds + L.dof);
  T* l_tn = reinterpret_cast<T*>(lds + L.tn);
  T* l_f0 = reinterpret_cast<T*>(lds + L.f0);
  T* l_f1 = reinterpret_cast<T*>(lds + L.f1);
  T* l_f1a = reinterpret_cast<T*>(lds + L.f1a);
  T* l_upa = reinterpret_cast<T*>(lds + L.upa);
  T* l_fs = reinterpret_cast<T*>(lds + L.fs);
  T* l_fsa = reinterpret_cast<T*>(lds + L.fsa);
  const int64_t el = list[blockIdx.x];
  for (int i = threadIdx.x; i < SZ; i += kThreads) l_dof[i] = dofs[el * SZ + i];
  for (int i = threadIdx.x; i < SZA; i += kThreads) l_upa[i] = T(0);
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
    for (int i = threadIdx.x; i < 54; i += kThreads) l_fsa[i] = s.fs_a[((s.n_el + el) * 4 + fa) * 54 + i];
    __syncthreads();
    gemm_ac<T, R>(9, NF, N, N, NF, NF, false, l_tn, s.dg.fluxLN + fm * N * NF, l_f0);
    __syncthreads();
    gemm_bc<T, R>(9, NF, 9, 9, NF, NF, false, l_fs, l_f0, l_f1);
    gemm_bc<T, R>(6, NF, 9, 9, NF, NF, false, l_fsa, l_f0, l_f1a);
    __syncthreads();
    gemm_ac<T, R>(9, N, NF, NF, N, N, true, l_f1, s.dg.fluxT + fa * NF * N, l_dof);
    gemm_ac<T, R>(6, N, NF, NF, N, N, true, l_f1a, s.dg.fluxT + fa * NF * N, l_upa);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < SZ; i += kThreads) dofs[el * SZ + i] = l_dof[i];
  for (int i = threadIdx.x; i < s.n_rms * SZA; i += kThreads) {
    const int rm = i / SZA;
    dofs_a[el * s.n_rms * SZA + i] += s.rfs[rm] * l_upa[i % SZA];
  }
}

template <typename T, int R>
int half_r(const SolverDev<T>& s, const int64_t* list, int64_t n, T h, const T* dofs, const T* dofs_a, T* t2,
           hipStream_t st) {
  auto go = [&](auto kern, int order) {
    const size_t lds = HalfViscoLds(order, R, s.n_rms, sizeof(T)).size;
    EDGE_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds));
    hipLaunchKernelGGL(kern, dim3((unsigned)n), dim3(kThreads), lds, st, s, list, h, dofs, dofs_a, t2);
    EDGE_HIP_TRY(hipGetLastError());
    return EDGE_OK;
  };
  switch (s.order) {
    case 2: return go(k_lts_half_visco<T, 2, R>, 2);
    case 3: return go(k_lts_half_visco<T, 3, R>, 3);
    case 4: return go(k_lts_half_visco<T, 4, R>, 4);
    case 5: return go(k_lts_half_visco<T, 5, R>, 5);
  }
  return fail(EDGE_ERR_UNSUPPORTED, "local time stepping: order");
}

template <typename T, int R>
int faces_r(const SolverDev<T>& s, const uint8_t* fa_lts, const int64_t* list, int64_t n, int first_ts, const T* t0,
            const T* t1, const T* t2, T* dofs, T* dofs_a, hipStream_t st) {
  auto go = [&](auto kern, int order) {
    const size_t lds = FacesViscoLds(order, R, sizeof(T)).size;
    EDGE_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds));
    hipLaunchKernelGGL(kern, dim3((unsigned)n), dim3(kThreads), lds, st, s, fa_lts, list, first_ts, t0, t1, t2, dofs,
                       dofs_a);
    EDGE_HIP_TRY(hipGetLastError());
    return EDGE_OK;
  };
  switch (s.order) {
    case 2: return go(k_lts_faces_visco<T, 2, R>, 2);
    case 3: return go(k_lts_faces_visco<T, 3, R>, 3);
    case 4: return go(k_lts_faces_visco<T, 4, R>, 4);
    case 5: return go(k_lts_faces_visco<T, 5, R>, 5);
  }
  return fail(EDGE_ERR_UNSUPPORTED, "local time stepping: order");
}

}  // namespace

bool lts_visco_supported(int order, int n_runs, int prec_bytes, int n_rms) {
  if (order < 2 || order > 5 || (n_runs != 1 && n_runs != 8) || n_rms < 1 || n_rms > 5) return false;
  return HalfViscoLds(order, n_runs, n_rms, prec_bytes).size <= 160 * 1024 &&
         FacesViscoLds(order, n_runs, prec_bytes).size <= 160 * 1024;
}

template <typename T>
int launch_lts_half_visco(const SolverDev<T>& s, const int64_t* list, int64_t n, T h, const T* dofs,
                          const T* dofs_a, T* t2, hipStream_t st) {
  if (n <= 0) return EDGE_OK;
  if (s.n_runs == 1) return half_r<T, 1>(s, list, n, h, dofs, dofs_a, t2, st);
  if (s.n_runs == 8) return half_r<T, 8>(s, list, n, h, dofs, dofs_a, t2, st);
  return fail(EDGE_ERR_UNSUPPORTED, "local time stepping: viscoelastic solvers take 1 or 8 runs");
}

template <typename T>
int launch_lts_faces_visco(const SolverDev<T>& s, const uint8_t* fa_lts, const int64_t* list, int64_t n, int first_ts,
                           const T* t0, const T* t1, const T* t2, T* dofs, T* dofs_a, hipStream_t st) {
  if (n <= 0) return EDGE_OK;
  if (s.n_runs == 1) return faces_r<T, 1>(s, fa_lts, list, n, first_ts, t0, t1, t2, dofs, dofs_a, st);
  if (s.n_runs == 8) return faces_r<T, 8>(s, fa_lts, list, n, first_ts, t0, t1, t2, dofs, dofs_a, st);
  return fail(EDGE_ERR_UNSUPPORTED, "local time stepping: viscoelastic solvers take 1 or 8 runs");
}

template int launch_lts_half_visco<float>(const SolverDev<float>&, const int64_t*, int64_t, float, const float*,
                                          const float*, float*, hipStream_t);
template int launch_lts_half_visco<double>(const SolverDev<double>&, const int64_t*, int64_t, double, const double*,
                                           const double*, double*, hipStream_t);
template int launch_lts_faces_visco<float>(const SolverDev<float>&, const uint8_t*, const int64_t*, int64_t, int,
                                           const float*, const float*, const float*, float*, float*, hipStream_t);
template int launch_lts_faces_visco<double>(const SolverDev<double>&, const uint8_t*, const int64_t*, int64_t, int,
                                            const double*, const double*, const double*, double*, double*,
                                            hipStream_t);

}  // namespace edge
