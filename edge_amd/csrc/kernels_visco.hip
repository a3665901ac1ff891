// Viscoelastic element kernels (N_RMS relaxation mechanisms): one workgroup
// per element with the elastic and anelastic tiles in LDS, threads over the
// outputs of every small GEMM, as the generic elastic kernels.
//
// Anelastic DOFs per element: [rm][6][N][R] (memory variables of the six
// stresses per mechanism).  Per stage of the local step (reference loops of
// src/impl/seismic/solvers/AderDg.hpp:362-533):
//   CK (kernels/TimePredVanilla.hpp:200-306, no mode truncation with
//   attenuation): per direction d
//     X = der_{de-1} (-K_d^T M^-1),  der_de += S_d X,  A += S^A_d X_vel
//   then per mechanism rm
//     der_de[stress] += E_rm derA_{rm,de-1},  derA_{rm,de} = w_rm (A + derA_{rm,de-1}),
//   and tInt / tIntA accumulate dt^{de+1}/(de+1)! of both;
//   volume (kernels/VolIntVanilla.hpp:189-241): Q += sum_d S_d (tInt K_d M^-1),
//     A = sum_d S^A_d (tInt K_d M^-1)_vel, Q[stress] += E_rm tIntA_rm,
//     QA_rm += w_rm (A - tIntA_rm);
//   local surface (kernels/SurfIntVanilla.hpp:212-262, SurfInt.hpp:139-156):
//     Q += (fs0 (tInt F^L)) F^T, QA_rm += w_rm (fsA0 (tInt F^L)) F^T.
// Neighbour step (SurfIntVanilla.hpp:274-357): Q += (fs1 (tInt_ne F^N)) F^T,
//   QA_rm += w_rm (fsA1 (tInt_ne F^N)) F^T.
#include <hip/hip_runtime.h>

#include "gemm_lds.hpp"
#include "kernels.hpp"

namespace edge {
namespace {

// LDS layout (elements of T) of the local kernel for nr mechanisms
struct ViscoLds {
  int64_t sz, sza, szf, dof, der0, der1, scr, tint, f0, f1, star, fs, star_a, fs_a, src, dera0, dera1, tinta, scra,
      upa, total;
  __host__ __device__ ViscoLds(int order, int runs, int nr) {
    const int64_t n = n_modes(order), nf = n_face_modes(order);
    sz = 9 * n * runs;
    sza = 6 * n * runs;
    szf = 9 * nf * runs;
    int64_t o = 0;
    auto take = [&](int64_t k) {
      const int64_t at = o;
      o += (k + 1) & ~int64_t(1);  // keep 8-byte alignment for any T
      return at;
    };
    dof = take(sz);
    der0 = take(sz);
    der1 = take(sz);
    scr = take(sz);
    tint = take(sz);
    f0 = take(szf);
    f1 = take(szf);
    star = take(243);
    fs = take(324);
    star_a = take(54);
    fs_a = take(216);
    src = take(36 * nr);
    dera0 = take(sza * nr);
    dera1 = take(sza * nr);
    tinta = take(sza * nr);
    scra = take(sza);
    upa = take(sza);
    total = o;
  }
};

template <typename T, int O, int R>
__global__ __launch_bounds__(kThreads) void k_local_visco(SolverDev<T> s, int64_t first, T dt, T* __restrict__ dofs,
                                                          T* __restrict__ dofs_a, T* __restrict__ t_int) {
  constexpr int N = n_modes(O), NF = n_face_modes(O), NV = n_modes(O - 1);
  constexpr int SZ = 9 * N * R, SZA = 6 * N * R, NQA = 6;
  extern __shared__ __align__(16) unsigned char smem[];
  T* lds = reinterpret_cast<T*>(smem);
  const int nr = s.n_rms;
  const ViscoLds L(O, R, nr);
  T *l_dof = lds + L.dof, *l_der[2] = {lds + L.der0, lds + L.der1}, *l_scr = lds + L.scr, *l_tint = lds + L.tint;
  T *l_f0 = lds + L.f0, *l_f1 = lds + L.f1, *l_star = lds + L.star, *l_fs = lds + L.fs, *l_star_a = lds + L.star_a;
  T *l_fs_a = lds + L.fs_a, *l_src = lds + L.src, *l_dera[2] = {lds + L.dera0, lds + L.dera1};
  T *l_tinta = lds + L.tinta, *l_scra = lds + L.scra, *l_upa = lds + L.upa;
  const int64_t el = first + blockIdx.x;
  const T* g_dof = dofs + el * SZ;
  T* g_dofa = dofs_a + el * nr * SZA;
  for (int i = threadIdx.x; i < SZ; i += kThreads) {
    const T v = g_dof[i];
    l_dof[i] = v;
    l_der[0][i] = v;
    l_tint[i] = dt * v;
  }
  for (int i = threadIdx.x; i < nr * SZA; i += kThreads) {
    const T v = g_dofa[i];
    l_dera[0][i] = v;
    l_tinta[i] = dt * v;
  }
  for (int i = threadIdx.x; i < 243; i += kThreads) l_star[i] = s.star[el * 243 + i];
  for (int i = threadIdx.x; i < 324; i += kThreads) l_fs[i] = s.fs[el * 324 + i];
  for (int i = threadIdx.x; i < 54; i += kThreads) l_star_a[i] = s.star_a[el * 54 + i];
  for (int i = threadIdx.x; i < 216; i += kThreads) l_fs_a[i] = s.fs_a[el * 216 + i];
  for (int i = threadIdx.x; i < 36 * nr; i += kThreads) l_src[i] = s.src_a[el * 36 * nr + i];
  __syncthreads();

  // Cauchy-Kowalevski with the anelastic terms
  T sca = dt;
  for (int de = 1; de < O; de++) {
    T* prev = l_der[(de - 1) & 1];
    T* cur = l_der[de & 1];
    T* pa = l_dera[(de - 1) & 1];
    T* ca = l_dera[de & 1];
    for (int i = threadIdx.x; i < SZ; i += kThreads) cur[i] = T(0);
    for (int i = threadIdx.x; i < SZA; i += kThreads) l_scra[i] = T(0);
    __syncthreads();
    for (int d = 0; d < 3; d++) {
      gemm_ac<T, R>(9, NV, N, N, N, N, false, prev, s.dg.stiffTn + d * N * N, l_scr);
      __syncthreads();
      gemm_bc<T, R>(9, NV, 9, 9, N, N, true, l_star + d * 81, l_scr, cur);
      gemm_bc<T, R>(NQA, NV, 3, 3, N, N, true, l_star_a + d * 18, l_scr + NQA * N * R, l_scra);
      __syncthreads();
    }
    sca *= dt / T(de + 1);
    for (int rm = 0; rm < nr; rm++) {
      gemm_bc<T, R>(NQA, N, NQA, NQA, N, N, true, l_src + rm * 36, pa + rm * SZA, cur);
      __syncthreads();
    }
    for (int i = threadIdx.x; i < nr * SZA; i += kThreads) {
      const int rm = i / SZA;
      const T v = s.rfs[rm] * (l_scra[i % SZA] + pa[i]);
      ca[i] = v;
      l_tinta[i] += sca * v;
    }
    for (int i = threadIdx.x; i < SZ; i += kThreads) l_tint[i] += sca * cur[i];
    __syncthreads();
  }
  T* g_tint = t_int + el * SZ;
  for (int i = threadIdx.x; i < SZ; i += kThreads) g_tint[i] = l_tint[i];

  // volume integral
  for (int i = threadIdx.x; i < SZA; i += kThreads) l_scra[i] = T(0);
  __syncthreads();
  for (int d = 0; d < 3; d++) {
    gemm_ac<T, R>(9, N, NV, N, N, N, false, l_tint, s.dg.stiffV + d * N * N, l_scr);
    __syncthreads();
    gemm_bc<T, R>(9, N, 9, 9, N, N, true, l_star + d * 81, l_scr, l_dof);
    gemm_bc<T, R>(NQA, N, 3, 3, N, N, true, l_star_a + d * 18, l_scr + NQA * N * R, l_scra);
    __syncthreads();
  }
  for (int rm = 0; rm < nr; rm++) {
    gemm_bc<T, R>(NQA, N, NQA, NQA, N, N, true, l_src + rm * 36, l_tinta + rm * SZA, l_dof);
    __syncthreads();
  }
  // local surface integral; the anelastic update sums in l_upa
  for (int i = threadIdx.x; i < SZA; i += kThreads) l_upa[i] = T(0);
  __syncthreads();
  for (int fa = 0; fa < 4; fa++) {
    gemm_ac<T, R>(9, NF, N, N, NF, NF, false, l_tint, s.dg.fluxLN + fa * N * NF, l_f0);
    __syncthreads();
    gemm_bc<T, R>(9, NF, 9, 9, NF, NF, false, l_fs + fa * 81, l_f0, l_f1);
    __syncthreads();
    gemm_ac<T, R>(9, N, NF, NF, N, N, true, l_f1, s.dg.fluxT + fa * NF * N, l_dof);
    __syncthreads();
    gemm_bc<T, R>(NQA, NF, 9, 9, NF, NF, false, l_fs_a + fa * 54, l_f0, l_f1);
    __syncthreads();
    gemm_ac<T, R>(NQA, N, NF, NF, N, N, true, l_f1, s.dg.fluxT + fa * NF * N, l_upa);
    __syncthreads();
  }
  T* g_out = dofs + el * SZ;
  for (int i = threadIdx.x; i < SZ; i += kThreads) g_out[i] = l_dof[i];
  // QA_rm += w_rm (A - tIntA_rm) from the volume and w_rm upA from the surface
  for (int i = threadIdx.x; i < nr * SZA; i += kThreads) {
    const int rm = i / SZA, j = i % SZA;
    g_dofa[i] += s.rfs[rm] * (l_scra[j] - l_tinta[i] + l_upa[j]);
  }
}

template <typename T, int O, int R>
__global__ __launch_bounds__(kThreads) void k_neigh_visco(SolverDev<T> s, int64_t first, const T* __restrict__ t_int,
                                                          T* __restrict__ dofs, T* __restrict__ dofs_a) {
  constexpr int N = n_modes(O), NF = n_face_modes(O), SZ = 9 * N * R, SZA = 6 * N * R, SZF = 9 * NF * R;
  constexpr int NQA = 6;
  __shared__ T l_dof[SZ], l_tn[SZ], l_f0[SZF], l_f1[SZF], l_fs[81], l_fs_a[54], l_upa[SZA];
  const int64_t el = first + blockIdx.x;
  const int nr = s.n_rms;
  for (int i = threadIdx.x; i < SZ; i += kThreads) l_dof[i] = dofs[el * SZ + i];
  for (int i = threadIdx.x; i < SZA; i += kThreads) l_upa[i] = T(0);
  for (int fa = 0; fa < 4; fa++) {
    const int32_t ne = s.nb.ne[el * 4 + fa];
    if (ne < 0) continue;  // outflow (uniform per block)
    const int fm = s.nb.fmat[el * 4 + fa];
    __syncthreads();
    for (int i = threadIdx.x; i < SZ; i += kThreads) l_tn[i] = t_int[(int64_t)ne * SZ + i];
    for (int i = threadIdx.x; i < 81; i += kThreads) l_fs[i] = s.fs[((s.n_el + el) * 4 + fa) * 81 + i];
    for (int i = threadIdx.x; i < 54; i += kThreads) l_fs_a[i] = s.fs_a[((s.n_el + el) * 4 + fa) * 54 + i];
    __syncthreads();
    gemm_ac<T, R>(9, NF, N, N, NF, NF, false, l_tn, s.dg.fluxLN + fm * N * NF, l_f0);
    __syncthreads();
    gemm_bc<T, R>(9, NF, 9, 9, NF, NF, false, l_fs, l_f0, l_f1);
    __syncthreads();
    gemm_ac<T, R>(9, N, NF, NF, N, N, true, l_f1, s.dg.fluxT + fa * NF * N, l_dof);
    __syncthreads();
    gemm_bc<T, R>(NQA, NF, 9, 9, NF, NF, false, l_fs_a, l_f0, l_f1);
    __syncthreads();
    gemm_ac<T, R>(NQA, N, NF, NF, N, N, true, l_f1, s.dg.fluxT + fa * NF * N, l_upa);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < SZ; i += kThreads) dofs[el * SZ + i] = l_dof[i];
  T* g_dofa = dofs_a + el * nr * SZA;
  for (int i = threadIdx.x; i < nr * SZA; i += kThreads) g_dofa[i] += s.rfs[i / SZA] * l_upa[i % SZA];
}

template <typename T, int O, int R>
constexpr bool neigh_fits() {
  constexpr size_t n = n_modes(O), nf = n_face_modes(O);
  return sizeof(T) * (2 * 9 * n * R + 2 * 9 * nf * R + 81 + 54 + 6 * n * R) <= 64 * 1024;
}

template <typename T, int O, int R>
int run_visco(const SolverDev<T>& s, int64_t first, int64_t n, T dt, T* dofs, T* dofs_a, T* t_int, const T* tin,
              bool local, hipStream_t st) {
  if constexpr (neigh_fits<T, O, R>()) {
    if (local) {
      const size_t lds = sizeof(T) * ViscoLds(O, R, s.n_rms).total;
      EDGE_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_local_visco<T, O, R>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      hipLaunchKernelGGL((k_local_visco<T, O, R>), dim3((unsigned)n), dim3(kThreads), lds, st, s, first, dt, dofs,
                         dofs_a, t_int);
    } else {
      hipLaunchKernelGGL((k_neigh_visco<T, O, R>), dim3((unsigned)n), dim3(kThreads), 0, st, s, first, tin, dofs,
                         dofs_a);
    }
    EDGE_HIP_TRY(hipGetLastError());
    return EDGE_OK;
  } else {
    return fail(EDGE_ERR_UNSUPPORTED, "viscoelastic kernel: tiles do not fit LDS");
  }
}

template <typename T, int R>
int dispatch_visco(const SolverDev<T>& s, int64_t first, int64_t n, T dt, T* dofs, T* dofs_a, T* t_int, const T* tin,
                   bool local, hipStream_t st) {
  switch (s.order) {
    case 2: return run_visco<T, 2, R>(s, first, n, dt, dofs, dofs_a, t_int, tin, local, st);
    case 3: return run_visco<T, 3, R>(s, first, n, dt, dofs, dofs_a, t_int, tin, local, st);
    case 4: return run_visco<T, 4, R>(s, first, n, dt, dofs, dofs_a, t_int, tin, local, st);
    case 5: return run_visco<T, 5, R>(s, first, n, dt, dofs, dofs_a, t_int, tin, local, st);
  }
  return fail(EDGE_ERR_UNSUPPORTED, "order");
}

template <typename T>
int launch_visco(const SolverDev<T>& s, int64_t first, int64_t n, T dt, T* dofs, T* dofs_a, T* t_int, const T* tin,
                 bool local, hipStream_t st) {
  if (n <= 0) return EDGE_OK;
  if (s.n_runs == 1) return dispatch_visco<T, 1>(s, first, n, dt, dofs, dofs_a, t_int, tin, local, st);
  if (s.n_runs == 8) return dispatch_visco<T, 8>(s, first, n, dt, dofs, dofs_a, t_int, tin, local, st);
  return fail(EDGE_ERR_UNSUPPORTED, "viscoelastic kernels: 1 or 8 fused runs");
}

}  // namespace

bool visco_supported(int order, int n_runs, int prec_bytes, int n_rms) {
  if (order < 2 || order > 5 || (n_runs != 1 && n_runs != 8) || n_rms < 1) return false;
  return (size_t)prec_bytes * ViscoLds(order, n_runs, n_rms).total <= 160 * 1024;
}

template <typename T>
int launch_local_visco(const SolverDev<T>& s, int64_t first, int64_t n, T dt, T* dofs, T* dofs_a, T* t_int,
                       hipStream_t st) {
  return launch_visco<T>(s, first, n, dt, dofs, dofs_a, t_int, nullptr, true, st);
}

template <typename T>
int launch_neigh_visco(const SolverDev<T>& s, int64_t first, int64_t n, const T* t_int, T* dofs, T* dofs_a,
                       hipStream_t st) {
  return launch_visco<T>(s, first, n, T(0), dofs, dofs_a, nullptr, t_int, false, st);
}

template int launch_local_visco<double>(const SolverDev<double>&, int64_t, int64_t, double, double*, double*,
                                        double*, hipStream_t);
template int launch_local_visco<float>(const SolverDev<float>&, int64_t, int64_t, float, float*, float*, float*,
                                       hipStream_t);
template int launch_neigh_visco<double>(const SolverDev<double>&, int64_t, int64_t, const double*, double*, double*,
                                        hipStream_t);
template int launch_neigh_visco<float>(const SolverDev<float>&, int64_t, int64_t, const float*, float*, float*,
                                       hipStream_t);

}  // namespace edge
