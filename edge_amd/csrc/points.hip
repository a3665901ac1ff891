// Receivers and point sources on the GPU.
//
// Receivers (src/impl/seismic/solvers/AderDg.hpp:475-496, src/io/Receivers.cpp:
// 235-290): at relative times tau in [0, dt) of a step, the time prediction of
// the receiver's element, evaluated from its Cauchy-Kowalevski derivatives
// (kernels/TimePred.hpp:130-156: Q(tau) = sum_de tau^de / de! der_de), at the
// receiver's point: out[q][r] = sum_md Q(tau)[q][md][r] phi_md(xi).  The
// derivatives follow the local step's CK (TimePredVanilla.hpp:200-306); they
// are taken from the DOFs at the start of the step, so a step evaluates its
// receivers before its local update.
//
// Point sources (src/impl/seismic/setups/PointSources.hpp:887-930): per step
// [t1, t2] and source, the integral of its sampled time series
// (linalg::Series<1>::integrate, src/linalg/Series.hpp:59-133) times the
// quantity scalings times the basis evaluations (scaled by the inverse mass
// and 1/|J|) is added to its element's DOFs of its run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
#include <numeric>
#include <vector>

#include "gemm_lds.hpp"
#include "kernels.hpp"

namespace edge {
namespace {

constexpr int kMaxTau = 8;  // output times per receiver and call

// One workgroup per receiver.  Dynamic LDS: der_0 (= Q, N modes), the odd
// derivatives (n_modes(O - 1) modes), the even ones from the second on
// (n_modes(O - 2)), the stiffness product (n_modes(O - 1)), the star
// matrices, the basis values and the FP64 output accumulators.
template <typename T, int O, int R>
struct RecvLds {
  static constexpr int N = n_modes(O), N1 = n_modes(O > 1 ? O - 1 : 1), N2 = n_modes(O > 2 ? O - 2 : 1);
  static constexpr size_t kQ = 0, kOdd = kQ + sizeof(T) * 9 * N * R, kEven = kOdd + sizeof(T) * 9 * N1 * R,
                          kScr = kEven + sizeof(T) * 9 * N2 * R, kStar = kScr + sizeof(T) * 9 * N1 * R,
                          kPhi = (kStar + sizeof(T) * 243 + 7) / 8 * 8, kAcc = kPhi + sizeof(double) * N,
                          kSize = kAcc + sizeof(double) * kMaxTau * 9 * R;
};

template <typename T, int O, int R>
__global__ __launch_bounds__(kThreads) void k_recv(SolverDev<T> s, const int64_t* __restrict__ el_ids,
                                                   const double* __restrict__ basis, int n_tau,
                                                   const double* __restrict__ tau, const T* __restrict__ dofs,
                                                   T* __restrict__ out) {
  using L = RecvLds<T, O, R>;
  constexpr int N = L::N, N1 = L::N1, N2 = L::N2;
  extern __shared__ __align__(16) unsigned char lds[];
  T* l_q = reinterpret_cast<T*>(lds + L::kQ);
  T* l_odd = reinterpret_cast<T*>(lds + L::kOdd);
  T* l_even = reinterpret_cast<T*>(lds + L::kEven);
  T* l_scr = reinterpret_cast<T*>(lds + L::kScr);
  T* l_star = reinterpret_cast<T*>(lds + L::kStar);
  double* l_phi = reinterpret_cast<double*>(lds + L::kPhi);
  double* l_acc = reinterpret_cast<double*>(lds + L::kAcc);
  const int64_t rc = blockIdx.x, el = el_ids[rc];
  for (int i = threadIdx.x; i < 9 * N * R; i += kThreads) l_q[i] = dofs[el * 9 * N * R + i];
  for (int i = threadIdx.x; i < 243; i += kThreads) l_star[i] = s.star[el * 243 + i];
  for (int i = threadIdx.x; i < N; i += kThreads) l_phi[i] = basis[rc * N + i];
  __syncthreads();
  // zeroth derivative: Q at the point, for every output time
  for (int i = threadIdx.x; i < 9 * R; i += kThreads) {
    const int q = i / R, r = i % R;
    double v = 0;
    for (int md = 0; md < N; md++) v += l_phi[md] * double(l_q[(q * N + md) * R + r]);
    for (int k = 0; k < n_tau; k++) l_acc[k * 9 * R + i] = v;
  }
  // der_de = sum_d S_d (der_{de-1} (-K_d^T M^-1)) over the modes CK keeps;
  // odd derivatives in one buffer, even ones in the other
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
    for (int i = threadIdx.x; i < 9 * R; i += kThreads) {
      const int q = i / R, r = i % R;
      double v = 0;
      for (int md = 0; md < nn; md++) v += l_phi[md] * double(cur[(q * ldc + md) * R + r]);
      double fac = 1.0;  // tau^de / de!
      for (int j = 1; j <= de; j++) fac /= j;
      for (int k = 0; k < n_tau; k++) {
        double c = fac;
        for (int j = 0; j < de; j++) c *= tau[rc * n_tau + k];
        l_acc[k * 9 * R + i] += c * v;
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n_tau * 9 * R; i += kThreads) out[rc * n_tau * 9 * R + i] = T(l_acc[i]);
}

template <typename T, int R>
int dispatch_recv(const SolverDev<T>& s, int64_t n, const int64_t* el, const double* basis, int n_tau,
                  const double* tau, const T* dofs, T* out, hipStream_t st) {
  auto go = [&](auto kern, size_t lds) {
    hipLaunchKernelGGL(kern, dim3((unsigned)n), dim3(kThreads), lds, st, s, el, basis, n_tau, tau, dofs, out);
    EDGE_HIP_TRY(hipGetLastError());
    return EDGE_OK;
  };
  switch (s.order) {
    case 2: return go(k_recv<T, 2, R>, RecvLds<T, 2, R>::kSize);
    case 3: return go(k_recv<T, 3, R>, RecvLds<T, 3, R>::kSize);
    case 4: return go(k_recv<T, 4, R>, RecvLds<T, 4, R>::kSize);
    case 5: return go(k_recv<T, 5, R>, RecvLds<T, 5, R>::kSize);
  }
  return fail(EDGE_ERR_UNSUPPORTED, "receivers: order");
}

// ---------------------------------------------------------------- sources
// Series<1>::integrate (src/linalg/Series.hpp:59-133), FP64, with the FP64
// tolerance of src/constants.hpp:289-300; one thread per source.
__global__ void k_src_integrate(int64_t n, const double* __restrict__ t0, const double* __restrict__ dts,
                                const int64_t* __restrict__ ptr, const double* __restrict__ ts, double t1, double t2,
                                double* __restrict__ ints) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double dx = dts[i], x0 = t0[i], tol = 1e-9;
  const double* v = ts + ptr[i];
  const int64_t np = ptr[i + 1] - ptr[i];
  const double a = (t1 - x0 + dx) / dx, b = (t2 - x0) / dx;
  int64_t first = (int64_t)(a > 0 ? a : 0), last = (int64_t)(b > 0 ? b : 0);
  if (last > np - 1) last = np - 1;
  if (first > last) first = last;
  double s = 0;
  for (int64_t k = first + 1; k <= last; k++) s += 0.5 * dx * (v[k] - v[k - 1]) + v[k - 1] * dx;
  const double t_end = x0 + (np - 1) * dx;
  const bool in1 = t1 + tol > x0 && t1 - tol < t_end, in2 = t2 + tol > x0 && t2 - tol < t_end;
  double d = x0 + first * dx - t1;
  if (first > 0 && in1) s += 0.5 * d * (d / dx) * (v[first - 1] - v[first]) + v[first] * d;
  d = t2 - (x0 + last * dx);
  if (last < np - 1 && in2) s += 0.5 * d * (d / dx) * (v[last + 1] - v[last]) + v[last] * d;
  ints[i] = s;  // outside the samples the series is zero
}

// One workgroup per source element: threads over its (q, mode, run) DOFs,
// each summing the element's sources of its run (no two workgroups share an
// element, so the update needs no atomics).
template <typename T>
__global__ __launch_bounds__(kThreads) void k_src_apply(int n_modes, int n_runs, const int64_t* __restrict__ sel_el,
                                                        const int64_t* __restrict__ sel_ptr,
                                                        const int32_t* __restrict__ run,
                                                        const double* __restrict__ bev,
                                                        const double* __restrict__ sca,
                                                        const double* __restrict__ ints, T* __restrict__ dofs) {
  const int64_t e = blockIdx.x, el = sel_el[e];
  const int sz = 9 * n_modes * n_runs;
  for (int i = threadIdx.x; i < sz; i += kThreads) {
    const int r = i % n_runs, md = (i / n_runs) % n_modes, q = i / (n_runs * n_modes);
    double acc = 0;
    bool any = false;
    for (int64_t k = sel_ptr[e]; k < sel_ptr[e + 1]; k++)
      if (run[k] == r) {
        acc += ints[k] * sca[k * 9 + q] * bev[k * n_modes + md];
        any = true;
      }
    if (any) dofs[el * sz + i] += T(acc);
  }
}

}  // namespace

namespace {

// ---------------------------------------------------------------- receivers with attenuation
// Viscoelastic solvers: the receiver's elastic derivatives come from the CK
// with the anelastic terms (TimePredVanilla.hpp:200-306 with TL_N_RMS > 0,
// all N modes per derivative): der_de = sum_d S_d (der_{de-1} (-K_d^T M^-1))
// + sum_rm E_rm derA_rm, derA_rm <- w_rm (sum_d S^A_d (der_{de-1} K)_vel +
// derA_rm).  Dynamic LDS: the previous and current derivative, the stiffness
// product, the anelastic combination, derA of every mechanism, the star /
// anelastic star / source matrices, the basis values and the accumulators.
struct RecvViscoLds {
  size_t prev, cur, scr, scra, da, star, star_a, src, phi, acc, size;
  __host__ __device__ RecvViscoLds(int order, int runs, int nr, int tb) {
    const size_t N = n_modes(order), sz = 9 * N * runs, sza = 6 * N * runs;
    size_t o = 0;
    auto take = [&](size_t bytes) {
      const size_t at = o;
      o += (bytes + 15) / 16 * 16;
      return at;
    };
    prev = take(tb * sz);
    cur = take(tb * sz);
    scr = take(tb * sz);
    scra = take(tb * sza);
    da = take(tb * sza * nr);
    star = take(tb * 243);
    star_a = take(tb * 54);
    src = take(tb * 36 * nr);
    phi = take(8 * N);
    acc = take(8 * kMaxTau * 9 * runs);
    size = o;
  }
};

template <typename T, int O, int R>
__global__ __launch_bounds__(kThreads) void k_recv_visco(SolverDev<T> s, const int64_t* __restrict__ el_ids,
                                                         const double* __restrict__ basis, int n_tau,
                                                         const double* __restrict__ tau, const T* __restrict__ dofs,
                                                         const T* __restrict__ dofs_a, T* __restrict__ out) {
  constexpr int N = n_modes(O), NV = n_modes(O - 1), SZ = 9 * N * R, SZA = 6 * N * R;
  const int nr = s.n_rms;
  const RecvViscoLds L(O, R, nr, sizeof(T));
  extern __shared__ __align__(16) unsigned char lds[];
  T* l_prev = reinterpret_cast<T*>(lds + L.prev);
  T* l_cur = reinterpret_cast<T*>(lds + L.cur);
  T* l_scr = reinterpret_cast<T*>(lds + L.scr);
  T* l_scra = reinterpret_cast<T*>(lds + L.scra);
  T* l_da = reinterpret_cast<T*>(lds + L.da);
  T* l_star = reinterpret_cast<T*>(lds + L.star);
  T* l_star_a = reinterpret_cast<T*>(lds + L.star_a);
  T* l_src = reinterpret_cast<T*>(lds + L.src);
  double* l_phi = reinterpret_cast<double*>(lds + L.phi);
  double* l_acc = reinterpret_cast<double*>(lds + L.acc);
  const int64_t rc = blockIdx.x, el = el_ids[rc];
  for (int i = threadIdx.x; i < SZ; i += kThreads) l_prev[i] = dofs[el * SZ + i];
  for (int i = threadIdx.x; i < nr * SZA; i += kThreads) l_da[i] = dofs_a[el * nr * SZA + i];
  for (int i = threadIdx.x; i < 243; i += kThreads) l_star[i] = s.star[el * 243 + i];
  for (int i = threadIdx.x; i < 54; i += kThreads) l_star_a[i] = s.star_a[el * 54 + i];
  for (int i = threadIdx.x; i < 36 * nr; i += kThreads) l_src[i] = s.src_a[el * 36 * nr + i];
  for (int i = threadIdx.x; i < N; i += kThreads) l_phi[i] = basis[rc * N + i];
  __syncthreads();
  for (int i = threadIdx.x; i < 9 * R; i += kThreads) {
    const int q = i / R, r = i % R;
    double v = 0;
    for (int md = 0; md < N; md++) v += l_phi[md] * double(l_prev[(q * N + md) * R + r]);
    for (int k = 0; k < n_tau; k++) l_acc[k * 9 * R + i] = v;
  }
  for (int de = 1; de < O; de++) {
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
    for (int i = threadIdx.x; i < 9 * R; i += kThreads) {
      const int q = i / R, r = i % R;
      double v = 0;
      for (int md = 0; md < N; md++) v += l_phi[md] * double(l_cur[(q * N + md) * R + r]);
      double fac = 1.0;  // tau^de / de!
      for (int j = 1; j <= de; j++) fac /= j;
      for (int k = 0; k < n_tau; k++) {
        double c = fac;
        for (int j = 0; j < de; j++) c *= tau[rc * n_tau + k];
        l_acc[k * 9 * R + i] += c * v;
      }
    }
    __syncthreads();
    T* t = l_prev;
    l_prev = l_cur;
    l_cur = t;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n_tau * 9 * R; i += kThreads) out[rc * n_tau * 9 * R + i] = T(l_acc[i]);
}

template <typename T, int R>
int dispatch_recv_visco(const SolverDev<T>& s, int64_t n, const int64_t* el, const double* basis, int n_tau,
                        const double* tau, const T* dofs, const T* dofs_a, T* out, hipStream_t st) {
  auto go = [&](auto kern, int order) {
    const size_t lds = RecvViscoLds(order, R, s.n_rms, sizeof(T)).size;
    EDGE_HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)lds));
    hipLaunchKernelGGL(kern, dim3((unsigned)n), dim3(kThreads), lds, st, s, el, basis, n_tau, tau, dofs, dofs_a,
                       out);
    EDGE_HIP_TRY(hipGetLastError());
    return EDGE_OK;
  };
  switch (s.order) {
    case 2: return go(k_recv_visco<T, 2, R>, 2);
    case 3: return go(k_recv_visco<T, 3, R>, 3);
    case 4: return go(k_recv_visco<T, 4, R>, 4);
    case 5: return go(k_recv_visco<T, 5, R>, 5);
  }
  return fail(EDGE_ERR_UNSUPPORTED, "receivers: order");
}

}  // namespace

bool recv_visco_supported(int order, int n_runs, int prec_bytes, int n_rms) {
  if (order < 2 || order > 5 || (n_runs != 1 && n_runs != 8) || n_rms < 1) return false;
  return RecvViscoLds(order, n_runs, n_rms, prec_bytes).size <= 160 * 1024;
}

template <typename T>
int launch_recv_eval_visco(const SolverDev<T>& s, int64_t n, const int64_t* el, const double* basis, int n_tau,
                           const double* tau, const T* dofs, const T* dofs_a, T* out, hipStream_t st) {
  if (n <= 0 || n_tau <= 0) return EDGE_OK;
  if (n_tau > kMaxTau) return fail(EDGE_ERR_ARG, "receivers: at most 8 output times per call");
  if (!recv_visco_supported(s.order, s.n_runs, sizeof(T), s.n_rms))
    return fail(EDGE_ERR_UNSUPPORTED, "receivers: configuration");
  if (s.n_runs == 1) return dispatch_recv_visco<T, 1>(s, n, el, basis, n_tau, tau, dofs, dofs_a, out, st);
  return dispatch_recv_visco<T, 8>(s, n, el, basis, n_tau, tau, dofs, dofs_a, out, st);
}

template int launch_recv_eval_visco<float>(const SolverDev<float>&, int64_t, const int64_t*, const double*, int,
                                           const double*, const float*, const float*, float*, hipStream_t);
template int launch_recv_eval_visco<double>(const SolverDev<double>&, int64_t, const int64_t*, const double*, int,
                                            const double*, const double*, const double*, double*, hipStream_t);

bool recv_supported(int order, int n_runs, int prec_bytes) {
  if (order < 2 || order > 5) return false;
  if (n_runs != 1 && n_runs != 8 && n_runs != 16) return false;
  const int N = n_modes(order), N1 = n_modes(order - 1), N2 = n_modes(order > 2 ? order - 2 : 1);
  const size_t lds = (size_t)prec_bytes * (9 * (N + 2 * N1 + N2) * n_runs + 243) + 8 * (N + kMaxTau * 9 * n_runs) + 8;
  return lds <= 64 * 1024;
}

template <typename T>
int launch_recv_eval(const SolverDev<T>& s, int64_t n, const int64_t* el, const double* basis, int n_tau,
                     const double* tau, const T* dofs, T* out, hipStream_t st) {
  if (n <= 0 || n_tau <= 0) return EDGE_OK;
  if (n_tau > kMaxTau) return fail(EDGE_ERR_ARG, "receivers: at most 8 output times per call");
  if (!recv_supported(s.order, s.n_runs, sizeof(T))) return fail(EDGE_ERR_UNSUPPORTED, "receivers: configuration");
  if (s.n_runs == 1) return dispatch_recv<T, 1>(s, n, el, basis, n_tau, tau, dofs, out, st);
  if (s.n_runs == 8) return dispatch_recv<T, 8>(s, n, el, basis, n_tau, tau, dofs, out, st);
  return dispatch_recv<T, 16>(s, n, el, basis, n_tau, tau, dofs, out, st);
}

template int launch_recv_eval<float>(const SolverDev<float>&, int64_t, const int64_t*, const double*, int,
                                     const double*, const float*, float*, hipStream_t);
template int launch_recv_eval<double>(const SolverDev<double>&, int64_t, const int64_t*, const double*, int,
                                      const double*, const double*, double*, hipStream_t);

}  // namespace edge

// ---------------------------------------------------------------- C ABI
struct edge_points {
  int kind = 0;  // 1 receivers, 2 sources
  int precision = 64, order = 0, n_runs = 0, n_modes = 0;
  int64_t n_el = 0, n = 0;
  // receivers
  int64_t* el = nullptr;   // [n]
  double* basis = nullptr; // [n][N]
  double* tau = nullptr;   // [n][8] staging of the output times
  // sources, sorted by element
  int64_t n_sel = 0;
  int64_t* sel_el = nullptr;   // [n_sel] source elements
  int64_t* sel_ptr = nullptr;  // [n_sel + 1] their sources
  int32_t* run = nullptr;      // [n]
  double* t0 = nullptr;        // [n]
  double* dts = nullptr;       // [n]
  double* bev = nullptr;       // [n][N]
  double* sca = nullptr;       // [n][9]
  int64_t* ts_ptr = nullptr;   // [n + 1]
  double* ts = nullptr;        // [ts_ptr[n]]
  double* ints = nullptr;      // [n] integrals of the current step
};

namespace {

int check_solver(const edge_points* p, const edge_el_solver* s, edge::SolverRef* r) {
  using edge::fail;
  if (!p || !s) return fail(EDGE_ERR_ARG, "points: handle missing");
  const int rc = edge::solver_ref(s, r);
  if (rc != EDGE_OK) return rc;
  if (r->order != p->order || r->n_runs != p->n_runs || r->n_el != p->n_el || r->precision != p->precision)
    return fail(EDGE_ERR_ARG, "points: solver differs from the one the points were created for");
  return EDGE_OK;
}

template <typename P>
hipError_t up(P** d, const P* h, size_t n) {
  hipError_t e = hipMalloc((void**)d, sizeof(P) * (n ? n : 1));
  if (e == hipSuccess && n) e = hipMemcpy(*d, h, sizeof(P) * n, hipMemcpyHostToDevice);
  return e;
}

}  // namespace

extern "C" int edge_points_destroy(edge_points* p) {
  if (!p) return EDGE_OK;
  for (void* d : {(void*)p->el, (void*)p->basis, (void*)p->tau, (void*)p->sel_el, (void*)p->sel_ptr, (void*)p->run,
                  (void*)p->t0, (void*)p->dts, (void*)p->bev, (void*)p->sca, (void*)p->ts_ptr, (void*)p->ts,
                  (void*)p->ints})
    if (d) (void)hipFree(d);
  delete p;
  return EDGE_OK;
}

extern "C" int edge_el_point_basis(const edge_el_solver* s, int64_t n, const int64_t* el, const double* xi,
                                   int32_t src_scaled, double* out) {
  using edge::fail;
  edge::SolverRef r;
  if (!s || !el || !xi || !out) return fail(EDGE_ERR_ARG, "point basis: argument missing");
  int rc = edge::solver_ref(s, &r);
  if (rc != EDGE_OK) return rc;
  const int N = r.n_modes;
  for (int64_t i = 0; i < n; i++) {
    if (el[i] < 0 || el[i] >= r.n_el) return fail(EDGE_ERR_ARG, "point basis: element out of range");
    for (int d = 0; d < 3; d++)
      if (!(xi[i * 3 + d] > -1e-9 && xi[i * 3 + d] < 1 + 1e-9)) return fail(EDGE_ERR_ARG, "point basis: point outside the reference tet");
    if (xi[i * 3] + xi[i * 3 + 1] + xi[i * 3 + 2] > 1 + 1e-9)
      return fail(EDGE_ERR_ARG, "point basis: point outside the reference tet");
  }
  rc = edge_dg_eval_basis(r.order, n, xi, out);
  if (rc != EDGE_OK || !src_scaled || n == 0) return rc;
  if (!r.jdet) return fail(EDGE_ERR_UNSUPPORTED, "point basis: the solver has no geometry (caller-given solvers)");
  std::vector<double> mass(N), jd(r.n_el);
  EDGE_HIP_TRY(hipMemcpy(mass.data(), r.mass, sizeof(double) * N, hipMemcpyDeviceToHost));
  EDGE_HIP_TRY(hipMemcpy(jd.data(), r.jdet, sizeof(double) * r.n_el, hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < n; i++)
    for (int md = 0; md < N; md++) out[i * N + md] *= 1.0 / mass[md] / jd[el[i]];
  return EDGE_OK;
}

extern "C" int edge_recv_create(const edge_el_solver* s, int64_t n, const int64_t* el, const double* basis,
                                edge_points** out) {
  using edge::fail;
  if (out) *out = nullptr;
  edge::SolverRef r;
  if (!s || !out || (n > 0 && (!el || !basis))) return fail(EDGE_ERR_ARG, "receivers: argument missing");
  int rc = edge::solver_ref(s, &r);
  if (rc != EDGE_OK) return rc;
  if (n < 0) return fail(EDGE_ERR_ARG, "receivers: negative count");
  if (r.n_rms > 0 ? !edge::recv_visco_supported(r.order, r.n_runs, r.precision / 8, r.n_rms)
                  : !edge::recv_supported(r.order, r.n_runs, r.precision / 8))
    return fail(EDGE_ERR_UNSUPPORTED, "receivers: order, runs and precision do not fit one workgroup's LDS");
  for (int64_t i = 0; i < n; i++)
    if (el[i] < 0 || el[i] >= r.n_el) return fail(EDGE_ERR_ARG, "receivers: element out of range");
  edge_points* p = new (std::nothrow) edge_points();
  if (!p) return fail(EDGE_ERR_NOMEM, "receivers");
  p->kind = 1;
  p->precision = r.precision;
  p->order = r.order;
  p->n_runs = r.n_runs;
  p->n_modes = r.n_modes;
  p->n_el = r.n_el;
  p->n = n;
  hipError_t e = up(&p->el, el, n);
  if (e == hipSuccess) e = up(&p->basis, basis, n * r.n_modes);
  if (e == hipSuccess) e = hipMalloc((void**)&p->tau, sizeof(double) * 8 * (n ? n : 1));
  if (e != hipSuccess) {
    edge_points_destroy(p);
    return fail(EDGE_ERR_HIP, std::string("receivers: ") + hipGetErrorString(e));
  }
  *out = p;
  return EDGE_OK;
}

extern "C" int edge_recv_eval(edge_points* p, const edge_el_solver* s, const void* dofs, int32_t n_tau,
                              const double* tau, void* out, void* stream) {
  using edge::fail;
  edge::SolverRef r;
  int rc = check_solver(p, s, &r);
  if (rc != EDGE_OK) return rc;
  if (p->kind != 1) return fail(EDGE_ERR_ARG, "points: not receivers");
  if (r.n_rms > 0) return fail(EDGE_ERR_ARG, "receivers: a viscoelastic solver needs dofs_a (edge_recv_eval_a)");
  if (!dofs || !out || (n_tau > 0 && !tau)) return fail(EDGE_ERR_ARG, "receivers: argument missing");
  if (n_tau < 0 || n_tau > 8) return fail(EDGE_ERR_ARG, "receivers: 0 to 8 output times per call");
  if (p->n == 0 || n_tau == 0) return EDGE_OK;
  hipStream_t st = (hipStream_t)stream;
  EDGE_HIP_TRY(hipMemcpyAsync(p->tau, tau, sizeof(double) * p->n * n_tau, hipMemcpyHostToDevice, st));
  if (r.precision == 64)
    return edge::launch_recv_eval<double>(r.d, p->n, p->el, p->basis, n_tau, p->tau, (const double*)dofs,
                                          (double*)out, st);
  return edge::launch_recv_eval<float>(r.f, p->n, p->el, p->basis, n_tau, p->tau, (const float*)dofs, (float*)out,
                                       st);
}

extern "C" int edge_recv_eval_a(edge_points* p, const edge_el_solver* s, const void* dofs, const void* dofs_a,
                                int32_t n_tau, const double* tau, void* out, void* stream) {
  using edge::fail;
  edge::SolverRef r;
  int rc = check_solver(p, s, &r);
  if (rc != EDGE_OK) return rc;
  if (r.n_rms == 0) return edge_recv_eval(p, s, dofs, n_tau, tau, out, stream);
  if (p->kind != 1) return fail(EDGE_ERR_ARG, "points: not receivers");
  if (!dofs || !dofs_a || !out || (n_tau > 0 && !tau)) return fail(EDGE_ERR_ARG, "receivers: argument missing");
  if (n_tau < 0 || n_tau > 8) return fail(EDGE_ERR_ARG, "receivers: 0 to 8 output times per call");
  if (p->n == 0 || n_tau == 0) return EDGE_OK;
  hipStream_t st = (hipStream_t)stream;
  EDGE_HIP_TRY(hipMemcpyAsync(p->tau, tau, sizeof(double) * p->n * n_tau, hipMemcpyHostToDevice, st));
  if (r.precision == 64)
    return edge::launch_recv_eval_visco<double>(r.d, p->n, p->el, p->basis, n_tau, p->tau, (const double*)dofs,
                                                (const double*)dofs_a, (double*)out, st);
  return edge::launch_recv_eval_visco<float>(r.f, p->n, p->el, p->basis, n_tau, p->tau, (const float*)dofs,
                                             (const float*)dofs_a, (float*)out, st);
}

extern "C" int edge_src_create(const edge_el_solver* s, int64_t n, const int64_t* el, const int32_t* run,
                               const double* t0, const double* dts, const double* bev, const double* sca,
                               const int64_t* ts_ptr, const double* ts, edge_points** out) {
  using edge::fail;
  if (out) *out = nullptr;
  edge::SolverRef r;
  if (!s || !out || (n > 0 && (!el || !run || !t0 || !dts || !bev || !sca || !ts_ptr || !ts)))
    return fail(EDGE_ERR_ARG, "sources: argument missing");
  int rc = edge::solver_ref(s, &r);
  if (rc != EDGE_OK) return rc;
  if (n < 0) return fail(EDGE_ERR_ARG, "sources: negative count");
  if (n > 0 && ts_ptr[0] != 0) return fail(EDGE_ERR_ARG, "sources: series offsets must start at 0");
  for (int64_t i = 0; i < n; i++) {
    if (el[i] < 0 || el[i] >= r.n_el) return fail(EDGE_ERR_ARG, "sources: element out of range");
    if (run[i] < 0 || run[i] >= r.n_runs) return fail(EDGE_ERR_ARG, "sources: run out of range");
    if (!(dts[i] > 0)) return fail(EDGE_ERR_ARG, "sources: sampling steps must be positive");
    if (ts_ptr[i + 1] - ts_ptr[i] < 1) return fail(EDGE_ERR_ARG, "sources: every series needs a sample");
  }
  // sources sorted by element (stable), grouped per source element
  std::vector<int64_t> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) { return el[a] < el[b]; });
  const int N = r.n_modes;
  std::vector<int64_t> sel_el, sel_ptr{0}, h_ptr{0};
  std::vector<int32_t> h_run(n);
  std::vector<double> h_t0(n), h_dts(n), h_bev((size_t)n * N), h_sca((size_t)n * 9), h_ts;
  for (int64_t k = 0; k < n; k++) {
    const int64_t i = ord[k];
    if (k == 0 || el[i] != sel_el.back()) {
      if (k > 0) sel_ptr.push_back(k);
      sel_el.push_back(el[i]);
    }
    h_run[k] = run[i];
    h_t0[k] = t0[i];
    h_dts[k] = dts[i];
    std::copy(bev + i * N, bev + (i + 1) * N, h_bev.begin() + k * N);
    std::copy(sca + i * 9, sca + (i + 1) * 9, h_sca.begin() + k * 9);
    h_ts.insert(h_ts.end(), ts + ts_ptr[i], ts + ts_ptr[i + 1]);
    h_ptr.push_back((int64_t)h_ts.size());
  }
  if (n > 0) sel_ptr.push_back(n);
  edge_points* p = new (std::nothrow) edge_points();
  if (!p) return fail(EDGE_ERR_NOMEM, "sources");
  p->kind = 2;
  p->precision = r.precision;
  p->order = r.order;
  p->n_runs = r.n_runs;
  p->n_modes = N;
  p->n_el = r.n_el;
  p->n = n;
  p->n_sel = (int64_t)sel_el.size();
  hipError_t e = up(&p->sel_el, sel_el.data(), sel_el.size());
  if (e == hipSuccess) e = up(&p->sel_ptr, sel_ptr.data(), sel_ptr.size());
  if (e == hipSuccess) e = up(&p->run, h_run.data(), n);
  if (e == hipSuccess) e = up(&p->t0, h_t0.data(), n);
  if (e == hipSuccess) e = up(&p->dts, h_dts.data(), n);
  if (e == hipSuccess) e = up(&p->bev, h_bev.data(), h_bev.size());
  if (e == hipSuccess) e = up(&p->sca, h_sca.data(), h_sca.size());
  if (e == hipSuccess) e = up(&p->ts_ptr, h_ptr.data(), h_ptr.size());
  if (e == hipSuccess) e = up(&p->ts, h_ts.data(), h_ts.size());
  if (e == hipSuccess) e = hipMalloc((void**)&p->ints, sizeof(double) * (n ? n : 1));
  if (e != hipSuccess) {
    edge_points_destroy(p);
    return fail(EDGE_ERR_HIP, std::string("sources: ") + hipGetErrorString(e));
  }
  *out = p;
  return EDGE_OK;
}

extern "C" int edge_src_apply(edge_points* p, const edge_el_solver* s, double t1, double t2, void* dofs,
                              void* stream) {
  using edge::fail;
  edge::SolverRef r;
  int rc = check_solver(p, s, &r);
  if (rc != EDGE_OK) return rc;
  if (p->kind != 2) return fail(EDGE_ERR_ARG, "points: not sources");
  if (!dofs) return fail(EDGE_ERR_ARG, "sources: DOFs missing");
  if (p->n == 0) return EDGE_OK;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(edge::k_src_integrate, dim3((unsigned)((p->n + 127) / 128)), dim3(128), 0, st, p->n, p->t0,
                     p->dts, p->ts_ptr, p->ts, t1, t2, p->ints);
  EDGE_HIP_TRY(hipGetLastError());
  if (r.precision == 64)
    hipLaunchKernelGGL(edge::k_src_apply<double>, dim3((unsigned)p->n_sel), dim3(edge::kThreads), 0, st, p->n_modes,
                       p->n_runs, p->sel_el, p->sel_ptr, p->run, p->bev, p->sca, p->ints, (double*)dofs);
  else
    hipLaunchKernelGGL(edge::k_src_apply<float>, dim3((unsigned)p->n_sel), dim3(edge::kThreads), 0, st, p->n_modes,
                       p->n_runs, p->sel_el, p->sel_ptr, p->run, p->bev, p->sca, p->ints, (float*)dofs);
  EDGE_HIP_TRY(hipGetLastError());
  return EDGE_OK;
}
