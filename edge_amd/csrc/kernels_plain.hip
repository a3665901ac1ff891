// FP64 single-run element kernels on FP64 MFMA (v_mfma_f64_16x16x4f64).
//
// With one run there is no run axis to fill an MFMA tile, so the quantities
// take the M dimension (9 of 16 rows) and the modes the N dimension: every
// product of an element's data X [q][k] with a shared reference-element
// matrix B [k][n] (stiffness, flux) is
//   A operand  lane (g, c) = X[c][4 s + g]          (c < 9, element data)
//   B operand  lane (g, c) = B[4 s + g][16 t + c]   (shared matrix)
//   D          lane (g, c), register v = C[g + 4 v][16 t + c]
// (the FP64 instruction interleaves the rows differently from the FP32 one).
// The per-element 9 x 9 matrices (star matrices, flux solvers) act on the
// quantity axis, across lanes: they are applied on the A side through LDS
// (S_d (X K_d) = (S_d X) K_d), with the elastic star's 24 non-zeros only.
// One wave per element.
//
// Operations (src/impl/seismic/solvers/AderDg.hpp:362-533, :553-722):
//   local: CK time prediction (TimePredVanilla.hpp:200-306), volume
//          integral (VolIntVanilla.hpp:189-241), local surface integral
//          (SurfIntVanilla.hpp:212-262);
//   neigh: neighbouring surface integral (SurfIntVanilla.hpp:274-357).
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace edge {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4d mfma64(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int tiles16(int n) { return (n + 15) / 16; }
constexpr int ksteps4(int n) { return (n + 3) / 4; }

__device__ __forceinline__ v4d zero4d() { return v4d{0.0, 0.0, 0.0, 0.0}; }

// Guarded loads without branches: clamped index, value masked afterwards.
__device__ __forceinline__ double ld(const double* __restrict__ p, int idx, bool ok) {
  const double v = p[ok ? idx : 0];
  return ok ? v : 0.0;
}

// acc[t] += A B for A [9][NK] (leading dimension lda, LDS or global) and
// B [NK][NN] (leading dimension ldb, global), output tiles t < NT; the
// operands of a k-step are loaded one k-step ahead
template <int NT, int NK, int NN>
__device__ __forceinline__ void mm(const double* __restrict__ a, int lda, const double* __restrict__ b, int ldb,
                                   v4d (&acc)[NT], int g, int c) {
  constexpr int KS = ksteps4(NK);
  double an, bn[NT];
  auto load = [&](int s) {
    const int k = 4 * s + g;
    an = ld(a, c * lda + k, c < 9 && k < NK);
#pragma unroll
    for (int t = 0; t < NT; t++) bn[t] = ld(b, k * ldb + 16 * t + c, k < NK && 16 * t + c < NN);
  };
  load(0);
#pragma unroll
  for (int s = 0; s < KS; s++) {
    const double av = an;
    double bv[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) bv[t] = bn[t];
    if (s + 1 < KS) load(s + 1);
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = mfma64(av, bv[t], acc[t]);
  }
}

// X[q][n] = acc (rows q < 9, columns n < nn) into LDS
template <int NT>
__device__ __forceinline__ void store_tiles(double* __restrict__ x, int ldx, int nn, const v4d (&acc)[NT], int g,
                                            int c) {
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int q = g + 4 * v, n = 16 * t + c;
      if (q < 9 && n < nn) x[q * ldx + n] = acc[t][v];
    }
}

// Y[q][k] = sum_p S[q][p] X[p][k] for k < NK (one lane per k): the elastic
// star's non-zeros (SP) or all 81 entries; S wave-uniform
template <bool SP, int NK>
__device__ __forceinline__ void mix(const double* __restrict__ sd, const double* __restrict__ x, int ldx,
                                    double* __restrict__ y, int ldy, int lane) {
  for (int k = lane; k < NK; k += 64) {
    double xv[9];
#pragma unroll
    for (int p = 0; p < 9; p++) xv[p] = x[p * ldx + k];
#pragma unroll
    for (int q = 0; q < 9; q++) {
      double z = 0.0;
      if constexpr (SP) {
#pragma unroll
        for (int j = 0; j < kStarNnzRow[q]; j++) z += sd[q * 9 + kStarCol[q][j]] * xv[kStarCol[q][j]];
      } else {
#pragma unroll
        for (int p = 0; p < 9; p++) z += sd[q * 9 + p] * xv[p];
      }
      y[q * ldy + k] = z;
    }
  }
}

template <int O>
struct PlainDims {
  static constexpr int N = n_modes(O), NF = n_face_modes(O), NT = tiles16(N), SZ = 9 * N;
};

// CK stage DE (and the following ones): prev holds der_{DE-1} (leading
// dimension N); the stage writes der_DE into the l_der buffer of its parity
// and adds sca dt / (DE + 1) der_DE to tInt
template <int O, int DE, bool SP>
__device__ __forceinline__ void ck_stage(const SolverDev<double>& s, const double* __restrict__ star, double dt,
                                         double sca, const double* prev, double (*l_der)[9 * n_modes(O)],
                                         double* l_a, double* l_ti, int lane, int g, int c) {
  if constexpr (DE < O) {
    constexpr int N = n_modes(O), NK = n_modes(O - DE + 1), NN = n_modes(O - DE), NT = tiles16(NN);
    v4d acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = zero4d();
#pragma unroll
    for (int d = 0; d < 3; d++) {
      mix<SP, NK>(star + d * 81, prev, N, l_a, N, lane);
      __syncthreads();
      mm<NT, NK, NN>(l_a, N, s.dg.stiffTn + d * N * N, N, acc, g, c);
      __syncthreads();
    }
    double* cur = l_der[DE & 1];
    store_tiles<NT>(cur, N, NN, acc, g, c);
    __syncthreads();
    sca *= dt / double(DE + 1);
    for (int i = lane; i < 9 * NN; i += 64) {
      const int q = i / NN, n = i % NN;
      l_ti[q * N + n] += sca * cur[q * N + n];
    }
    __syncthreads();
    ck_stage<O, DE + 1, SP>(s, star, dt, sca, cur, l_der, l_a, l_ti, lane, g, c);
  }
}

template <int O, bool SP>
__global__ __launch_bounds__(64) void k_local_p64(SolverDev<double> s, int64_t first, double dt,
                                                  double* __restrict__ dofs, double* __restrict__ t_int) {
  using D = PlainDims<O>;
  constexpr int N = D::N, NF = D::NF, NT = D::NT, SZ = D::SZ;
  __shared__ double l_q[SZ], l_ti[SZ], l_der[2][SZ], l_a[SZ];
  const int64_t el = first + blockIdx.x;
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const double* star = s.star + el * 243;
  for (int i = lane; i < SZ; i += 64) {
    const double v = dofs[el * SZ + i];
    l_q[i] = v;
    l_ti[i] = dt * v;
  }
  __syncthreads();
  // ---- CK: der_de = sum_d (S_d der_{de-1}) (-K_d^T M^-1), tInt += dt^{de+1}/(de+1)! der_de
  ck_stage<O, 1, SP>(s, star, dt, dt, l_q, l_der, l_a, l_ti, lane, g, c);
  for (int i = lane; i < SZ; i += 64) t_int[el * SZ + i] = l_ti[i];

  // ---- volume: dofs += sum_d (S_d tInt) (K_d M^-1)
  v4d acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int q = g + 4 * v, n = 16 * t + c;
      acc[t][v] = (q < 9 && n < N) ? l_q[q * N + n] : 0.0;
    }
  constexpr int NV = n_modes(O - 1);
  for (int d = 0; d < 3; d++) {
    mix<SP, NV>(star + d * 81, l_ti, N, l_a, N, lane);
    __syncthreads();
    mm<NT, NV, N>(l_a, N, s.dg.stiffV + d * N * N, N, acc, g, c);
    __syncthreads();
  }
  // ---- local surface: dofs += ((fs0_fa (tInt F^L_fa)) F^T_fa)
  double* l_s0 = l_der[0];  // [9][NF]
  const double* fs0 = s.fs + el * 324;
  for (int fa = 0; fa < 4; fa++) {
    v4d s0[1] = {zero4d()};
    mm<1, N, NF>(l_ti, N, s.dg.fluxLN + fa * N * NF, NF, s0, g, c);
    store_tiles<1>(l_s0, NF, NF, s0, g, c);
    __syncthreads();
    mix<false, NF>(fs0 + fa * 81, l_s0, NF, l_a, NF, lane);
    __syncthreads();
    mm<NT, NF, N>(l_a, NF, s.dg.fluxT + fa * NF * N, N, acc, g, c);
    __syncthreads();
  }
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int q = g + 4 * v, n = 16 * t + c;
      if (q < 9 && n < N) dofs[el * SZ + q * N + n] = acc[t][v];
    }
}

template <int O>
__global__ __launch_bounds__(64) void k_neigh_p64(SolverDev<double> s, int64_t first,
                                                  const double* __restrict__ t_int, double* __restrict__ dofs) {
  using D = PlainDims<O>;
  constexpr int N = D::N, NF = D::NF, NT = D::NT, SZ = D::SZ;
  __shared__ double l_s0[9 * NF], l_s1[9 * NF];
  const int64_t el = first + blockIdx.x;
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  v4d acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int q = g + 4 * v, n = 16 * t + c;
      acc[t][v] = ld(dofs + el * SZ, q * N + n, q < 9 && n < N);
    }
  for (int fa = 0; fa < 4; fa++) {
    const int32_t ne = s.nb.ne[el * 4 + fa];
    if (ne < 0) continue;  // outflow: no neighbouring contribution (uniform per wave)
    const int fm = s.nb.fmat[el * 4 + fa];
    v4d s0[1] = {zero4d()};
    mm<1, N, NF>(t_int + (int64_t)ne * SZ, N, s.dg.fluxLN + fm * N * NF, NF, s0, g, c);
    store_tiles<1>(l_s0, NF, NF, s0, g, c);
    __syncthreads();
    mix<false, NF>(s.fs + ((s.n_el + el) * 4 + fa) * 81, l_s0, NF, l_s1, NF, lane);
    __syncthreads();
    mm<NT, NF, N>(l_s1, NF, s.dg.fluxT + fa * NF * N, N, acc, g, c);
    __syncthreads();
  }
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int q = g + 4 * v, n = 16 * t + c;
      if (q < 9 && n < N) dofs[el * SZ + q * N + n] = acc[t][v];
    }
}

template <int O>
int local_o(const SolverDev<double>& s, int64_t first, int64_t n, double dt, double* dofs, double* t_int,
            hipStream_t st) {
  if (s.sparse_star)
    hipLaunchKernelGGL((k_local_p64<O, true>), dim3((unsigned)n), dim3(64), 0, st, s, first, dt, dofs, t_int);
  else
    hipLaunchKernelGGL((k_local_p64<O, false>), dim3((unsigned)n), dim3(64), 0, st, s, first, dt, dofs, t_int);
  EDGE_HIP_TRY(hipGetLastError());
  return EDGE_OK;
}

template <int O>
int neigh_o(const SolverDev<double>& s, int64_t first, int64_t n, const double* t_int, double* dofs,
            hipStream_t st) {
  hipLaunchKernelGGL((k_neigh_p64<O>), dim3((unsigned)n), dim3(64), 0, st, s, first, t_int, dofs);
  EDGE_HIP_TRY(hipGetLastError());
  return EDGE_OK;
}

}  // namespace

bool plain64_supported(int order, int n_runs, int prec_bytes) {
  return order >= 2 && order <= 5 && n_runs == 1 && prec_bytes == 8;
}

int launch_local_plain64(const SolverDev<double>& s, int64_t first, int64_t n, double dt, double* dofs,
                         double* t_int, hipStream_t st) {
  if (n <= 0) return EDGE_OK;
  switch (s.order) {
    case 2: return local_o<2>(s, first, n, dt, dofs, t_int, st);
    case 3: return local_o<3>(s, first, n, dt, dofs, t_int, st);
    case 4: return local_o<4>(s, first, n, dt, dofs, t_int, st);
    case 5: return local_o<5>(s, first, n, dt, dofs, t_int, st);
  }
  return fail(EDGE_ERR_UNSUPPORTED, "order");
}

int launch_neigh_plain64(const SolverDev<double>& s, int64_t first, int64_t n, const double* t_int, double* dofs,
                         hipStream_t st) {
  if (n <= 0) return EDGE_OK;
  switch (s.order) {
    case 2: return neigh_o<2>(s, first, n, t_int, dofs, st);
    case 3: return neigh_o<3>(s, first, n, t_int, dofs, st);
    case 4: return neigh_o<4>(s, first, n, t_int, dofs, st);
    case 5: return neigh_o<5>(s, first, n, t_int, dofs, st);
  }
  return fail(EDGE_ERR_UNSUPPORTED, "order");
}

}  // namespace edge
