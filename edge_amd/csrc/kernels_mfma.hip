// MFMA-tiled element kernels for 16 fused runs in FP32 (gfx950).
//
// The fused-run axis (16 runs, innermost in the DOF layout [q][mode][run])
// is the N dimension of v_mfma_f32_16x16x4f32.  Every product with a shared
// reference-element matrix M ([k][n], stiffness or flux) becomes, per
// quantity q, a 16 x 16 output tile C[n][r] += sum_k M[k][n] X_q[k][r]:
//   A operand (16 x 4)  lane (g, c) = M[k(g)][16 t + c]      (shared matrix)
//   B operand (4 x 16)  lane (g, c) = X_q[k(g)][c]           (element data, run c)
//   D (16 x 16)         lane (g, c), register v = C[16 t + 4 g + v][c]
// with g = lane / 16, c = lane % 16.  B operands come either from memory in
// row order (k = 4 s + g) or straight from a previous D tile with the
// k-order permuted (k = 16 u + 4 g + v), so GEMM chains stay in registers.
// One wave handles one element at a time: the 9 quantities of an (n, r)
// position then live in the same lane, and the per-element 9 x 9 star and
// flux-solver products are lane-local VALU FMAs with wave-uniform
// coefficients (scalar loads).
//
// Operations (src/impl/seismic/solvers/AderDg.hpp:362-533, :553-722):
//   local: CK time prediction (TimePredVanilla.hpp:200-306), volume
//          integral (VolIntVanilla.hpp:189-241), local surface integral
//          (SurfIntVanilla.hpp:212-262);
//   neigh: neighbouring surface integral (SurfIntVanilla.hpp:274-357).
// Summation order differs from the reference's loops (tile order, the star
// applied after the stiffness product); results agree to rounding.
#include <hip/hip_runtime.h>

#include "kernels_mfma.hpp"

namespace edge {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
constexpr int R = 16;

__device__ __forceinline__ v4f mfma(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

constexpr int tiles(int rows) { return (rows + 15) / 16; }
constexpr int ksteps(int rows) { return (rows + 3) / 4; }

template <int O>
struct Dims {
  static constexpr int N = n_modes(O);
  static constexpr int NF = n_face_modes(O);
  static constexpr int NT = tiles(N);
  static constexpr int SZ = 9 * N * R;
};

// blockIdx -> element offset: each of the 8 XCDs (blocks dispatched round
// robin) gets a contiguous element range, so neighbours share its L2.
__device__ __forceinline__ int64_t xcd_element(int64_t n) {
  const int64_t chunk = (n + 7) / 8;
  const int64_t b = blockIdx.x;
  return (b & 7) * chunk + (b >> 3);
}

__device__ __forceinline__ v4f zero4() { return v4f{0.f, 0.f, 0.f, 0.f}; }

// Guarded load without a branch: the index is clamped into the valid range
// (so the load is always in bounds) and the value masked afterwards.
__device__ __forceinline__ float ldg(const float* __restrict__ p, int idx, int clamp_idx, bool ok) {
  const float v = p[ok ? idx : clamp_idx];
  return ok ? v : 0.f;
}

// D-layout load of rows [16 t + 4 g + v] (< rows) of X[k][r] (row stride R)
__device__ __forceinline__ v4f load_d(const float* __restrict__ x, int rows, int t, int g, int c) {
  v4f o;
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int k = 16 * t + 4 * g + v;
    o[v] = ldg(x, k * R + c, c, k < rows);
  }
  return o;
}

__device__ __forceinline__ void store_d(float* __restrict__ x, int rows, int t, int g, int c, v4f o) {
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int k = 16 * t + 4 * g + v;
    if (k < rows) x[k * R + c] = o[v];
  }
}

// ---------------------------------------------------------------- local
template <int O, int DE>
__device__ __forceinline__ void ck_stage(const float* __restrict__ stiffTn, const float* __restrict__ star,
                                         const float* __restrict__ q0, float* __restrict__ l_der,
                                         float* __restrict__ l_tint, float dt, float& sca, int g, int c) {
  constexpr int N = Dims<O>::N;
  constexpr int NK = n_modes(O - DE + 1);  // rows of the previous derivative
  constexpr int NN = n_modes(O - DE);      // rows of this derivative
  constexpr int TN = tiles(NN);
  constexpr int KS = ksteps(NK);
  // der_q = sum_d sum_k (-K_d^T M^-1)[k][n] (sum_p S_d[q][p] der_prev_p[k][r]): the
  // star matrix is applied to the B operands (row order k = 4 s + g)
  v4f der[9][TN];
#pragma unroll
  for (int q = 0; q < 9; q++)
#pragma unroll
    for (int t = 0; t < TN; t++) der[q][t] = zero4();
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const float* m = stiffTn + d * N * N;
    const float* sd = star + d * 81;
#pragma unroll 1
    for (int s = 0; s < KS; s++) {
      const int k = 4 * s + g;
      float a[TN], b[9];
#pragma unroll
      for (int t = 0; t < TN; t++) a[t] = ldg(m, k * N + 16 * t + c, 0, k < NK && 16 * t + c < NN);
#pragma unroll
      for (int p = 0; p < 9; p++) {
        if constexpr (DE == 1)
          b[p] = ldg(q0, (p * N + k) * R + c, c, k < NK);
        else
          b[p] = ldg(l_der, (p * NK + k) * R + c, c, k < NK);
      }
#pragma unroll
      for (int q = 0; q < 9; q++) {
        float z = 0.f;
#pragma unroll
        for (int p = 0; p < 9; p++) z += sd[q * 9 + p] * b[p];
#pragma unroll
        for (int t = 0; t < TN; t++) der[q][t] = mfma(a[t], z, der[q][t]);
      }
    }
  }
  sca *= dt / float(DE + 1);
  // tInt += dt^{DE+1}/(DE+1)! der_DE (each lane updates the entries it holds)
#pragma unroll
  for (int q = 0; q < 9; q++)
#pragma unroll
    for (int t = 0; t < TN; t++) {
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int k = 16 * t + 4 * g + v;
        if (k < NN) l_tint[(q * N + k) * R + c] += sca * der[q][t][v];
      }
    }
  if constexpr (DE + 1 < O) {
    __syncthreads();  // previous readers of l_der are done
#pragma unroll
    for (int q = 0; q < 9; q++)
#pragma unroll
      for (int t = 0; t < TN; t++) store_d(l_der + q * NN * R, NN, t, g, c, der[q][t]);
    __syncthreads();
  }
}

template <int O, int DE>
__device__ __forceinline__ void ck_all(const float* stiffTn, const float* star, const float* q0, float* l_der,
                                       float* l_tint, float dt, float& sca, int g, int c) {
  if constexpr (DE < O) {
    ck_stage<O, DE>(stiffTn, star, q0, l_der, l_tint, dt, sca, g, c);
    ck_all<O, DE + 1>(stiffTn, star, q0, l_der, l_tint, dt, sca, g, c);
  }
}

template <int O>
__global__ __launch_bounds__(64) void k_local(SolverDev<float> s, int64_t first, int64_t n, float dt,
                                              float* __restrict__ dofs, float* __restrict__ t_int) {
  using D = Dims<O>;
  constexpr int N = D::N, NF = D::NF, NT = D::NT;
  constexpr int NV = n_modes(O - 1);  // rows of tInt the volume integral reads
  // time-integrated DOFs and the derivative handed from one CK stage to the next
  __shared__ float l_buf[9 * N * R];
  __shared__ float l_der[9 * (O > 1 ? n_modes(O - 1) : 1) * R];
  const int64_t off = xcd_element(n);
  if (off >= n) return;
  const int64_t el = first + off;
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const float* q0 = dofs + el * D::SZ;
  const float* star = s.star + el * 243;

  // ---- CK: tInt = dt Q + sum_de dt^{de+1}/(de+1)! der_de, accumulated in LDS
#pragma unroll
  for (int q = 0; q < 9; q++)
#pragma unroll
    for (int t = 0; t < NT; t++) store_d(l_buf + q * N * R, N, t, g, c, dt * load_d(q0 + q * N * R, N, t, g, c));
  float sca = dt;
  ck_all<O, 1>(s.dg.stiffTn, star, q0, l_der, l_buf, dt, sca, g, c);
  __syncthreads();
  float* gti = t_int + el * D::SZ;
  for (int i = lane; i < 9 * N * R; i += 64) gti[i] = l_buf[i];

  // ---- volume: acc_q += sum_d stiffV_d^T (sum_p S_d[q][p] tInt_p)
  v4f acc[9][NT];
#pragma unroll
  for (int q = 0; q < 9; q++)
#pragma unroll
    for (int t = 0; t < NT; t++) acc[q][t] = load_d(q0 + q * N * R, N, t, g, c);
  constexpr int KV = ksteps(NV);
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const float* m = s.dg.stiffV + d * N * N;
    const float* sd = star + d * 81;
#pragma unroll 1
    for (int ks = 0; ks < KV; ks++) {
      const int k = 4 * ks + g;
      float a[NT], b[9];
#pragma unroll
      for (int t = 0; t < NT; t++) a[t] = ldg(m, k * N + 16 * t + c, 0, k < NV && 16 * t + c < N);
#pragma unroll
      for (int p = 0; p < 9; p++) b[p] = ldg(l_buf, (p * N + k) * R + c, c, k < NV);
#pragma unroll
      for (int q = 0; q < 9; q++) {
        float z = 0.f;
#pragma unroll
        for (int p = 0; p < 9; p++) z += sd[q * 9 + p] * b[p];
#pragma unroll
        for (int t = 0; t < NT; t++) acc[q][t] = mfma(a[t], z, acc[q][t]);
      }
    }
  }

  // ---- local surface integral
  const float* fs0 = s.fs + el * 324;
  constexpr int KS = ksteps(N);
#pragma unroll 1
  for (int fa = 0; fa < 4; fa++) {
    const float* fl = s.dg.fluxLN + fa * N * NF;
    v4f s0[9];
#pragma unroll
    for (int q = 0; q < 9; q++) s0[q] = zero4();
#pragma unroll 1
    for (int ks = 0; ks < KS; ks++) {
      const int k = 4 * ks + g;
      const float a = ldg(fl, k * NF + c, 0, k < N && c < NF);
#pragma unroll
      for (int q = 0; q < 9; q++) {
        const float b = ldg(l_buf, (q * N + k) * R + c, c, k < N);
        s0[q] = mfma(a, b, s0[q]);
      }
    }
    const float* f = fs0 + fa * 81;
    v4f s1[9];
#pragma unroll
    for (int q = 0; q < 9; q++) {
      s1[q] = zero4();
#pragma unroll
      for (int p = 0; p < 9; p++) s1[q] += f[q * 9 + p] * s0[p];
    }
    const float* ft = s.dg.fluxT + fa * NF * N;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int fr = 4 * g + v;  // permuted k: face mode held in register v
#pragma unroll
      for (int t = 0; t < NT; t++) {
        const float a = ldg(ft, fr * N + 16 * t + c, 0, fr < NF && 16 * t + c < N);
#pragma unroll
        for (int q = 0; q < 9; q++) acc[q][t] = mfma(a, s1[q][v], acc[q][t]);
      }
    }
  }
  float* out = dofs + el * D::SZ;
#pragma unroll
  for (int q = 0; q < 9; q++)
#pragma unroll
    for (int t = 0; t < NT; t++) store_d(out + q * N * R, N, t, g, c, acc[q][t]);
}

// ---------------------------------------------------------------- neigh
template <int O>
__global__ __launch_bounds__(64) void k_neigh(SolverDev<float> s, int64_t first, int64_t n,
                                              const float* __restrict__ t_int, float* __restrict__ dofs) {
  using D = Dims<O>;
  constexpr int N = D::N, NF = D::NF, NT = D::NT;
  const int64_t off = xcd_element(n);
  if (off >= n) return;
  const int64_t el = first + off;
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  float* q0 = dofs + el * D::SZ;
  v4f acc[9][NT];
#pragma unroll
  for (int q = 0; q < 9; q++)
#pragma unroll
    for (int t = 0; t < NT; t++) acc[q][t] = load_d(q0 + q * N * R, N, t, g, c);
  constexpr int KS = ksteps(N);
#pragma unroll 1
  for (int fa = 0; fa < 4; fa++) {
    const int32_t ne = s.nb.ne[el * 4 + fa];
    if (ne < 0) continue;  // outflow (uniform across the wave)
    const int fm = s.nb.fmat[el * 4 + fa];
    const float* fl = s.dg.fluxLN + fm * N * NF;
    const float* tn = t_int + (int64_t)ne * D::SZ;
    v4f s0[9];
#pragma unroll
    for (int q = 0; q < 9; q++) s0[q] = zero4();
#pragma unroll 1
    for (int ks = 0; ks < KS; ks++) {
      const int k = 4 * ks + g;
      const float a = ldg(fl, k * NF + c, 0, k < N && c < NF);
#pragma unroll
      for (int q = 0; q < 9; q++) {
        const float b = ldg(tn, (q * N + k) * R + c, c, k < N);
        s0[q] = mfma(a, b, s0[q]);
      }
    }
    const float* f = s.fs + ((s.n_el + el) * 4 + fa) * 81;
    v4f s1[9];
#pragma unroll
    for (int q = 0; q < 9; q++) {
      s1[q] = zero4();
#pragma unroll
      for (int p = 0; p < 9; p++) s1[q] += f[q * 9 + p] * s0[p];
    }
    const float* ft = s.dg.fluxT + fa * NF * N;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int fr = 4 * g + v;
#pragma unroll
      for (int t = 0; t < NT; t++) {
        const float a = ldg(ft, fr * N + 16 * t + c, 0, fr < NF && 16 * t + c < N);
#pragma unroll
        for (int q = 0; q < 9; q++) acc[q][t] = mfma(a, s1[q][v], acc[q][t]);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 9; q++)
#pragma unroll
    for (int t = 0; t < NT; t++) store_d(q0 + q * N * R, N, t, g, c, acc[q][t]);
}

unsigned grid_of(int64_t n) { return (unsigned)(8 * ((n + 7) / 8)); }

}  // namespace

bool mfma_supported(int order, int n_runs, int precision) {
  return precision == 32 && n_runs == 16 && order >= 2 && order <= 5;
}

template <>
int launch_local_mfma<float>(const SolverDev<float>& s, int64_t first, int64_t n, float dt, float* dofs,
                             float* t_int, hipStream_t st) {
  if (n <= 0) return EDGE_OK;
  const dim3 grid(grid_of(n)), block(64);
  switch (s.order) {
    case 2: hipLaunchKernelGGL(k_local<2>, grid, block, 0, st, s, first, n, dt, dofs, t_int); break;
    case 3: hipLaunchKernelGGL(k_local<3>, grid, block, 0, st, s, first, n, dt, dofs, t_int); break;
    case 4: hipLaunchKernelGGL(k_local<4>, grid, block, 0, st, s, first, n, dt, dofs, t_int); break;
    case 5: hipLaunchKernelGGL(k_local<5>, grid, block, 0, st, s, first, n, dt, dofs, t_int); break;
    default: return fail(EDGE_ERR_UNSUPPORTED, "order");
  }
  EDGE_HIP_TRY(hipGetLastError());
  return EDGE_OK;
}

template <>
int launch_neigh_mfma<float>(const SolverDev<float>& s, int64_t first, int64_t n, const float* t_int, float* dofs,
                             hipStream_t st) {
  if (n <= 0) return EDGE_OK;
  const dim3 grid(grid_of(n)), block(64);
  switch (s.order) {
    case 2: hipLaunchKernelGGL(k_neigh<2>, grid, block, 0, st, s, first, n, t_int, dofs); break;
    case 3: hipLaunchKernelGGL(k_neigh<3>, grid, block, 0, st, s, first, n, t_int, dofs); break;
    case 4: hipLaunchKernelGGL(k_neigh<4>, grid, block, 0, st, s, first, n, t_int, dofs); break;
    case 5: hipLaunchKernelGGL(k_neigh<5>, grid, block, 0, st, s, first, n, t_int, dofs); break;
    default: return fail(EDGE_ERR_UNSUPPORTED, "order");
  }
  EDGE_HIP_TRY(hipGetLastError());
  return EDGE_OK;
}

template <>
int launch_local_mfma<double>(const SolverDev<double>&, int64_t, int64_t, double, double*, double*, hipStream_t) {
  return fail(EDGE_ERR_UNSUPPORTED, "MFMA-tiled kernels are FP32 with 16 runs");
}
template <>
int launch_neigh_mfma<double>(const SolverDev<double>&, int64_t, int64_t, const double*, double*, hipStream_t) {
  return fail(EDGE_ERR_UNSUPPORTED, "MFMA-tiled kernels are FP32 with 16 runs");
}

}  // namespace edge
