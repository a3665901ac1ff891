// Single-run element kernels with 16 elements per wave, FP64 and FP32 MFMA.
//
// With one run per element there is no run axis to fill an MFMA tile.  These
// kernels put 16 elements on the N dimension of v_mfma_f64_16x16x4f64 /
// v_mfma_f32_16x16x4f32 instead: the shared reference-element matrices are
// the A operands, and every product with one is a full 16 x 16 tile per
// quantity:
//   A operand (16 x 4)  lane (g, c) = M[k(s, g)][16 t + c]     (shared, packed)
//   B operand (4 x 16)  lane (g, c) = X_e[q][k(s, g)]          (element e = c)
//   D (16 x 16)         lane (g, c), register v = Out_e[q][row(t, g, v)]
// with g = lane / 16, c = lane % 16, row = 16 t + g + 4 v (FP64) or
// 16 t + 4 g + v (FP32), and k(s, g) = 4 s + g for B operands from LDS.
// Register v of a D tile t is the B operand of k-step 4 t + v (the packed
// A operands follow the D rows, ec64_pack.hpp), so product chains (the CK's
// derivatives, the lifts of the face projections) stay in registers.  The 9 quantities of an
// (element, mode) position live in one lane: the per-element star matrices
// and flux-solver factors act lane-locally, with per-lane coefficients.
// The elements' DOFs pass through an LDS tile [q][mode][16] (40 KiB in FP64:
// one wave per SIMD; 20 KiB in FP32: two).
//
// The neighbouring surface integral uses face projections, as the FP32
// MFMA kernels do: the local kernel projects each face's tInt with the
// local flux matrix fluxL[f] and the receiving side's V^T (rank-3 flux
// solvers), and the neighbour kernel turns that trace into its own face
// coordinates with the map R_v of the face orientation (fluxN[4 v + f] =
// fluxL[f] R_v; the reference indexes fluxN by vId and fId,
// src/impl/seismic/kernels/SurfInt.hpp:127-130) before the lift.
//
// Operations (src/impl/seismic/solvers/AderDg.hpp:362-533, :553-722):
//   local: CK time prediction (TimePredVanilla.hpp:200-306), volume
//          integral (VolIntVanilla.hpp:189-241), local surface integral
//          (SurfIntVanilla.hpp:212-262);
//   neigh: neighbouring surface integral (SurfIntVanilla.hpp:274-357).
#include <hip/hip_runtime.h>

#include "dg_masks.hpp"
#include "ec64_pack.hpp"
#include "kernels.hpp"

namespace edge {
namespace {

// The MFMA of each precision and its D layout.
template <typename T>
struct Mx;
template <>
struct Mx<double> {
  typedef double V __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ V mfma(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static constexpr __host__ __device__ int row(int t, int g, int v) { return 16 * t + g + 4 * v; }
  static constexpr bool f32 = false;
};
template <>
struct Mx<float> {
  typedef float V __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ V mfma(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static constexpr __host__ __device__ int row(int t, int g, int v) { return 16 * t + 4 * g + v; }
  static constexpr bool f32 = true;
};
template <typename T>
using V4 = typename Mx<T>::V;

template <typename T>
__device__ __forceinline__ V4<T> zero4() {
  return V4<T>{T(0), T(0), T(0), T(0)};
}

using ec::ksteps;
using ec::pairs;
using ec::tiles;

// blockIdx -> wave (16 elements) index: each of the 8 XCDs (workgroups
// dispatched round robin) gets a contiguous range, so neighbours share its
// L2; or workgroup i -> wave i.  The grid holds 8 ceil(n / 8) workgroups
// with the XCD-aware schedule.
__device__ __forceinline__ int64_t wave_index(int64_t n, bool xcd) {
  if (!xcd) return blockIdx.x;
  const int64_t chunk = (n + 7) / 8;
  const int64_t b = blockIdx.x;
  return (b & 7) * chunk + (b >> 3);
}

// Elastic star pattern (the rows' non-zero columns, src/impl/seismic/setups/
// Elasticity.cpp:172-250): velocity row 6 + r reads stress columns
// kVel[r][0..2]; stress row q reads velocity columns 6 + kStr[q][0..nnz-1].
constexpr int kVel[3][3] = {{0, 3, 5}, {1, 3, 4}, {2, 4, 5}};
constexpr int kStrNnz[6] = {3, 3, 3, 2, 2, 2};
constexpr int kStr[6][3] = {{0, 1, 2}, {0, 1, 2}, {0, 1, 2}, {0, 1, 0}, {1, 2, 0}, {0, 2, 0}};

// The velocity rows of one star matrix (this lane's element): 9 coefficients.
template <typename T>
struct VelRows {
  T v[9];
  __device__ __forceinline__ void load(const T* __restrict__ sd) {
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int j = 0; j < 3; j++) v[r * 3 + j] = sd[(6 + r) * 9 + kVel[r][j]];
  }
  __device__ __forceinline__ T row(int r, const T (&b)[9]) const {
    return v[r * 3] * b[kVel[r][0]] + v[r * 3 + 1] * b[kVel[r][1]] + v[r * 3 + 2] * b[kVel[r][2]];
  }
};

// out_q (+)= sum_i S_d[q][6 + i] P_i for the stress rows q = 0..5 (D tiles)
template <typename T, int TN>
__device__ __forceinline__ void add_stress(const T* __restrict__ sd, const V4<T> (&p)[3][TN], V4<T> (&out)[9][TN],
                                           bool first) {
#pragma unroll
  for (int q = 0; q < 6; q++) {
    T cf[3];
#pragma unroll
    for (int j = 0; j < kStrNnz[q]; j++) cf[j] = sd[q * 9 + 6 + kStr[q][j]];
#pragma unroll
    for (int t = 0; t < TN; t++) {
      V4<T> z = first ? zero4<T>() : out[q][t];
#pragma unroll
      for (int j = 0; j < kStrNnz[q]; j++) z += cf[j] * p[kStr[q][j]][t];
      out[q][t] = z;
    }
  }
}

// A operand of k-step s, tile t of a packed product (ec::Layout) with np
// k-step pairs per tile
template <typename T>
__device__ __forceinline__ T a_op(const T* __restrict__ blk, int np, int t, int s, int lane) {
  return blk[((t * np + (s >> 1)) * 64 + lane) * 2 + (s & 1)];
}

// B values of k-step s from the LDS tile: the 9 quantities' rows 4 s + g of
// column c.  Rows past a quantity's N read the next quantity's first rows,
// or the zero rows past the tile: always finite, and the A operand is zero
// there.
template <typename T, int N>
__device__ __forceinline__ void b_lds(const T* __restrict__ l_x, int s, int g, int c, T (&b)[9]) {
#pragma unroll
  for (int p = 0; p < 9; p++) b[p] = l_x[(p * N + 4 * s + g) * 16 + c];
}

// CK stage DE >= 2 from the previous derivative in registers (k-step s is
// register s % 4 of tile s / 4), elastic pattern: per direction the three
// raw velocity products and the three star-mixed velocity rows.
template <typename T, int O, int DE>
__device__ __forceinline__ void ck_next(const T* __restrict__ pk, const T* __restrict__ star,
                                        const V4<T> (&prev)[9][tiles(n_modes(O - DE + 1))], V4<T> (&cur)[9][1],
                                        int lane) {
  constexpr ec::Layout L{O, Mx<T>::f32};
  constexpr int NK = n_modes(O - DE + 1), KS = L.chain(NK), NP = pairs(KS);
  constexpr masks::Tab NZ = masks::Nz<O, Mx<T>::f32>::tab;
  static_assert(tiles(n_modes(O - DE)) == 1, "derivatives of stage 2 and up fit one tile");
#pragma unroll
  for (int r = 6; r < 9; r++) cur[r][0] = zero4<T>();
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const T* blk = pk + L.ck(DE, d);
    VelRows<T> sv;
    sv.load(star + d * 81);
    V4<T> pv[3][1];
#pragma unroll
    for (int i = 0; i < 3; i++) pv[i][0] = zero4<T>();
    const uint32_t nz = NZ.ck[DE][d][0];  // k-steps with a non-zero A block
    T an = (nz & 1) ? a_op(blk, NP, 0, 0, lane) : T(0);
#pragma unroll
    for (int s = 0; s < KS; s++) {
      const T a = an;
      if (s + 1 < KS && ((nz >> (s + 1)) & 1)) an = a_op(blk, NP, 0, s + 1, lane);  // one k-step ahead
      __builtin_amdgcn_sched_barrier(0);
      if (!((nz >> s) & 1)) continue;
      T b[9];
#pragma unroll
      for (int p = 0; p < 9; p++) b[p] = prev[p][s / 4][s % 4];
#pragma unroll
      for (int i = 0; i < 3; i++) pv[i][0] = Mx<T>::mfma(a, b[6 + i], pv[i][0]);
#pragma unroll
      for (int r = 0; r < 3; r++) cur[6 + r][0] = Mx<T>::mfma(a, sv.row(r, b), cur[6 + r][0]);
    }
    add_stress<T, 1>(star + d * 81, pv, cur, d == 0);
  }
}

// tInt (LDS) += x for the D rows of tile t below `rows`, column c
template <typename T, int N>
__device__ __forceinline__ void add_tint(T* __restrict__ l_x, int rows, int t, const V4<T> (&x)[9], int g, int c) {
#pragma unroll
  for (int q = 0; q < 9; q++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int m = Mx<T>::row(t, g, v);
      if (m < rows) l_x[(q * N + m) * 16 + c] += x[q][v];
    }
}

// CK stages DE >= 2; each adds sca dt^{DE} / (DE + 1)! der_DE to tInt (LDS).
template <typename T, int O, int DE>
__device__ __forceinline__ void ck_rest(const T* __restrict__ pk, const T* __restrict__ star,
                                        const V4<T> (&prev)[9][tiles(n_modes(O - DE + 1))], T* __restrict__ l_x, T dt,
                                        T sca, int g, int c, int lane) {
  if constexpr (DE < O) {
    V4<T> cur[9][1];
    ck_next<T, O, DE>(pk, star, prev, cur, lane);
    sca *= dt / T(DE + 1);
    V4<T> xt[9];
#pragma unroll
    for (int q = 0; q < 9; q++) xt[q] = sca * cur[q][0];
    add_tint<T, n_modes(O)>(l_x, n_modes(O - DE), 0, xt, g, c);
    ck_rest<T, O, DE + 1>(pk, star, cur, l_x, dt, sca, g, c, lane);
  }
}

// The per-element tiles [9][N] of elements e0 .. e0 + ne - 1 (contiguous in
// the reference layout) into the LDS tile [q][mode][16], zeros for missing
// columns.
template <typename T, int SZ>
__device__ __forceinline__ void gather(const T* __restrict__ src, int ne, T* __restrict__ l_x, int lane) {
  constexpr int IT = (SZ * 16 + 63) / 64, B = 16;
#pragma unroll 1
  for (int i0 = 0; i0 < IT; i0 += B) {
    T v[B];
#pragma unroll
    for (int j = 0; j < B; j++) {
      const int idx = (i0 + j) * 64 + lane, i = idx >> 4, e = idx & 15;
      v[j] = (i0 + j < IT && idx < SZ * 16 && e < ne) ? src[(size_t)e * SZ + i] : T(0);
    }
#pragma unroll
    for (int j = 0; j < B; j++) {
      const int idx = (i0 + j) * 64 + lane;
      if (i0 + j < IT && idx < SZ * 16) l_x[idx] = v[j];
    }
  }
}

// The LDS tile [q][mode][16] back to the elements' tiles (transposed).
template <typename T, int SZ>
__device__ __forceinline__ void scatter(const T* __restrict__ l_x, int ne, T* __restrict__ dst, int lane) {
  constexpr int IT = (SZ * 16 + 63) / 64;
#pragma unroll 8
  for (int j = 0; j < IT; j++) {
    const int idx = j * 64 + lane, i = idx >> 4, e = idx & 15;
    if (idx < SZ * 16 && e < ne) dst[(size_t)e * SZ + i] = l_x[idx];
  }
}

// dofs of this lane's element (the D rows) += acc
template <typename T, int N, int NT>
__device__ __forceinline__ void add_out(T* __restrict__ q0, const V4<T> (&acc)[9][NT], int g) {
  T o[9][NT][4];
#pragma unroll
  for (int q = 0; q < 9; q++)
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int m = Mx<T>::row(t, g, v);
        o[q][t][v] = m < N ? q0[q * N + m] : T(0);
      }
#pragma unroll
  for (int q = 0; q < 9; q++)
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int m = Mx<T>::row(t, g, v);
        if (m < N) q0[q * N + m] = o[q][t][v] + acc[q][t][v];
      }
}

// Waves per SIMD: the FP64 LDS tile (40 KiB) allows one, the FP32 one two.
template <typename T>
constexpr int kWaves = sizeof(T) == 8 ? 1 : 2;

// ---------------------------------------------------------------- local
template <typename T, int O>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kWaves<T>, kWaves<T>)))
void k_local_ec(SolverDev<T> s, EcDev<T> x, int64_t first, int64_t n, T dt, T* __restrict__ dofs,
                T* __restrict__ t_int) {
  constexpr int N = n_modes(O), NF = n_face_modes(O), NT = tiles(N), SZ = 9 * N;
  constexpr int NN1 = n_modes(O - 1), T1 = tiles(NN1), KS = ksteps(N), NV = n_modes(O - 1), KV = ksteps(NV);
  constexpr ec::Layout L{O, Mx<T>::f32};
  // Q, then tInt; zero rows past the tile for the k-step tail of the last
  // quantity (rows up to 8 N + 4 KS - 1)
  constexpr int PADR = 4 * KS - N > 0 ? 4 * KS - N : 1;
  __shared__ __align__(16) T l_x[(SZ + PADR) * 16];
  const int64_t nw = (n + 15) / 16;
  const int64_t w = wave_index(nw, s.xcd_map);
  if (w >= nw) return;  // one wave per workgroup: no barrier is left waiting
  const int64_t e0 = first + 16 * w;
  const int ne = (int)min((int64_t)16, first + n - e0);
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const bool col = c < ne;
  const int64_t el = e0 + (col ? c : 0);  // this lane's element (a valid one on padding columns)
  const T* star = s.star + el * 243;
  const T* pk = x.pk;

  VelRows<T> sv1[3];  // CK stage 1's star rows, requested before the tile
#pragma unroll
  for (int d = 0; d < 3; d++) sv1[d].load(star + d * 81);
  gather<T, SZ>(dofs + e0 * SZ, ne, l_x, lane);
  for (int i = lane; i < PADR * 16; i += 64) l_x[SZ * 16 + i] = T(0);
  __syncthreads();

  // ---- CK stage 1: per direction the raw velocity products and the
  // star-mixed velocity rows; the LDS rows become dt Q (tInt's first term)
  V4<T> der[9][T1];
  {
    V4<T> pv[3][3][T1];
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int t = 0; t < T1; t++) pv[d][i][t] = zero4<T>();
#pragma unroll
    for (int r = 6; r < 9; r++)
#pragma unroll
      for (int t = 0; t < T1; t++) der[r][t] = zero4<T>();
    constexpr int NP = pairs(KS);
    constexpr masks::Tab NZ = masks::Nz<O, Mx<T>::f32>::tab;  // blocks with non-zero A
    // operands of k-step ks + 1 are requested while k-step ks runs
    auto load_a = [&](int ks, T (&a)[3][T1]) {
#pragma unroll
      for (int d = 0; d < 3; d++)
#pragma unroll
        for (int t = 0; t < T1; t++)
          a[d][t] = ((NZ.ck[1][d][t] >> ks) & 1) ? a_op(pk + L.ck(1, d), NP, t, ks, lane) : T(0);
    };
    T an[3][T1], bn[9];
    load_a(0, an);
    b_lds<T, N>(l_x, 0, g, c, bn);
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
      T a[3][T1], b[9];
#pragma unroll
      for (int d = 0; d < 3; d++)
#pragma unroll
        for (int t = 0; t < T1; t++) a[d][t] = an[d][t];
#pragma unroll
      for (int p = 0; p < 9; p++) b[p] = bn[p];
      if (ks + 1 < KS) {  // rows 4 (ks + 1) + g: not the rows this step rewrites
        load_a(ks + 1, an);
        b_lds<T, N>(l_x, ks + 1, g, c, bn);
      }
      __builtin_amdgcn_sched_barrier(0);  // bounds register pressure: no hoisting across k-steps
      const int k = 4 * ks + g;
      if (k < N) {
#pragma unroll
        for (int p = 0; p < 9; p++) l_x[(p * N + k) * 16 + c] = dt * b[p];
      }
#pragma unroll
      for (int d = 0; d < 3; d++) {
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
          for (int t = 0; t < T1; t++)
            if ((NZ.ck[1][d][t] >> ks) & 1) pv[d][i][t] = Mx<T>::mfma(a[d][t], b[6 + i], pv[d][i][t]);
#pragma unroll
        for (int r = 0; r < 3; r++) {
          const T z = sv1[d].row(r, b);
#pragma unroll
          for (int t = 0; t < T1; t++)
            if ((NZ.ck[1][d][t] >> ks) & 1) der[6 + r][t] = Mx<T>::mfma(a[d][t], z, der[6 + r][t]);
        }
      }
    }
#pragma unroll
    for (int d = 0; d < 3; d++) add_stress<T, T1>(star + d * 81, pv[d], der, d == 0);
  }
  // tInt = dt Q + sum_de dt^{de+1} / (de+1)! der_de, summed in LDS
  {
    const T sca = dt * dt / T(2);
#pragma unroll
    for (int t = 0; t < T1; t++) {
      V4<T> xt[9];
#pragma unroll
      for (int q = 0; q < 9; q++) xt[q] = sca * der[q][t];
      add_tint<T, N>(l_x, NN1, t, xt, g, c);
    }
    ck_rest<T, O, 2>(pk, star, der, l_x, dt, sca, g, c, lane);
  }
  __syncthreads();
  scatter<T, SZ>(l_x, ne, t_int + e0 * SZ, lane);

  // ---- volume: acc_q += sum_d stiffV_d^T (S_d tInt)_q over tInt's first NV modes
  V4<T> acc[9][NT];
#pragma unroll
  for (int q = 0; q < 9; q++)
#pragma unroll
    for (int t = 0; t < NT; t++) acc[q][t] = zero4<T>();
  constexpr int NPV = pairs(KV);
  constexpr masks::Tab NZV = masks::Nz<O, Mx<T>::f32>::tab;  // blocks with non-zero A
  T avn[NT], bvn[9];
#pragma unroll
  for (int t = 0; t < NT; t++) avn[t] = (NZV.vol[0][t] & 1) ? a_op(pk + L.vol(0), NPV, t, 0, lane) : T(0);
  b_lds<T, N>(l_x, 0, g, c, bvn);
#pragma unroll
  for (int d = 0; d < 3; d++) {
    VelRows<T> sv;
    sv.load(star + d * 81);
    V4<T> pv[3][NT];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int t = 0; t < NT; t++) pv[i][t] = zero4<T>();
#pragma unroll
    for (int ks = 0; ks < KV; ks++) {
      T b[9], a[NT];
#pragma unroll
      for (int t = 0; t < NT; t++) a[t] = avn[t];
#pragma unroll
      for (int p = 0; p < 9; p++) b[p] = bvn[p];
      // the next k-step: (d, ks + 1) or (d + 1, 0)
      if (ks + 1 < KV || d + 1 < 3) {
        const int dn = ks + 1 < KV ? d : d + 1, kn = ks + 1 < KV ? ks + 1 : 0;
#pragma unroll
        for (int t = 0; t < NT; t++)
          avn[t] = ((NZV.vol[dn][t] >> kn) & 1) ? a_op(pk + L.vol(dn), NPV, t, kn, lane) : T(0);
        b_lds<T, N>(l_x, kn, g, c, bvn);
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int t = 0; t < NT; t++)
          if ((NZV.vol[d][t] >> ks) & 1) pv[i][t] = Mx<T>::mfma(a[t], b[6 + i], pv[i][t]);
#pragma unroll
      for (int r = 0; r < 3; r++) {
        const T z = sv.row(r, b);
#pragma unroll
        for (int t = 0; t < NT; t++)
          if ((NZV.vol[d][t] >> ks) & 1) acc[6 + r][t] = Mx<T>::mfma(a[t], z, acc[6 + r][t]);
      }
    }
    add_stress<T, NT>(star + d * 81, pv, acc, false);
  }

  // ---- local surface (rank-3 solvers U V^T) and the face projections for
  // the neighbours: per face f, y_j = (V^T tInt)_j projected with fluxL[f]
  // and lifted, acc += U lift(y); then z_j = (V^T_rcv tInt)_j (the receiving
  // side's V^T) projected with fluxL[f] into the projection buffer.  Two
  // passes over tInt per face keep one coefficient set in registers.
  constexpr int NPF = pairs(KS), KF = L.chain(NF), NPL = pairs(KF);
#pragma unroll 1
  for (int f = 0; f < 4; f++) {
    const T* blk = pk + L.prj(f);
    const int32_t rv = col ? s.snd_v[el * 4 + f] : -1;
    T al[KF][NT];  // this face's lift operands, requested during the projection
    auto load_lift = [&]() {
#pragma unroll
      for (int ks = 0; ks < KF; ks++)
#pragma unroll
        for (int t = 0; t < NT; t++) al[ks][t] = a_op(pk + L.lfd(f), NPL, t, ks, lane);
    };
    auto project = [&](const T* __restrict__ vt, bool on, V4<T> (&w)[3], bool lift_ahead) {
      T cv[27];
#pragma unroll
      for (int i = 0; i < 27; i++) cv[i] = on ? vt[i] : T(0);
#pragma unroll
      for (int j = 0; j < 3; j++) w[j] = zero4<T>();
      T bn[9], an = a_op(blk, NPF, 0, 0, lane);
      b_lds<T, N>(l_x, 0, g, c, bn);
#pragma unroll
      for (int ks = 0; ks < KS; ks++) {
        T b[9];
#pragma unroll
        for (int p = 0; p < 9; p++) b[p] = bn[p];
        const T a = an;
        if (ks + 1 < KS) {
          b_lds<T, N>(l_x, ks + 1, g, c, bn);
          an = a_op(blk, NPF, 0, ks + 1, lane);
        }
        if (lift_ahead && ks == (KS > 3 ? KS - 3 : 0)) load_lift();
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < 3; j++) {
          T y = T(0);
#pragma unroll
          for (int p = 0; p < 9; p++) y += cv[j * 9 + p] * b[p];
          w[j] = Mx<T>::mfma(a, y, w[j]);
        }
      }
    };
    {
      V4<T> wl[3];
      // FP64 requests the lift operands during the projection (one wave
      // per SIMD hides no latency); FP32 has no registers to spare for it
      project(s.fsf + (el * 4 + f) * 54 + 27, true, wl, sizeof(T) == 8);
      if constexpr (sizeof(T) == 4) load_lift();
      // lift: L_j = fluxT_f^T y_j (B operands straight from the y tiles),
      // then acc_q += sum_j U[q][j] L_j
      V4<T> lj[3][NT];
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int t = 0; t < NT; t++) lj[j][t] = zero4<T>();
#pragma unroll
      for (int ks = 0; ks < KF; ks++)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
          for (int t = 0; t < NT; t++) lj[j][t] = Mx<T>::mfma(al[ks][t], wl[j][ks], lj[j][t]);
      const T* fu = s.fsf + (el * 4 + f) * 54;
#pragma unroll
      for (int q = 0; q < 9; q++) {
        const T u0 = fu[q * 3], u1 = fu[q * 3 + 1], u2 = fu[q * 3 + 2];
#pragma unroll
        for (int t = 0; t < NT; t++) acc[q][t] += u0 * lj[0][t] + u1 * lj[1][t] + u2 * lj[2][t];
      }
    }
    {
      V4<T> ws[3];
      project(s.fsf + (s.n_el * 4 + (rv >= 0 ? rv : 0)) * 54 + 27, rv >= 0, ws, false);
      if (rv >= 0) {
        T* pj = x.proj + (el * 4 + f) * 48;
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const int m = Mx<T>::row(0, g, v);
            if (m < NF) pj[j * 16 + m] = ws[j][v];
          }
      }
    }
  }

  // ---- DOFs: Q + acc
  if (col) add_out<T, N, NT>(dofs + el * SZ, acc, g);
}

// ---------------------------------------------------------------- neigh
template <typename T, int O>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kWaves<T>, kWaves<T>)))
void k_neigh_ec(SolverDev<T> s, EcDev<T> x, int64_t first, int64_t n, T* __restrict__ dofs) {
  constexpr int N = n_modes(O), NF = n_face_modes(O), NT = tiles(N), SZ = 9 * N;
  constexpr int KF = ksteps(NF), NPL = pairs(KF);
  constexpr ec::Layout L{O, Mx<T>::f32};
  __shared__ __align__(16) T l_r[ec::kRotSize];  // face-orientation maps [4][16][16]
  __shared__ __align__(16) T l_w[3 * 16 * 16];   // rotated projections [j][face mode][16]
  const int64_t nw = (n + 15) / 16;
  const int64_t w = wave_index(nw, s.xcd_map);
  if (w >= nw) return;
  const int64_t e0 = first + 16 * w;
  const int ne = (int)min((int64_t)16, first + n - e0);
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const bool col = c < ne;
  const int64_t el = e0 + (col ? c : 0);
  for (int i = lane; i < ec::kRotSize; i += 64) l_r[i] = x.rot[i];
  V4<T> acc[9][NT];
#pragma unroll
  for (int q = 0; q < 9; q++)
#pragma unroll
    for (int t = 0; t < NT; t++) acc[q][t] = zero4<T>();
  // lanes < 48 hold (j, e): the projection tile j of element e's face,
  // requested one face ahead
  const int jy = lane >> 4, ey = lane & 15;
  auto fetch = [&](int fa, T (&y)[16], int& rot) {
    int32_t slot = -1;
    rot = 0;
    if (lane < 48 && ey < ne) {
      slot = s.rcv[(e0 + ey) * 4 + fa];
      const int fm = s.nb.fmat[(e0 + ey) * 4 + fa];
      rot = fm < 4 ? 0 : 1 + (fm - 4) / 4;
    }
    const T* py = x.proj + (int64_t)(slot >= 0 ? slot : 0) * 48 + jy * 16;
#pragma unroll
    for (int i = 0; i < 16; i++) y[i] = slot >= 0 ? py[i] : T(0);
  };
  T yn[16];
  int rotn;
  fetch(0, yn, rotn);
#pragma unroll 1
  for (int fa = 0; fa < 4; fa++) {
    T y[16];
#pragma unroll
    for (int i = 0; i < 16; i++) y[i] = yn[i];
    const int rot = rotn;
    T al[KF][NT];  // lift operands and the neighbour-side U of this face
#pragma unroll
    for (int ks = 0; ks < KF; ks++)
#pragma unroll
      for (int t = 0; t < NT; t++) al[ks][t] = a_op(x.pk + L.lfn(fa), NPL, t, ks, lane);
    T u[27];
    const T* fu = s.fsf + ((s.n_el + el) * 4 + fa) * 54;
#pragma unroll
    for (int i = 0; i < 27; i++) u[i] = fu[i];
    if (fa + 1 < 4) fetch(fa + 1, yn, rotn);
    __syncthreads();  // the maps are in; the previous face's lifts have read l_w
    if (lane < 48) {
      const T* rm = l_r + rot * 256;
      // w[k'] = sum_k y[k] R[k][k'] within each degree block
#pragma unroll
      for (int p = 0; p < O; p++) {
        const int b0 = p * (p + 1) / 2;
#pragma unroll
        for (int i = 0; i <= p; i++) {
          T z = T(0);
#pragma unroll
          for (int jj = 0; jj <= p; jj++) z += y[b0 + jj] * rm[(b0 + jj) * 16 + b0 + i];
          l_w[(jy * 16 + b0 + i) * 16 + ey] = z;
        }
      }
#pragma unroll
      for (int k = NF; k < 16; k++) l_w[(jy * 16 + k) * 16 + ey] = T(0);
    }
    __syncthreads();
    V4<T> lj[3][NT];
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int t = 0; t < NT; t++) lj[j][t] = zero4<T>();
#pragma unroll
    for (int ks = 0; ks < KF; ks++) {
      T b[3];
#pragma unroll
      for (int j = 0; j < 3; j++) b[j] = l_w[(j * 16 + 4 * ks + g) * 16 + c];
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int t = 0; t < NT; t++) lj[j][t] = Mx<T>::mfma(al[ks][t], b[j], lj[j][t]);
    }
#pragma unroll
    for (int q = 0; q < 9; q++)
#pragma unroll
      for (int t = 0; t < NT; t++) acc[q][t] += u[q * 3] * lj[0][t] + u[q * 3 + 1] * lj[1][t] + u[q * 3 + 2] * lj[2][t];
  }
  if (col) add_out<T, N, NT>(dofs + el * SZ, acc, g);
}

template <typename T, int O>
int launch_local_o(const SolverDev<T>& s, const EcDev<T>& x, int64_t first, int64_t n, T dt, T* dofs, T* t_int,
                   hipStream_t st) {
  const int64_t nw = (n + 15) / 16;
  const int64_t grid = s.xcd_map ? 8 * ((nw + 7) / 8) : nw;
  hipLaunchKernelGGL((k_local_ec<T, O>), dim3((unsigned)grid), dim3(64), 0, st, s, x, first, n, dt, dofs, t_int);
  return hipGetLastError() == hipSuccess ? EDGE_OK : fail(EDGE_ERR_HIP, "element-batched local launch");
}

template <typename T, int O>
int launch_neigh_o(const SolverDev<T>& s, const EcDev<T>& x, int64_t first, int64_t n, T* dofs, hipStream_t st) {
  const int64_t nw = (n + 15) / 16;
  const int64_t grid = s.xcd_map ? 8 * ((nw + 7) / 8) : nw;
  hipLaunchKernelGGL((k_neigh_ec<T, O>), dim3((unsigned)grid), dim3(64), 0, st, s, x, first, n, dofs);
  return hipGetLastError() == hipSuccess ? EDGE_OK : fail(EDGE_ERR_HIP, "element-batched neighbour launch");
}

}  // namespace

bool ec_supported(int order, int n_runs) { return n_runs == 1 && order >= 2 && order <= 5; }

template <typename T>
int launch_local_ec(const SolverDev<T>& s, const EcDev<T>& x, int64_t first, int64_t n, T dt, T* dofs, T* t_int,
                    hipStream_t st) {
  if (s.n_runs != 1 || !x.pk || !x.rot || !x.proj)
    return fail(EDGE_ERR_UNSUPPORTED, "element-batched kernels: one run and their packed operands");
  if (n <= 0) return EDGE_OK;
  switch (s.order) {
    case 2: return launch_local_o<T, 2>(s, x, first, n, dt, dofs, t_int, st);
    case 3: return launch_local_o<T, 3>(s, x, first, n, dt, dofs, t_int, st);
    case 4: return launch_local_o<T, 4>(s, x, first, n, dt, dofs, t_int, st);
    case 5: return launch_local_o<T, 5>(s, x, first, n, dt, dofs, t_int, st);
    default: return fail(EDGE_ERR_UNSUPPORTED, "element-batched kernels: orders 2-5");
  }
}

template <typename T>
int launch_neigh_ec(const SolverDev<T>& s, const EcDev<T>& x, int64_t first, int64_t n, T* dofs, hipStream_t st) {
  if (s.n_runs != 1 || !x.pk || !x.rot || !x.proj)
    return fail(EDGE_ERR_UNSUPPORTED, "element-batched kernels: one run and their packed operands");
  if (n <= 0) return EDGE_OK;
  switch (s.order) {
    case 2: return launch_neigh_o<T, 2>(s, x, first, n, dofs, st);
    case 3: return launch_neigh_o<T, 3>(s, x, first, n, dofs, st);
    case 4: return launch_neigh_o<T, 4>(s, x, first, n, dofs, st);
    case 5: return launch_neigh_o<T, 5>(s, x, first, n, dofs, st);
    default: return fail(EDGE_ERR_UNSUPPORTED, "element-batched kernels: orders 2-5");
  }
}

template int launch_local_ec<double>(const SolverDev<double>&, const EcDev<double>&, int64_t, int64_t, double,
                                     double*, double*, hipStream_t);
template int launch_local_ec<float>(const SolverDev<float>&, const EcDev<float>&, int64_t, int64_t, float, float*,
                                    float*, hipStream_t);
template int launch_neigh_ec<double>(const SolverDev<double>&, const EcDev<double>&, int64_t, int64_t, double*,
                                     hipStream_t);
template int launch_neigh_ec<float>(const SolverDev<float>&, const EcDev<float>&, int64_t, int64_t, float*,
                                    hipStream_t);

}  // namespace edge
