// FP64 single-run element kernels with 16 elements per wave (FP64 MFMA).
//
// With one run per element there is no run axis to fill an MFMA tile.  These
// kernels put 16 elements on the N dimension of v_mfma_f64_16x16x4f64
// instead: the shared reference-element matrices are the A operands, and
// every product with one is a full 16 x 16 tile per quantity:
//   A operand (16 x 4)  lane (g, c) = M[4 s + g][16 t + c]     (shared, packed)
//   B operand (4 x 16)  lane (g, c) = X_e[q][4 s + g]          (element e = c)
//   D (16 x 16)         lane (g, c), register v = Out_e[q][16 t + g + 4 v]
// with g = lane / 16, c = lane % 16.  Register v of a D tile t is the B
// operand of k-step 4 t + v, so product chains (the CK's derivatives, the
// lifts of the face projections) stay in registers.  The 9 quantities of an
// (element, mode) position live in one lane: the per-element star matrices
// and flux-solver factors act lane-locally, with per-lane coefficients.
// The elements' DOFs pass through a 40 KiB LDS tile [q][mode][16] (one wave
// per SIMD).
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

#include "ec64_pack.hpp"
#include "kernels.hpp"

namespace edge {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));

__device__ __forceinline__ v4d mfma(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ v4d zero4() { return v4d{0.0, 0.0, 0.0, 0.0}; }

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
struct VelRows {
  double v[9];
  __device__ __forceinline__ void load(const double* __restrict__ sd) {
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int j = 0; j < 3; j++) v[r * 3 + j] = sd[(6 + r) * 9 + kVel[r][j]];
  }
  __device__ __forceinline__ double row(int r, const double (&b)[9]) const {
    return v[r * 3] * b[kVel[r][0]] + v[r * 3 + 1] * b[kVel[r][1]] + v[r * 3 + 2] * b[kVel[r][2]];
  }
};

// out_q (+)= sum_i S_d[q][6 + i] P_i for the stress rows q = 0..5 (D tiles)
template <int TN>
__device__ __forceinline__ void add_stress(const double* __restrict__ sd, const v4d (&p)[3][TN], v4d (&out)[9][TN],
                                           bool first) {
#pragma unroll
  for (int q = 0; q < 6; q++) {
    double cf[3];
#pragma unroll
    for (int j = 0; j < kStrNnz[q]; j++) cf[j] = sd[q * 9 + 6 + kStr[q][j]];
#pragma unroll
    for (int t = 0; t < TN; t++) {
      v4d z = first ? zero4() : out[q][t];
#pragma unroll
      for (int j = 0; j < kStrNnz[q]; j++) z += cf[j] * p[kStr[q][j]][t];
      out[q][t] = z;
    }
  }
}

// A operand of k-step s, tile t of a packed product (ec::Layout) with np
// k-step pairs per tile
__device__ __forceinline__ double a_op(const double* __restrict__ blk, int np, int t, int s, int lane) {
  return blk[((t * np + (s >> 1)) * 64 + lane) * 2 + (s & 1)];
}

// B values of k-step s from the LDS tile: the 9 quantities' rows 4 s + g of
// column c.  Rows past a quantity's N read the next quantity's first rows,
// or the zero rows past the tile: always finite, and the A operand is zero
// there.
template <int N>
__device__ __forceinline__ void b_lds(const double* __restrict__ l_x, int s, int g, int c, double (&b)[9]) {
#pragma unroll
  for (int p = 0; p < 9; p++) b[p] = l_x[(p * N + 4 * s + g) * 16 + c];
}

// CK stage DE >= 2 from the previous derivative in registers (row 4 s + g of
// k-step s is register s % 4 of tile s / 4), elastic pattern: per direction
// the three raw velocity products and the three star-mixed velocity rows.
template <int O, int DE>
__device__ __forceinline__ void ck_next(const double* __restrict__ pk, const double* __restrict__ star,
                                        const v4d (&prev)[9][tiles(n_modes(O - DE + 1))], v4d (&cur)[9][1],
                                        int lane) {
  constexpr int NK = n_modes(O - DE + 1), KS = ksteps(NK), NP = pairs(KS);
  static_assert(tiles(n_modes(O - DE)) == 1, "derivatives of stage 2 and up fit one tile");
  constexpr ec::Layout L{O};
#pragma unroll
  for (int r = 6; r < 9; r++) cur[r][0] = zero4();
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double* blk = pk + L.ck(DE, d);
    VelRows sv;
    sv.load(star + d * 81);
    v4d pv[3][1];
#pragma unroll
    for (int i = 0; i < 3; i++) pv[i][0] = zero4();
#pragma unroll
    for (int s = 0; s < KS; s++) {
      __builtin_amdgcn_sched_barrier(0);
      const double a = a_op(blk, NP, 0, s, lane);
      double b[9];
#pragma unroll
      for (int p = 0; p < 9; p++) b[p] = prev[p][s / 4][s % 4];
#pragma unroll
      for (int i = 0; i < 3; i++) pv[i][0] = mfma(a, b[6 + i], pv[i][0]);
#pragma unroll
      for (int r = 0; r < 3; r++) cur[6 + r][0] = mfma(a, sv.row(r, b), cur[6 + r][0]);
    }
    add_stress<1>(star + d * 81, pv, cur, d == 0);
  }
}

// tInt (LDS) += x for rows 16 t + g + 4 v below `rows` of column c
template <int N>
__device__ __forceinline__ void add_tint(double* __restrict__ l_x, int rows, int t, const v4d (&x)[9], int g,
                                         int c) {
#pragma unroll
  for (int q = 0; q < 9; q++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int m = 16 * t + g + 4 * v;
      if (m < rows) l_x[(q * N + m) * 16 + c] += x[q][v];
    }
}

// CK stages DE >= 2; each adds sca dt^{DE} / (DE + 1)! der_DE to tInt (LDS).
template <int O, int DE>
__device__ __forceinline__ void ck_rest(const double* __restrict__ pk, const double* __restrict__ star,
                                        const v4d (&prev)[9][tiles(n_modes(O - DE + 1))], double* __restrict__ l_x,
                                        double dt, double sca, int g, int c, int lane) {
  if constexpr (DE < O) {
    v4d cur[9][1];
    ck_next<O, DE>(pk, star, prev, cur, lane);
    sca *= dt / double(DE + 1);
    v4d xt[9];
#pragma unroll
    for (int q = 0; q < 9; q++) xt[q] = sca * cur[q][0];
    add_tint<n_modes(O)>(l_x, n_modes(O - DE), 0, xt, g, c);
    ck_rest<O, DE + 1>(pk, star, cur, l_x, dt, sca, g, c, lane);
  }
}

// The per-element tiles [9][N] of elements e0 .. e0 + ne - 1 (contiguous in
// the reference layout) into the LDS tile [q][mode][16], zeros for missing
// columns.
template <int SZ>
__device__ __forceinline__ void gather(const double* __restrict__ src, int ne, double* __restrict__ l_x, int lane) {
  constexpr int IT = (SZ * 16 + 63) / 64, B = 8;
#pragma unroll 1
  for (int i0 = 0; i0 < IT; i0 += B) {
    double v[B];
#pragma unroll
    for (int j = 0; j < B; j++) {
      const int idx = (i0 + j) * 64 + lane, i = idx >> 4, e = idx & 15;
      v[j] = (i0 + j < IT && idx < SZ * 16 && e < ne) ? src[(size_t)e * SZ + i] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < B; j++) {
      const int idx = (i0 + j) * 64 + lane;
      if (i0 + j < IT && idx < SZ * 16) l_x[idx] = v[j];
    }
  }
}

// The LDS tile [q][mode][16] back to the elements' tiles (transposed).
template <int SZ>
__device__ __forceinline__ void scatter(const double* __restrict__ l_x, int ne, double* __restrict__ dst, int lane) {
  constexpr int IT = (SZ * 16 + 63) / 64;
#pragma unroll 8
  for (int j = 0; j < IT; j++) {
    const int idx = j * 64 + lane, i = idx >> 4, e = idx & 15;
    if (idx < SZ * 16 && e < ne) dst[(size_t)e * SZ + i] = l_x[idx];
  }
}

// dofs of this lane's element (rows 16 t + g + 4 v) += acc
template <int N, int NT>
__device__ __forceinline__ void add_out(double* __restrict__ q0, const v4d (&acc)[9][NT], int g) {
  double o[9][NT][4];
#pragma unroll
  for (int q = 0; q < 9; q++)
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int m = 16 * t + g + 4 * v;
        o[q][t][v] = m < N ? q0[q * N + m] : 0.0;
      }
#pragma unroll
  for (int q = 0; q < 9; q++)
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int m = 16 * t + g + 4 * v;
        if (m < N) q0[q * N + m] = o[q][t][v] + acc[q][t][v];
      }
}

// ---------------------------------------------------------------- local
template <int O>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_local_ec64(SolverDev<double> s, Ec64Dev x, int64_t first, int64_t n, double dt, double* __restrict__ dofs,
                  double* __restrict__ t_int) {
  constexpr int N = n_modes(O), NF = n_face_modes(O), NT = tiles(N), SZ = 9 * N;
  constexpr int NN1 = n_modes(O - 1), T1 = tiles(NN1), KS = ksteps(N), NV = n_modes(O - 1), KV = ksteps(NV);
  constexpr ec::Layout L{O};
  // Q, then tInt; zero rows past the tile for the k-step tail of the last
  // quantity (rows up to 8 N + 4 KS - 1)
  constexpr int PADR = 4 * KS - N > 0 ? 4 * KS - N : 1;
  __shared__ __align__(16) double l_x[(SZ + PADR) * 16];
  const int64_t nw = (n + 15) / 16;
  const int64_t w = wave_index(nw, s.xcd_map);
  if (w >= nw) return;  // one wave per workgroup: no barrier is left waiting
  const int64_t e0 = first + 16 * w;
  const int ne = (int)min((int64_t)16, first + n - e0);
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const bool col = c < ne;
  const int64_t el = e0 + (col ? c : 0);  // this lane's element (a valid one on padding columns)
  const double* star = s.star + el * 243;
  const double* pk = x.pk;

  gather<SZ>(dofs + e0 * SZ, ne, l_x, lane);
  for (int i = lane; i < PADR * 16; i += 64) l_x[SZ * 16 + i] = 0.0;
  __syncthreads();

  // ---- CK stage 1: per direction the raw velocity products and the
  // star-mixed velocity rows; the LDS rows become dt Q (tInt's first term)
  v4d der[9][T1];
  {
    v4d pv[3][3][T1];
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int t = 0; t < T1; t++) pv[d][i][t] = zero4();
#pragma unroll
    for (int r = 6; r < 9; r++)
#pragma unroll
      for (int t = 0; t < T1; t++) der[r][t] = zero4();
    VelRows sv[3];
#pragma unroll
    for (int d = 0; d < 3; d++) sv[d].load(star + d * 81);
    constexpr int NP = pairs(KS);
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
      __builtin_amdgcn_sched_barrier(0);  // bounds register pressure: no hoisting across k-steps
      double b[9];
      b_lds<N>(l_x, ks, g, c, b);
      const int k = 4 * ks + g;
      if (k < N) {
#pragma unroll
        for (int p = 0; p < 9; p++) l_x[(p * N + k) * 16 + c] = dt * b[p];
      }
#pragma unroll
      for (int d = 0; d < 3; d++) {
        double a[T1];
#pragma unroll
        for (int t = 0; t < T1; t++) a[t] = a_op(pk + L.ck(1, d), NP, t, ks, lane);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
          for (int t = 0; t < T1; t++) pv[d][i][t] = mfma(a[t], b[6 + i], pv[d][i][t]);
#pragma unroll
        for (int r = 0; r < 3; r++) {
          const double z = sv[d].row(r, b);
#pragma unroll
          for (int t = 0; t < T1; t++) der[6 + r][t] = mfma(a[t], z, der[6 + r][t]);
        }
      }
    }
#pragma unroll
    for (int d = 0; d < 3; d++) add_stress<T1>(star + d * 81, pv[d], der, d == 0);
  }
  // tInt = dt Q + sum_de dt^{de+1} / (de+1)! der_de, summed in LDS
  {
    const double sca = dt * dt / 2.0;
#pragma unroll
    for (int t = 0; t < T1; t++) {
      v4d xt[9];
#pragma unroll
      for (int q = 0; q < 9; q++) xt[q] = sca * der[q][t];
      add_tint<N>(l_x, NN1, t, xt, g, c);
    }
    ck_rest<O, 2>(pk, star, der, l_x, dt, sca, g, c, lane);
  }
  __syncthreads();
  scatter<SZ>(l_x, ne, t_int + e0 * SZ, lane);

  // ---- volume: acc_q += sum_d stiffV_d^T (S_d tInt)_q over tInt's first NV modes
  v4d acc[9][NT];
#pragma unroll
  for (int q = 0; q < 9; q++)
#pragma unroll
    for (int t = 0; t < NT; t++) acc[q][t] = zero4();
#pragma unroll
  for (int d = 0; d < 3; d++) {
    constexpr int NP = pairs(KV);
    const double* blk = pk + L.vol(d);
    VelRows sv;
    sv.load(star + d * 81);
    v4d pv[3][NT];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int t = 0; t < NT; t++) pv[i][t] = zero4();
#pragma unroll
    for (int ks = 0; ks < KV; ks++) {
      __builtin_amdgcn_sched_barrier(0);
      double b[9], a[NT];
      b_lds<N>(l_x, ks, g, c, b);
#pragma unroll
      for (int t = 0; t < NT; t++) a[t] = a_op(blk, NP, t, ks, lane);
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int t = 0; t < NT; t++) pv[i][t] = mfma(a[t], b[6 + i], pv[i][t]);
#pragma unroll
      for (int r = 0; r < 3; r++) {
        const double z = sv.row(r, b);
#pragma unroll
        for (int t = 0; t < NT; t++) acc[6 + r][t] = mfma(a[t], z, acc[6 + r][t]);
      }
    }
    add_stress<NT>(star + d * 81, pv, acc, false);
  }

  // ---- local surface (rank-3 solvers U V^T) and the face projections for
  // the neighbours: per face f, y_j = (V^T tInt)_j projected with fluxL[f]
  // and lifted, acc += U lift(y); then z_j = (V^T_rcv tInt)_j (the receiving
  // side's V^T) projected with fluxL[f] into the projection buffer.  Two
  // passes over tInt per face keep one coefficient set in registers.
  constexpr int NPF = pairs(KS), KF = ksteps(NF), NPL = pairs(KF);
#pragma unroll 1
  for (int f = 0; f < 4; f++) {
    const double* blk = pk + L.prj(f);
    const int32_t rv = col ? s.snd_v[el * 4 + f] : -1;
    auto project = [&](const double* __restrict__ vt, bool on, v4d (&w)[3]) {
      double cv[27];
#pragma unroll
      for (int i = 0; i < 27; i++) cv[i] = on ? vt[i] : 0.0;
#pragma unroll
      for (int j = 0; j < 3; j++) w[j] = zero4();
#pragma unroll
      for (int ks = 0; ks < KS; ks++) {
        __builtin_amdgcn_sched_barrier(0);
        double b[9];
        b_lds<N>(l_x, ks, g, c, b);
        const double a = a_op(blk, NPF, 0, ks, lane);
#pragma unroll
        for (int j = 0; j < 3; j++) {
          double y = 0.0;
#pragma unroll
          for (int p = 0; p < 9; p++) y += cv[j * 9 + p] * b[p];
          w[j] = mfma(a, y, w[j]);
        }
      }
    };
    {
      v4d wl[3];
      project(s.fsf + (el * 4 + f) * 54 + 27, true, wl);
      // lift: L_j = fluxT_f^T y_j (B operands straight from the y tiles),
      // then acc_q += sum_j U[q][j] L_j
      v4d lj[3][NT];
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int t = 0; t < NT; t++) lj[j][t] = zero4();
      const double* blt = pk + L.lft(f);
#pragma unroll
      for (int ks = 0; ks < KF; ks++) {
        double a[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) a[t] = a_op(blt, NPL, t, ks, lane);
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
          for (int t = 0; t < NT; t++) lj[j][t] = mfma(a[t], wl[j][ks], lj[j][t]);
      }
      const double* fu = s.fsf + (el * 4 + f) * 54;
#pragma unroll
      for (int q = 0; q < 9; q++) {
        const double u0 = fu[q * 3], u1 = fu[q * 3 + 1], u2 = fu[q * 3 + 2];
#pragma unroll
        for (int t = 0; t < NT; t++) acc[q][t] += u0 * lj[0][t] + u1 * lj[1][t] + u2 * lj[2][t];
      }
    }
    {
      v4d ws[3];
      project(s.fsf + (s.n_el * 4 + (rv >= 0 ? rv : 0)) * 54 + 27, rv >= 0, ws);
      if (rv >= 0) {
        double* pj = x.proj + (el * 4 + f) * 48;
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const int m = g + 4 * v;
            if (m < NF) pj[j * 16 + m] = ws[j][v];
          }
      }
    }
  }

  // ---- DOFs: Q + acc
  if (col) add_out<N, NT>(dofs + el * SZ, acc, g);
}

// ---------------------------------------------------------------- neigh
template <int O>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_neigh_ec64(SolverDev<double> s, Ec64Dev x, int64_t first, int64_t n, double* __restrict__ dofs) {
  constexpr int N = n_modes(O), NF = n_face_modes(O), NT = tiles(N), SZ = 9 * N;
  constexpr int KF = ksteps(NF), NPL = pairs(KF);
  constexpr ec::Layout L{O};
  __shared__ __align__(16) double l_r[ec::kRotSize];  // face-orientation maps [4][16][16]
  __shared__ __align__(16) double l_w[3 * 16 * 16];   // rotated projections [j][face mode][16]
  const int64_t nw = (n + 15) / 16;
  const int64_t w = wave_index(nw, s.xcd_map);
  if (w >= nw) return;
  const int64_t e0 = first + 16 * w;
  const int ne = (int)min((int64_t)16, first + n - e0);
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const bool col = c < ne;
  const int64_t el = e0 + (col ? c : 0);
  for (int i = lane; i < ec::kRotSize; i += 64) l_r[i] = x.rot[i];
  v4d acc[9][NT];
#pragma unroll
  for (int q = 0; q < 9; q++)
#pragma unroll
    for (int t = 0; t < NT; t++) acc[q][t] = zero4();
#pragma unroll 1
  for (int fa = 0; fa < 4; fa++) {
    __syncthreads();  // the maps are in; the previous face's lifts have read l_w
    if (lane < 48) {  // (j, e): the projection tile j of element e's face fa
      const int j = lane >> 4, e = lane & 15;
      int32_t slot = -1;
      int rot = 0;
      if (e < ne) {
        slot = s.rcv[(e0 + e) * 4 + fa];
        const int fm = s.nb.fmat[(e0 + e) * 4 + fa];
        rot = fm < 4 ? 0 : 1 + (fm - 4) / 4;
      }
      double y[16];
      if (slot >= 0) {
        const v2d* py = reinterpret_cast<const v2d*>(x.proj + (int64_t)slot * 48 + j * 16);
#pragma unroll
        for (int i = 0; i < 8; i++) {
          const v2d t = py[i];
          y[2 * i] = t[0];
          y[2 * i + 1] = t[1];
        }
      } else {
#pragma unroll
        for (int i = 0; i < 16; i++) y[i] = 0.0;
      }
      const double* rm = l_r + rot * 256;
      // w[k'] = sum_k y[k] R[k][k'] within each degree block
#pragma unroll
      for (int p = 0; p < O; p++) {
        const int b0 = p * (p + 1) / 2;
#pragma unroll
        for (int i = 0; i <= p; i++) {
          double z = 0.0;
#pragma unroll
          for (int jj = 0; jj <= p; jj++) z += y[b0 + jj] * rm[(b0 + jj) * 16 + b0 + i];
          l_w[(j * 16 + b0 + i) * 16 + e] = z;
        }
      }
#pragma unroll
      for (int k = NF; k < 16; k++) l_w[(j * 16 + k) * 16 + e] = 0.0;
    }
    __syncthreads();
    v4d lj[3][NT];
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int t = 0; t < NT; t++) lj[j][t] = zero4();
    const double* blt = x.pk + L.lft(fa);
#pragma unroll
    for (int ks = 0; ks < KF; ks++) {
      double a[NT], b[3];
#pragma unroll
      for (int t = 0; t < NT; t++) a[t] = a_op(blt, NPL, t, ks, lane);
#pragma unroll
      for (int j = 0; j < 3; j++) b[j] = l_w[(j * 16 + 4 * ks + g) * 16 + c];
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int t = 0; t < NT; t++) lj[j][t] = mfma(a[t], b[j], lj[j][t]);
    }
    const double* fu = s.fsf + ((s.n_el + el) * 4 + fa) * 54;
#pragma unroll
    for (int q = 0; q < 9; q++) {
      const double u0 = fu[q * 3], u1 = fu[q * 3 + 1], u2 = fu[q * 3 + 2];
#pragma unroll
      for (int t = 0; t < NT; t++) acc[q][t] += u0 * lj[0][t] + u1 * lj[1][t] + u2 * lj[2][t];
    }
  }
  if (col) add_out<N, NT>(dofs + el * SZ, acc, g);
}

template <int O>
int launch_local_o(const SolverDev<double>& s, const Ec64Dev& x, int64_t first, int64_t n, double dt, double* dofs,
                   double* t_int, hipStream_t st) {
  const int64_t nw = (n + 15) / 16;
  const int64_t grid = s.xcd_map ? 8 * ((nw + 7) / 8) : nw;
  hipLaunchKernelGGL(k_local_ec64<O>, dim3((unsigned)grid), dim3(64), 0, st, s, x, first, n, dt, dofs, t_int);
  return hipGetLastError() == hipSuccess ? EDGE_OK : fail(EDGE_ERR_HIP, "element-batched FP64 local launch");
}

template <int O>
int launch_neigh_o(const SolverDev<double>& s, const Ec64Dev& x, int64_t first, int64_t n, double* dofs,
                   hipStream_t st) {
  const int64_t nw = (n + 15) / 16;
  const int64_t grid = s.xcd_map ? 8 * ((nw + 7) / 8) : nw;
  hipLaunchKernelGGL(k_neigh_ec64<O>, dim3((unsigned)grid), dim3(64), 0, st, s, x, first, n, dofs);
  return hipGetLastError() == hipSuccess ? EDGE_OK : fail(EDGE_ERR_HIP, "element-batched FP64 neighbour launch");
}

}  // namespace

bool ec64_supported(int order, int n_runs, int prec_bytes) {
  return prec_bytes == 8 && n_runs == 1 && order >= 2 && order <= 5;
}

int launch_local_ec64(const SolverDev<double>& s, const Ec64Dev& x, int64_t first, int64_t n, double dt,
                      double* dofs, double* t_int, hipStream_t st) {
  if (n <= 0) return EDGE_OK;
  switch (s.order) {
    case 2: return launch_local_o<2>(s, x, first, n, dt, dofs, t_int, st);
    case 3: return launch_local_o<3>(s, x, first, n, dt, dofs, t_int, st);
    case 4: return launch_local_o<4>(s, x, first, n, dt, dofs, t_int, st);
    case 5: return launch_local_o<5>(s, x, first, n, dt, dofs, t_int, st);
    default: return fail(EDGE_ERR_UNSUPPORTED, "element-batched FP64 kernels: orders 2-5");
  }
}

int launch_neigh_ec64(const SolverDev<double>& s, const Ec64Dev& x, int64_t first, int64_t n, double* dofs,
                      hipStream_t st) {
  if (n <= 0) return EDGE_OK;
  switch (s.order) {
    case 2: return launch_neigh_o<2>(s, x, first, n, dofs, st);
    case 3: return launch_neigh_o<3>(s, x, first, n, dofs, st);
    case 4: return launch_neigh_o<4>(s, x, first, n, dofs, st);
    case 5: return launch_neigh_o<5>(s, x, first, n, dofs, st);
    default: return fail(EDGE_ERR_UNSUPPORTED, "element-batched FP64 kernels: orders 2-5");
  }
}

}  // namespace edge
