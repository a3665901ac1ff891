// MFMA viscoelastic element kernels for 8 fused runs in FP32 (gfx950).
//
// One workgroup per element.  The products with the reference-element
// matrices (stiffness and flux matrices) run on v_mfma_f32_16x16x4f32 with
// the runs on the N dimension.  Eight runs fill half of the 16 columns, so
// every MFMA carries two products that share its A operand (the same
// reference matrix) in its column halves: lanes with c = lane % 16 < 8 one
// product, lanes with c >= 8 another, both for run c % 8.
//   A operand (16 x 4)  lane (g, c) = M[k(g)][16 t + c]
//   B operand (4 x 16)  lane (g, c) = X_h[k(g)][c % 8],  h = c / 8
//   D (16 x 16)         lane (g, c), register v: row 16 t + 4 g + v, product h, run c % 8
// with g = lane / 16.  Everything that mixes quantities at one (mode, run)
// position (star matrices, anelastic source and relaxation terms, flux-solver
// factors) is elementwise: the positions are spread over all threads of the
// workgroup, with the anelastic state in registers, and the raw products
// move between the MFMA and the elementwise phases through LDS.
//
// Local step, eight waves (one position per thread at O5); six of them take
// a direction and a half of the output row tiles each in the CK and volume
// products, and four a face each in the surface integral
// (src/impl/seismic/solvers/AderDg.hpp:362-533):
//  CK (kernels/TimePredVanilla.hpp:200-306; with attenuation every
//  derivative keeps all N modes): per stage de the raw products of
//  der_{de-1} with the negated transposed stiffness matrix of direction d,
//    P_{d,i} = der_{de-1}[6 + i] (-K_d^T M^-1)                  (i = 0..2)
//    V_{d,r} = (sum_p S_d[6 + r][p] der_{de-1}[p]) (-K_d^T M^-1)  (r = 0..2)
//  (the elastic star pattern: stress rows read only velocities, velocity
//  rows only stresses), then per position
//    der_de[q < 6] = sum_d sum_i S_d[q][6 + i] P_{d,i} + sum_rm E_rm derA_rm,
//    der_de[6 + r] = sum_d V_{d,r},   A = sum_d S^A_d P_d,
//    derA_rm <- w_rm (A + derA_rm),   tInt += c_de der_de,   tIntA_rm += c_de derA_rm;
//  volume (kernels/VolIntVanilla.hpp:189-241): the same raw products of tInt
//  with K_d M^-1, Q[q < 6] += sum_d S_d P_d + sum_rm E_rm tIntA_rm,
//  Q[6 + r] += sum_d V_d, QA_rm += w_rm (sum_d S^A_d P_d - tIntA_rm);
//  local surface (kernels/SurfIntVanilla.hpp:212-262, SurfInt.hpp:139-156)
//  with the rank-3 factors fs = U V^T and fsA = U_A V^T:
//    w_j = (V^T tInt)_j F^L_fa,  L_j = w_j F^T_fa,
//    Q += sum_j U[.][j] L_j,  QA_rm += w_rm sum_j U_A[.][j] L_j.
// Neighbour step (kernels/SurfIntVanilla.hpp:274-357), four waves, wave f
// lifts face f's projection of the neighbour's tInt (F^N and the
// neighbour-side factors): the projection the neighbour's local step wrote
// (the local kernel's waves with spare issue make them beside the volume and
// surface products), or one made here from the neighbour's tInt tile; the lifted
// tiles of the four faces are combined per position.
// Summation order differs from the reference's loops; results agree to
// rounding.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "kernels.hpp"

namespace edge {
namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(4))) float cfloat;
typedef const __attribute__((address_space(4))) int32_t cint32;
typedef const __attribute__((address_space(4))) uint8_t cuint8;
constexpr int R = 8;
// Local kernel: eight waves, so that two workgroups fill a CU's four SIMDs
// evenly (two waves of each per SIMD): waves are placed round robin from the
// first SIMD, and six-wave workgroups left two SIMDs with twice the waves of
// the other two.
constexpr int kLocalThreads = 512, kNeighThreads = 256;

__device__ __forceinline__ v4f mfma(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4f zero4() { return v4f{0.f, 0.f, 0.f, 0.f}; }

// Row tiles with at most 4 rows (the last tile of 20 or 35 modes) run on
// v_mfma_f32_4x4x1f32 (16 blocks of 4 x 4 x 1, half the cycles of the
// 16 x 16 x 4 tile it replaces): block 4 g + c / 4 takes the lane group's k
// (4 s + g), rows 0..3 and columns 4 (c / 4) .. + 3, so the A operand of
// lane (g, c) is M[4 s + g][16 t + c % 4] and the B operand is unchanged;
// the four k-groups are summed across g at the end, after which lane
// (g, c), register v holds row 16 t + v, column c.
__device__ __forceinline__ v4f mfma44(float a, float b, v4f c) {
  return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 0, 0, 0);
}
constexpr bool narrow_tile(int nn, int t) { return nn - 16 * t <= 4; }
constexpr int a_col(int nn, int t, int c) { return narrow_tile(nn, t) ? (c & 3) : c; }
__device__ __forceinline__ v4f sum_groups(v4f x) {
#pragma unroll
  for (int v = 0; v < 4; v++) {
    x[v] += __shfl_xor(x[v], 16);
    x[v] += __shfl_xor(x[v], 32);
  }
  return x;
}
constexpr int tiles(int rows) { return (rows + 15) / 16; }
constexpr int ksteps(int rows) { return (rows + 3) / 4; }

// Per-element tables that no kernel writes come in through the constant
// address space: wave-uniform values by scalar loads.
__device__ __forceinline__ const cfloat* as_const(const float* p) { return (const cfloat*)p; }

// blockIdx -> element offset: each of the 8 XCDs (workgroups dispatched round
// robin) gets a contiguous element range, so neighbours share its L2 (the
// autotuner may switch this schedule off: workgroup i -> element i).  The
// grid holds 8 ceil(n / 8) workgroups.
__device__ __forceinline__ int64_t xcd_element(int64_t n) {
  const int64_t chunk = (n + 7) / 8;
  const int64_t b = blockIdx.x;
  return (b & 7) * chunk + (b >> 3);
}

// Guarded load without a branch: the index is clamped into the valid range
// and the value masked afterwards.
__device__ __forceinline__ float ldg(const float* __restrict__ p, int idx, bool ok) {
  const float v = p[ok ? idx : 0];
  return ok ? v : 0.f;
}

// Elastic star pattern (src/impl/seismic/setups/Elasticity.cpp:172-250):
// velocity row 6 + r reads the stresses kVelCols[r], stress row q the
// velocities 6 + kStrVel[q][0 .. kStrNnz[q] - 1].
constexpr int kVelCols[3][3] = {{0, 3, 5}, {1, 3, 4}, {2, 4, 5}};
constexpr int kStrNnz[6] = {3, 3, 3, 2, 2, 2};
constexpr int kStrVel[6][3] = {{0, 1, 2}, {0, 1, 2}, {0, 1, 2}, {0, 1, 0}, {1, 2, 0}, {0, 2, 0}};
// Anelastic star pattern S^A_d [6][3] (src/impl/seismic/setups/
// ViscoElasticity.cpp:233-290): row q reads the velocities kAnCol[q][..].
constexpr int kAnNnz[6] = {1, 1, 1, 2, 2, 2};
constexpr int kAnCol[6][2] = {{0, 0}, {1, 0}, {2, 0}, {0, 1}, {1, 2}, {0, 2}};
// Anelastic source pattern E_rm [6][6] (src/impl/seismic/setups/
// ViscoElasticity.h:379-458): the normal-stress block rows 0..2 x columns
// 0..2 and the diagonal of rows 3..5.

template <int O>
struct VDims {
  static constexpr int N = n_modes(O), NV = n_modes(O - 1), NF = n_face_modes(O);
  static constexpr int NP = N * R;  // positions (mode, run) of one quantity
  static constexpr int SZ = 9 * NP, SZA = 6 * NP;
};

// The per-element coefficients the elementwise phases use, gathered in LDS
// in the order they are used (each workgroup copies its element's at the
// start).  Read from the star, anelastic star and source matrices where
// they lie, the pattern's scattered entries became a chain of small scalar
// loads, each waited for in full before the next FMAs; from LDS every
// thread reads the same address (a broadcast) and the reads pipeline.
//   [kCS + 15 d ..)  stress rows: S_d[q][6 + kStrVel[q][j]]   (15 per direction)
//   [kCA + 9 d ..)   anelastic star: S^A_d[q][kAnCol[q][j]]   (9 per direction)
//   [kCZ + 9 d ..)   star-combined velocity rows: S_d[6 + i][kVelCols[i][j]]
//   [kCE + 12 rm ..) source matrix E_rm: normal block [3][3], then the diagonal 3..5
constexpr int kCS = 0, kCA = 45, kCZ = 72, kCE = 99;
constexpr int n_coef(int nrm) { return kCE + 12 * nrm; }
constexpr int kStrOff[6] = {0, 3, 6, 9, 11, 13}, kAnOff[6] = {0, 1, 2, 3, 5, 7};
struct CoefTab {
  // index into [star 243 | anelastic star 54 | source matrices 36 n_rms]
  short src[n_coef(3)] = {};
  constexpr CoefTab() {
    int i = 0;
    for (int d = 0; d < 3; d++)
      for (int q = 0; q < 6; q++)
        for (int j = 0; j < kStrNnz[q]; j++) src[i++] = (short)(d * 81 + q * 9 + 6 + kStrVel[q][j]);
    for (int d = 0; d < 3; d++)
      for (int q = 0; q < 6; q++)
        for (int j = 0; j < kAnNnz[q]; j++) src[i++] = (short)(243 + d * 18 + q * 3 + kAnCol[q][j]);
    for (int d = 0; d < 3; d++)
      for (int r = 0; r < 3; r++)
        for (int j = 0; j < 3; j++) src[i++] = (short)(d * 81 + (6 + r) * 9 + kVelCols[r][j]);
    for (int rm = 0; rm < 3; rm++) {
      for (int q = 0; q < 3; q++)
        for (int j = 0; j < 3; j++) src[i++] = (short)(297 + rm * 36 + q * 6 + j);
      for (int q = 3; q < 6; q++) src[i++] = (short)(297 + rm * 36 + q * 6 + q);
    }
  }
};
constexpr CoefTab kCoef{};

// A face's rank-3 solver factors for the surface terms, gathered in LDS the
// same way: U (9 x 3) then U_A (6 x 3), 45 values per face.
constexpr int kUF = 45;
__device__ __forceinline__ void gather_u(float* __restrict__ lu, const float* __restrict__ fs,
                                         const float* __restrict__ fsa, int i) {
  const int f = i / kUF, k = i - f * kUF;
  lu[i] = k < 27 ? fs[f * 54 + k] : fsa[f * 18 + k - 27];
}

// The coefficients' LDS base, opaque to the optimiser: read where they are
// used instead of hoisted out of the time-derivative loop into registers.
__device__ __forceinline__ const float* coef_base(const float* l_c) {
  int z = 0;
  asm volatile("" : "+s"(z));
  return l_c + z;
}

// Velocity rows of the star-combined tile of direction d at one position:
// z_{d,i} = sum_j S_d[6 + i][kVelCols[i][j]] y[kVelCols[i][j]].
__device__ __forceinline__ float vel_row(const float* __restrict__ lc, int d, int i, const float (&y)[9]) {
  const float* z = lc + kCZ + 9 * d + 3 * i;
  return z[0] * y[kVelCols[i][0]] + z[1] * y[kVelCols[i][1]] + z[2] * y[kVelCols[i][2]];
}

// The B rows of the products at position p of an LDS tile x ([18][N][R]):
// the velocity rows 6..8 of y, the stress rows 0..5 when `stress`, and the
// star-combined velocity rows z_{d,i} of the three directions in rows
// 9 + 3 d + i.  One thread per position forms each z once; formed inside
// the products, every k-step of both waves of a direction would form them
// again (four times per position).
__device__ __forceinline__ void store_rows(float* __restrict__ x, const float* __restrict__ lc,
                                           const float (&y)[9], int np, int p, bool stress) {
#pragma unroll
  for (int q = 0; q < 9; q++)
    if (q >= 6 || stress) x[q * np + p] = y[q];
#pragma unroll
  for (int d = 0; d < 3; d++)
#pragma unroll
    for (int i = 0; i < 3; i++) x[(9 + 3 * d + i) * np + p] = vel_row(lc, d, i, y);
}

// raw_products (below) with the B rows ready-made by store_rows: three LDS
// reads per k-step and lane, no VALU work.
template <int O, int NK, int NN, int TT>
__device__ __forceinline__ void raw_products_zr(const float* __restrict__ m, const float* __restrict__ x,
                                                float* __restrict__ out, int dir, int t0, int g, int c,
                                                const float* apre) {
  using D = VDims<O>;
  constexpr int N = D::N, NP = D::NP, KS = ksteps(NK);
  const int h = c >> 3, r = c & 7;
  // this lane's B rows: (P_0 | P_1) 6 | 7, (P_2 | V_0) 8 | z_0, (V_1 | V_2) z_1 | z_2
  const float* xb[3] = {x + (h ? 7 : 6) * NP + r, x + (h ? 9 + 3 * dir : 8) * NP + r,
                        x + (10 + 3 * dir + h) * NP + r};
  v4f acc[TT][3];
#pragma unroll
  for (int t = 0; t < TT; t++)
#pragma unroll
    for (int p = 0; p < 3; p++) acc[t][p] = zero4();
  float aall[apre ? 1 : KS][TT];
  if (!apre) {
#pragma unroll
    for (int s = 0; s < KS; s++)
#pragma unroll
      for (int t = 0; t < TT; t++) {
        const int k = 4 * s + g;
        const int col = 16 * (t0 + t) + (TT == 1 ? a_col(NN, t0, c) : c);
        aall[s][t] = ldg(m, k * N + col, k < NK && col < NN);
      }
  }
  auto load = [&](int s, float (&a)[TT], float (&b)[3]) {
    const int k = 4 * s + g;
    if (apre) {
      a[0] = apre[s];
    } else {
#pragma unroll
      for (int t = 0; t < TT; t++) a[t] = aall[s][t];
    }
#pragma unroll
    for (int p = 0; p < 3; p++) {
      if (4 * s + 3 < NK)  // constant after unrolling
        b[p] = xb[p][k * R];
      else
        b[p] = k < NK ? xb[p][(k < NK ? k : 0) * R] : 0.f;
    }
  };
  // NAR: one narrow tile (TT = 1, at most 4 rows), on the 4 x 4 blocks
  auto run = [&](auto nar) {
    constexpr bool NAR = decltype(nar)::value;
    float an[TT], bn[3];
    load(0, an, bn);
#pragma unroll
    for (int s = 0; s < KS; s++) {
      float a[TT], b[3];
#pragma unroll
      for (int t = 0; t < TT; t++) a[t] = an[t];
#pragma unroll
      for (int p = 0; p < 3; p++) b[p] = bn[p];
      if (s + 1 < KS) load(s + 1, an, bn);
      __builtin_amdgcn_sched_barrier(0);  // bounds register pressure: no hoisting across k-steps
#pragma unroll
      for (int t = 0; t < TT; t++)
#pragma unroll
        for (int p = 0; p < 3; p++) acc[t][p] = NAR ? mfma44(a[t], b[p], acc[t][p]) : mfma(a[t], b[p], acc[t][p]);
    }
    if constexpr (NAR) {
#pragma unroll
      for (int p = 0; p < 3; p++) acc[0][p] = sum_groups(acc[0][p]);
    }
#pragma unroll
    for (int t = 0; t < TT; t++)
#pragma unroll
      for (int p = 0; p < 3; p++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          const int n = 16 * (t0 + t) + (NAR ? 0 : 4 * g) + v;
          if (n < NN && (!NAR || g == 0)) out[(2 * p + h) * NP + n * R + r] = acc[t][p][v];
        }
  };
  if (TT == 1 && narrow_tile(NN, t0))
    run(std::true_type{});
  else
    run(std::false_type{});
}

// Raw products of one direction d: rows k < NK of the tile x (LDS,
// [9][N][R], and with ZR the star-combined rows of store_rows) with the
// matrix m ([N][N], rows k, columns n < NN), output row tiles t0 .. t0 +
// TT - 1, written to out (LDS, [6][N][R]) as P_0, P_1, P_2 (velocity tiles)
// and V_0, V_1, V_2 (velocity rows of the star-combined tile).  Column-half
// pairs: (P_0 | P_1), (P_2 | V_0), (V_1 | V_2), so product 2 p + h sits in
// half h of pair p.  Without ZR each lane combines the star rows from the
// nine quantities per k-step.
// apre (optional): the A operands of every k-step of one tile (TT = 1),
// loaded by the caller ahead of time.
template <int O, int NK, int NN, int TT, bool ZR>
__device__ __forceinline__ void raw_products(const float* __restrict__ m, const cfloat* __restrict__ sd,
                                             const float* __restrict__ x, float* __restrict__ out, int dir, int t0,
                                             int g, int c, const float* apre = nullptr) {
  if constexpr (ZR) {
    raw_products_zr<O, NK, NN, TT>(m, x, out, dir, t0, g, c, apre);
    return;
  }
  using D = VDims<O>;
  constexpr int N = D::N, NP = D::NP, KS = ksteps(NK);
  const int h = c >> 3, r = c & 7;
  float sv[9];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) sv[i * 3 + j] = sd[(6 + i) * 9 + kVelCols[i][j]];
  v4f acc[TT][3];
#pragma unroll
  for (int t = 0; t < TT; t++)
#pragma unroll
    for (int p = 0; p < 3; p++) acc[t][p] = zero4();
  // without apre every k-step's A operands are requested up front: one
  // memory round trip instead of one per k-step
  float aall[apre ? 1 : KS][TT];
  if (!apre) {
#pragma unroll
    for (int s = 0; s < KS; s++)
#pragma unroll
      for (int t = 0; t < TT; t++) {
        const int k = 4 * s + g;
        aall[s][t] = ldg(m, k * N + 16 * (t0 + t) + c, k < NK && 16 * (t0 + t) + c < NN);
      }
  }
  auto load = [&](int s, float (&a)[TT], float (&b)[9]) {
    const int k = 4 * s + g;
    if (apre) {
      a[0] = apre[s];
    } else {
#pragma unroll
      for (int t = 0; t < TT; t++) a[t] = aall[s][t];
    }
#pragma unroll
    for (int q = 0; q < 9; q++) {
      if (4 * s + 3 < NK)  // constant after unrolling
        b[q] = x[(q * N + k) * R + r];
      else
        b[q] = k < NK ? x[(q * N + (k < NK ? k : 0)) * R + r] : 0.f;
    }
  };
  float an[TT], bn[9];
  load(0, an, bn);
#pragma unroll
  for (int s = 0; s < KS; s++) {
    float a[TT], b[9];
#pragma unroll
    for (int t = 0; t < TT; t++) a[t] = an[t];
#pragma unroll
    for (int q = 0; q < 9; q++) b[q] = bn[q];
    if (s + 1 < KS) load(s + 1, an, bn);
    __builtin_amdgcn_sched_barrier(0);  // bounds register pressure: no hoisting across k-steps
    float z[3];
#pragma unroll
    for (int i = 0; i < 3; i++)
      z[i] = sv[i * 3] * b[kVelCols[i][0]] + sv[i * 3 + 1] * b[kVelCols[i][1]] + sv[i * 3 + 2] * b[kVelCols[i][2]];
    const float b0 = h ? b[7] : b[6], b1 = h ? z[0] : b[8], b2 = h ? z[2] : z[1];
#pragma unroll
    for (int t = 0; t < TT; t++) {
      acc[t][0] = mfma(a[t], b0, acc[t][0]);
      acc[t][1] = mfma(a[t], b1, acc[t][1]);
      acc[t][2] = mfma(a[t], b2, acc[t][2]);
    }
  }
#pragma unroll
  for (int t = 0; t < TT; t++)
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int n = 16 * (t0 + t) + 4 * g + v;
        if (n < NN) out[(2 * p + h) * NP + n * R + r] = acc[t][p][v];
      }
}

// Face projections w_j = (V^T x)_j F for one face: x ([9][N][R], LDS or
// global), the projection matrix F ([N][NF]) and the rows V^T ([3][9],
// wave-uniform) of the face's rank-3 solver factors.  Column-half pairs:
// w01 = (w_0 | w_1) and w2 = (w_2 | w_2) as D tiles (face mode 4 g + v).
template <int O>
__device__ __forceinline__ void face_project(const float* __restrict__ fl, const cfloat* __restrict__ vt,
                                             const float* x, v4f& w01, v4f& w2, int g, int c) {
  using D = VDims<O>;
  constexpr int N = D::N, NF = D::NF, KS = ksteps(N);
  const int h = c >> 3, r = c & 7;
  // per lane the row of its column half (V^T row h) and row 2
  float ch[9], c2[9];
#pragma unroll
  for (int p = 0; p < 9; p++) {
    const float v0 = vt[p], v1 = vt[9 + p];
    ch[p] = h ? v1 : v0;
    c2[p] = vt[18 + p];
  }
  // every k-step's A operand requested up front (one memory round trip)
  float aall[KS];
#pragma unroll
  for (int s = 0; s < KS; s++) {
    const int k = 4 * s + g;
    aall[s] = ldg(fl, k * NF + c, k < N && c < NF);
  }
  auto load = [&](int s, float& a, float (&b)[9]) {
    const int k = 4 * s + g;
    a = aall[s];
#pragma unroll
    for (int q = 0; q < 9; q++) {
      if (4 * s + 3 < N)  // constant after unrolling
        b[q] = x[(q * N + k) * R + r];
      else
        b[q] = k < N ? x[(q * N + (k < N ? k : 0)) * R + r] : 0.f;
    }
  };
  w01 = zero4();
  w2 = zero4();
  float an, bn[9];
  load(0, an, bn);
#pragma unroll
  for (int s = 0; s < KS; s++) {
    const float a = an;
    float b[9];
#pragma unroll
    for (int q = 0; q < 9; q++) b[q] = bn[q];
    if (s + 1 < KS) load(s + 1, an, bn);
    __builtin_amdgcn_sched_barrier(0);
    float yh = 0.f, y2 = 0.f;
#pragma unroll
    for (int q = 0; q < 9; q++) {
      yh += ch[q] * b[q];
      y2 += c2[q] * b[q];
    }
    w01 = mfma(a, yh, w01);
    w2 = mfma(a, y2, w2);
  }
}

// Lifted face tiles L_j = w_j F^T ([3][N][R] into out, LDS) from the paired
// projections (face_project) and the lift F^T ([NF][N]); the B operands come
// straight from the projections' D tiles.
template <int O>
__device__ __forceinline__ void face_lift(const float* __restrict__ ft, v4f w01, v4f w2, float* out, int g, int c) {
  using D = VDims<O>;
  constexpr int N = D::N, NF = D::NF, NP = D::NP, NT = tiles(N);
  const int h = c >> 3, r = c & 7;
  v4f l01[NT], l2[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) {
    l01[t] = zero4();
    l2[t] = zero4();
  }
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int fm = 4 * g + v;
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const int col = 16 * t + a_col(N, t, c);
      const float at = ldg(ft, fm * N + col, fm < NF && col < N);
      if (narrow_tile(N, t)) {  // constant after unrolling
        l01[t] = mfma44(at, w01[v], l01[t]);
        l2[t] = mfma44(at, w2[v], l2[t]);
      } else {
        l01[t] = mfma(at, w01[v], l01[t]);
        l2[t] = mfma(at, w2[v], l2[t]);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const bool NAR = narrow_tile(N, t);  // constant after unrolling
    if (NAR) {
      l01[t] = sum_groups(l01[t]);
      l2[t] = sum_groups(l2[t]);
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int n = 16 * t + (NAR ? 0 : 4 * g) + v;
      if (n < N && (!NAR || g == 0)) {
        out[h * NP + n * R + r] = l01[t][v];
        if (h == 0) out[2 * NP + n * R + r] = l2[t][v];
      }
    }
  }
}

// Face projections exchanged between the local and the neighbour step: per
// element face the two paired D tiles, lane-major (lane l holds its four
// values of w01 at [l][0..3] and of w2 at [64 + l][0..3]).
constexpr int kProjFace = 512;

__device__ __forceinline__ void store_proj(float* __restrict__ p, v4f w01, v4f w2, int lane) {
  reinterpret_cast<v4f*>(p)[lane] = w01;
  reinterpret_cast<v4f*>(p)[64 + lane] = w2;
}

// Per position: stress rows (6), velocity rows (3) and the anelastic
// combination A (6) from the raw products of the three directions
// (l_p: [3][6][N][R]).
template <int NP>
__device__ __forceinline__ void combine_raw(const float* __restrict__ l_p, const float* __restrict__ lc, int p,
                                            float (&y)[9], float (&a)[6]) {
#pragma unroll
  for (int q = 0; q < 9; q++) y[q] = 0.f;
#pragma unroll
  for (int q = 0; q < 6; q++) a[q] = 0.f;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    float pv[3];
#pragma unroll
    for (int i = 0; i < 3; i++) pv[i] = l_p[(d * 6 + i) * NP + p];
#pragma unroll
    for (int r = 0; r < 3; r++) y[6 + r] += l_p[(d * 6 + 3 + r) * NP + p];
    const float* sd = lc + kCS + 15 * d;
    const float* sa = lc + kCA + 9 * d;
#pragma unroll
    for (int q = 0; q < 6; q++) {
#pragma unroll
      for (int j = 0; j < kStrNnz[q]; j++) y[q] += sd[kStrOff[q] + j] * pv[kStrVel[q][j]];
#pragma unroll
      for (int j = 0; j < kAnNnz[q]; j++) a[q] += sa[kAnOff[q] + j] * pv[kAnCol[q][j]];
    }
  }
}

// y[0..5] += E x for an anelastic source matrix E of the source pattern
// (its 12 entries as gathered in LDS at lc + kCE + 12 rm)
__device__ __forceinline__ void add_source(const float* __restrict__ e, const float (&x)[6], float (&y)[9]) {
#pragma unroll
  for (int q = 0; q < 3; q++) y[q] += e[q * 3] * x[0] + e[q * 3 + 1] * x[1] + e[q * 3 + 2] * x[2];
#pragma unroll
  for (int q = 3; q < 6; q++) y[q] += e[6 + q] * x[q];
}

// TL: tIntA in LDS (each thread's own positions, no barriers) instead of
// registers, so that five waves per SIMD fit the register file (the
// occupancy variant; LDS then allows three workgroups per CU at O5).
// NB: first the previous time step's neighbouring surface integral from the
// projections in proj_in (written by the previous launch; s.proj receives
// this step's, so the two alternate between two buffers): its lifted tiles
// update Q and QA before the CK and go back to the DOFs at once, where the
// output stage reads them, so the neighbour step of step k - 1 and the
// local step of step k run as one launch.
// Registers: at most 128 VGPRs (four waves per SIMD), so that two eight-wave
// workgroups share a CU; at 129 only one fits and the launch takes 1.5x.
template <int O, int NRM, bool TL, bool NB = false>
__global__ __launch_bounds__(kLocalThreads) __attribute__((amdgpu_waves_per_eu(TL ? 5 : 4)))
void k_local_vm(SolverDev<float> s, int64_t first, int64_t n, float dt, float* __restrict__ dofs,
                float* __restrict__ dofs_a, float* __restrict__ t_int, const float* __restrict__ proj_in) {
  static_assert(!(NB && TL), "the fused step keeps tIntA in registers");
  using D = VDims<O>;
  constexpr int N = D::N, NV = D::NV, NP = D::NP, SZ = D::SZ, SZA = D::SZA;
  // elementwise phases: one (mode, run) position per thread, compact in the
  // first waves (fewer waves issue the phase's instructions than with the
  // positions spread out)
  constexpr int PJ = 1;
  static_assert(NP <= kLocalThreads, "one position per thread");
  // ZR: the products' star-combined B rows formed per position (store_rows);
  // the occupancy variant keeps the smaller tile (three workgroups per CU)
  constexpr bool ZR = !TL;
  __shared__ __align__(16) float l_x[ZR ? 2 * SZ : SZ];  // Q, then der_{de-1}, then tInt (+ z rows)
  __shared__ __align__(16) float l_p[18 * NP];  // raw products [3][6][N][R]; lifted tiles [4][3][N][R]
  __shared__ __align__(16) float l_ta[TL ? NRM * 6 * NP : 1];  // tIntA [rm][6][N][R] (TL)
  __shared__ __align__(16) float l_c[n_coef(NRM)];              // the element's coefficients (CoefTab order)
  __shared__ __align__(16) float l_u[(NB ? 8 : 4) * kUF];        // U, U_A per face: local side, (NB) neighbour side
  const int64_t off = s.xcd_map ? xcd_element(n) : (int64_t)blockIdx.x;
  if (off >= n) return;
  const int64_t el = first + off;
  const int tid = threadIdx.x, lane = tid & 63, g = lane >> 4, c = lane & 15;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave index, uniform
  // The six product tasks (three directions x two halves of the row tiles)
  // run on logical waves 0..5: waves 0..5 in even workgroups and waves
  // 0..3, 6, 7 in odd ones (by blockIdx / 256, which alternates between the
  // workgroups of one CU in the first dispatch round), so that the two
  // workgroups of a CU put three task waves on every SIMD; logical waves
  // 6.. have no task.
  const int odd = (int)((blockIdx.x >> 8) & 1);
  const int lw = (odd && wv >= 4) ? (wv >= 6 ? wv - 2 : wv + 4) : wv;
  const int pos = tid;  // this thread's position (none from NP on)
  const int dir = lw % 3, t_ck = lw / 3;  // CK / volume products: direction, half of the row tiles
  const cfloat* star = as_const(s.star + el * 243);  // the products' own star rows (without ZR)
  if (tid < n_coef(NRM)) {
    const int v = kCoef.src[tid];
    l_c[tid] = v < 243 ? s.star[el * 243 + v]
                       : (v < 297 ? s.star_a[el * 54 + v - 243] : s.src_a[el * NRM * 36 + v - 297]);
  }
  if (tid < 4 * kUF) gather_u(l_u, s.fsf + el * 4 * 54, s.fsfa + el * 4 * 18, tid);
  if (NB && tid >= 4 * kUF && tid < 8 * kUF)
    gather_u(l_u + 4 * kUF, s.fsf + (s.n_el + el) * 4 * 54, s.fsfa + (s.n_el + el) * 4 * 18, tid - 4 * kUF);
  const cfloat* rfs = as_const(s.rfs);
  const float* q0 = dofs + el * SZ;
  float* qa0 = dofs_a + el * NRM * SZA;
  const cfloat* fsf = as_const(s.fsf + el * 4 * 54);

  // Q: per position into registers (ZR), or copied to LDS as a whole
  float qv[ZR ? 9 : 1][PJ];
  if constexpr (ZR) {
#pragma unroll
    for (int j = 0; j < PJ; j++) {
      const int p = pos + j;
#pragma unroll
      for (int q = 0; q < 9; q++) qv[q][j] = p < NP ? q0[q * NP + p] : 0.f;
    }
  } else {
    const float4* gq = reinterpret_cast<const float4*>(q0);
    float4* lq = reinterpret_cast<float4*>(l_x);
    for (int i = tid; i < SZ / 4; i += kLocalThreads) lq[i] = gq[i];
  }
  // the previous step's neighbour lifts, face wv < 4 by wave wv (NB)
  if constexpr (NB) {
    if (lw < 4) {
      const int32_t slot = ((const cint32*)s.rcv)[el * 4 + wv];
      if (slot >= 0) {  // else outflow (uniform)
        const v4f* pw = reinterpret_cast<const v4f*>(proj_in + (int64_t)slot * kProjFace);
        face_lift<O>(s.dg.fluxT + wv * D::NF * N, pw[lane], pw[64 + lane], l_p + wv * 3 * NP, g, c);
      }
    }
  }
  // anelastic state per position: derA (CK), tIntA (registers, or l_ta)
  float da[NRM][6][PJ], ta_r[TL ? 1 : NRM][6][PJ], ti[9][PJ];
  auto ta = [&](int rm, int q, int j) -> float& {
    if constexpr (TL)
      return l_ta[(rm * 6 + q) * NP + pos + j];
    else
      return ta_r[rm][q][j];
  };
#pragma unroll
  for (int j = 0; j < PJ; j++) {
    const int p = pos + j;
    if (p < NP) {
#pragma unroll
      for (int rm = 0; rm < NRM; rm++)
#pragma unroll
        for (int q = 0; q < 6; q++) da[rm][q][j] = qa0[(rm * 6 + q) * NP + p];
    }
  }
  __syncthreads();  // the coefficients (and the lifted tiles, or Q in LDS)
  // NB: Q += sum_f U_f L_f, QA_rm += w_rm sum_f U_A,f L_f (the neighbour-side
  // factors), written back to the DOFs at once: the output stage then reads
  // them as it reads Q and QA otherwise, and no register holds the lifted
  // sums across the CK
  if constexpr (NB) {
    float nq[9][PJ], nqa[6][PJ];
    const cint32* rcv = (const cint32*)(s.rcv + el * 4);
    const float* lun = coef_base(l_u) + 4 * kUF;
#pragma unroll
    for (int j = 0; j < PJ; j++) {
      const int p = pos + j;
#pragma unroll
      for (int q = 0; q < 9; q++) nq[q][j] = 0.f;
#pragma unroll
      for (int q = 0; q < 6; q++) nqa[q][j] = 0.f;
      if (p < NP) {
#pragma unroll
        for (int f = 0; f < 4; f++) {
          if (rcv[f] < 0) continue;  // outflow (uniform)
          const float l0 = l_p[(f * 3) * NP + p], l1 = l_p[(f * 3 + 1) * NP + p], l2 = l_p[(f * 3 + 2) * NP + p];
          const float* u = lun + f * kUF;
          const float* uan = u + 27;
#pragma unroll
          for (int q = 0; q < 9; q++) nq[q][j] += u[q * 3] * l0 + u[q * 3 + 1] * l1 + u[q * 3 + 2] * l2;
#pragma unroll
          for (int q = 0; q < 6; q++) nqa[q][j] += uan[q * 3] * l0 + uan[q * 3 + 1] * l1 + uan[q * 3 + 2] * l2;
        }
        float* gw = dofs + el * SZ;
#pragma unroll
        for (int q = 0; q < 9; q++) {
          qv[q][j] += nq[q][j];
          gw[q * NP + p] = qv[q][j];
        }
#pragma unroll
        for (int rm = 0; rm < NRM; rm++) {
          const float w = rfs[rm];
#pragma unroll
          for (int q = 0; q < 6; q++) {
            da[rm][q][j] += w * nqa[q][j];
            qa0[(rm * 6 + q) * NP + p] = da[rm][q][j];
          }
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < PJ; j++) {
    const int p = pos + j;
    if (p < NP) {
#pragma unroll
      for (int rm = 0; rm < NRM; rm++)
#pragma unroll
        for (int q = 0; q < 6; q++) ta(rm, q, j) = dt * da[rm][q][j];
    }
  }
#pragma unroll
  for (int j = 0; j < PJ; j++) {
    const int p = pos + j;
    if constexpr (ZR) {
      float y[9];
#pragma unroll
      for (int q = 0; q < 9; q++) {
        y[q] = qv[q][j];
        ti[q][j] = dt * y[q];
      }
      if (p < NP) store_rows(l_x, l_c, y, NP, p, false);
    } else {
#pragma unroll
      for (int q = 0; q < 9; q++) ti[q][j] = p < NP ? dt * l_x[q * NP + p] : 0.f;
    }
  }
  if constexpr (ZR) __syncthreads();  // the CK's first B rows (and, NB, the lifted tiles read)

  // ---- CK; the wave's stiffness A operands (one row tile, every k-step)
  // are the same in every stage
  float a_ck[ksteps(N)];
#pragma unroll
  for (int st = 0; st < ksteps(N); st++) {
    const int k = 4 * st + g;
    const int col = 16 * t_ck + (ZR ? a_col(NV, t_ck, c) : c);
    a_ck[st] = ldg(s.dg.stiffTn + dir * N * N, k * N + col, k < N && col < NV);
  }
  float sca = dt;
#pragma unroll 1
  for (int de = 1; de < O; de++) {
    if (t_ck < tiles(NV))
      raw_products<O, N, NV, 1, ZR>(s.dg.stiffTn + dir * N * N, star + dir * 81, l_x, l_p + dir * 6 * NP, dir,
                                    t_ck, g, c, a_ck);
    __syncthreads();
    sca *= dt / float(de + 1);
    const float* lc = coef_base(l_c);
#pragma unroll
    for (int j = 0; j < PJ; j++) {
      const int p = pos + j;
      if (p < NP) {
        float y[9], a[6];
        if (p < NV * R) {
          combine_raw<NP>(l_p, lc, p, y, a);
        } else {  // top-degree modes: no stiffness contribution
#pragma unroll
          for (int q = 0; q < 9; q++) y[q] = 0.f;
#pragma unroll
          for (int q = 0; q < 6; q++) a[q] = 0.f;
        }
#pragma unroll
        for (int rm = 0; rm < NRM; rm++) {
          float x[6];
#pragma unroll
          for (int q = 0; q < 6; q++) x[q] = da[rm][q][j];
          add_source(lc + kCE + 12 * rm, x, y);
          const float w = rfs[rm];
#pragma unroll
          for (int q = 0; q < 6; q++) {
            const float v = w * (a[q] + x[q]);
            da[rm][q][j] = v;
            ta(rm, q, j) += sca * v;
          }
        }
#pragma unroll
        for (int q = 0; q < 9; q++) ti[q][j] += sca * y[q];
        if constexpr (ZR) {
          store_rows(l_x, lc, y, NP, p, false);
        } else {
#pragma unroll
          for (int q = 0; q < 9; q++) l_x[q * NP + p] = y[q];
        }
      }
    }
    __syncthreads();
  }

  // ---- tInt to LDS and global memory
  float* gti = t_int + el * SZ;
#pragma unroll
  for (int j = 0; j < PJ; j++) {
    const int p = pos + j;
    if (p < NP) {
      float y[9];
#pragma unroll
      for (int q = 0; q < 9; q++) {
        y[q] = ti[q][j];
        l_x[q * NP + p] = y[q];
        gti[q * NP + p] = y[q];
      }
      if constexpr (ZR) store_rows(l_x, l_c, y, NP, p, false);
    }
  }
  __syncthreads();

  // this element's tInt as the neighbour through face f sees it (its flux
  // matrix and neighbour-side solver), for the neighbour step
  auto nb_proj = [&](int f) {
    const int32_t rv = ((const cint32*)s.snd_v)[el * 4 + f];
    if (rv >= 0) {  // somebody receives through this face (uniform)
      const int fm = ((const cuint8*)s.snd_fm)[el * 4 + f];
      v4f w01, w2;
      face_project<O>(s.dg.fluxLN + fm * N * D::NF, as_const(s.fsf + ((int64_t)s.n_el * 4 + rv) * 54 + 27), l_x,
                      w01, w2, g, c);
      store_proj(s.proj + (el * 4 + f) * kProjFace, w01, w2, lane);
    }
  };

  // ---- volume; the waves with fewer output tiles also project faces 0..2
  // for the neighbours
  {
    // output tiles: the first half of the waves [0, TV0), the second [TV0, tiles(N))
    constexpr int TN = tiles(N), TV0 = (TN + 1) / 2;
    if (t_ck == 0) {
      raw_products<O, NV, N, TV0, ZR>(s.dg.stiffV + dir * N * N, star + dir * 81, l_x, l_p + dir * 6 * NP, dir, 0,
                                      g, c);
    } else if (t_ck == 1) {
      if constexpr (TN > TV0)
        raw_products<O, NV, N, TN - TV0, ZR>(s.dg.stiffV + dir * N * N, star + dir * 81, l_x, l_p + dir * 6 * NP,
                                             dir, TV0, g, c);
      if (s.proj) nb_proj(dir);
    }
  }
  __syncthreads();
  float aq[9][PJ], aa[6][PJ];
#pragma unroll
  for (int j = 0; j < PJ; j++) {
    const int p = pos + j;
    float y[9], a[6];
    if (p < NP) {
      combine_raw<NP>(l_p, l_c, p, y, a);
#pragma unroll
      for (int rm = 0; rm < NRM; rm++) {
        float x[6];
#pragma unroll
        for (int q = 0; q < 6; q++) x[q] = ta(rm, q, j);
        add_source(l_c + kCE + 12 * rm, x, y);
      }
    } else {
#pragma unroll
      for (int q = 0; q < 9; q++) y[q] = 0.f;
#pragma unroll
      for (int q = 0; q < 6; q++) a[q] = 0.f;
    }
#pragma unroll
    for (int q = 0; q < 9; q++) aq[q][j] = y[q];
#pragma unroll
    for (int q = 0; q < 6; q++) aa[q][j] = a[q];
  }
  __syncthreads();

  // ---- Q and QA for the output, requested before the surface integral
  float* gq = dofs + el * SZ;
  float qin[9][PJ], qain[NRM][6][PJ];
#pragma unroll
  for (int j = 0; j < PJ; j++) {
    const int p = pos + j;
    if (p < NP) {
#pragma unroll
      for (int q = 0; q < 9; q++) qin[q][j] = gq[q * NP + p];
#pragma unroll
      for (int rm = 0; rm < NRM; rm++)
#pragma unroll
        for (int q = 0; q < 6; q++) qain[rm][q][j] = qa0[(rm * 6 + q) * NP + p];
    }
  }

  // ---- local surface: waves 0..3 the lifted tiles of face wv in
  // l_p[wv][3][N][R]; wave 4 the neighbours' projection through face 3
  if (lw < 4) {
    v4f w01, w2;
    face_project<O>(s.dg.fluxLN + wv * N * D::NF, fsf + wv * 54 + 27, l_x, w01, w2, g, c);
    face_lift<O>(s.dg.fluxT + wv * D::NF * N, w01, w2, l_p + wv * 3 * NP, g, c);
  } else if (lw == 4 && s.proj) {
    nb_proj(3);
  }
  __syncthreads();

  // ---- Q += volume + source + surface; QA_rm += w_rm (A - tIntA_rm + surface)
#pragma unroll
  for (int j = 0; j < PJ; j++) {
    const int p = pos + j;
    if (p < NP) {
      float y[9], ua[6];
#pragma unroll
      for (int q = 0; q < 9; q++) y[q] = aq[q][j];
#pragma unroll
      for (int q = 0; q < 6; q++) ua[q] = aa[q][j];
#pragma unroll
      for (int f = 0; f < 4; f++) {
        const float l0 = l_p[(f * 3) * NP + p], l1 = l_p[(f * 3 + 1) * NP + p], l2 = l_p[(f * 3 + 2) * NP + p];
        const float* u = l_u + f * kUF;
        const float* uan = u + 27;
#pragma unroll
        for (int q = 0; q < 9; q++) y[q] += u[q * 3] * l0 + u[q * 3 + 1] * l1 + u[q * 3 + 2] * l2;
#pragma unroll
        for (int q = 0; q < 6; q++) ua[q] += uan[q * 3] * l0 + uan[q * 3 + 1] * l1 + uan[q * 3 + 2] * l2;
      }
#pragma unroll
      for (int q = 0; q < 9; q++) gq[q * NP + p] = qin[q][j] + y[q];
#pragma unroll
      for (int rm = 0; rm < NRM; rm++) {
        const float w = rfs[rm];
#pragma unroll
        for (int q = 0; q < 6; q++) qa0[(rm * 6 + q) * NP + p] = qain[rm][q][j] + w * (ua[q] - ta(rm, q, j));
      }
    }
  }
}

// Neighbour step.  PROJ: the face projections the local step wrote (no tInt
// reads); else the neighbours' tInt tiles, projected here.
template <int O, int NRM, bool PROJ>
__global__ __launch_bounds__(kNeighThreads) void k_neigh_vm(SolverDev<float> s, int64_t first, int64_t n,
                                                            const float* __restrict__ t_int,
                                                            float* __restrict__ dofs, float* __restrict__ dofs_a) {
  using D = VDims<O>;
  constexpr int N = D::N, NF = D::NF, NP = D::NP, SZ = D::SZ, SZA = D::SZA;
  constexpr int PJ = (NP + kNeighThreads - 1) / kNeighThreads;
  // per face: (without projections) the neighbour's tInt, then (aliased) its
  // lifted tiles [3][N][R]
  __shared__ __align__(16) float l_n[4][PROJ ? 3 * NP : SZ];
  const int64_t off = s.xcd_map ? xcd_element(n) : (int64_t)blockIdx.x;
  if (off >= n) return;
  const int64_t el = first + off;
  const int tid = threadIdx.x, lane = tid & 63, g = lane >> 4, c = lane & 15;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave index = face, uniform
  const cint32* nes = (const cint32*)(s.nb.ne + el * 4);
  const cint32* rcv = (const cint32*)(s.rcv + el * 4);
  const cfloat* fsf = as_const(s.fsf + (s.n_el + el) * 4 * 54);
  // the neighbour-side U, U_A per face, gathered in LDS (read by broadcast)
  __shared__ __align__(16) float l_u[4 * kUF];
  if (tid < 4 * kUF) gather_u(l_u, s.fsf + (s.n_el + el) * 4 * 54, s.fsfa + (s.n_el + el) * 4 * 18, tid);
  float* gq = dofs + el * SZ;
  float* gqa = dofs_a + el * NRM * SZA;
  // this element's Q and QA at the thread's first position, requested
  // before the lifts; the further positions (a few threads at O5) are read
  // after them, so that fewer registers are held across the lifts
  float qin[9], qain[NRM][6];
  auto load_q = [&](int p) {
#pragma unroll
    for (int q = 0; q < 9; q++) qin[q] = gq[q * NP + p];
#pragma unroll
    for (int rm = 0; rm < NRM; rm++)
#pragma unroll
      for (int q = 0; q < 6; q++) qain[rm][q] = gqa[(rm * 6 + q) * NP + p];
  };
  if (tid < NP) load_q(tid);
  if constexpr (PROJ) {
    const int32_t slot = rcv[wv];
    if (slot >= 0) {  // else outflow (uniform)
      const v4f* p = reinterpret_cast<const v4f*>(s.proj + (int64_t)slot * kProjFace);
      face_lift<O>(s.dg.fluxT + wv * NF * N, p[lane], p[64 + lane], l_n[wv], g, c);
    }
  } else {
    const int32_t ne = nes[wv];
    if (ne >= 0) {  // uniform across the wave
      const float4* gt = reinterpret_cast<const float4*>(t_int + (int64_t)ne * SZ);
      float4* lt = reinterpret_cast<float4*>(l_n[wv]);
      for (int i = lane; i < SZ / 4; i += 64) lt[i] = gt[i];
    }
    __syncthreads();
    if (ne >= 0) {
      const int fm = ((const cuint8*)s.nb.fmat)[el * 4 + wv];
      v4f w01, w2;
      face_project<O>(s.dg.fluxLN + fm * N * NF, fsf + wv * 54 + 27, l_n[wv], w01, w2, g, c);
      face_lift<O>(s.dg.fluxT + wv * NF * N, w01, w2, l_n[wv], g, c);
    }
  }
  __syncthreads();
  const cfloat* rfs = as_const(s.rfs);
#pragma unroll
  for (int j = 0; j < PJ; j++) {
    const int p = tid + j * kNeighThreads;
    if (p < NP) {
      float y[9], ua[6];
#pragma unroll
      for (int q = 0; q < 9; q++) y[q] = 0.f;
#pragma unroll
      for (int q = 0; q < 6; q++) ua[q] = 0.f;
#pragma unroll
      for (int f = 0; f < 4; f++) {
        if ((PROJ ? rcv[f] : nes[f]) < 0) continue;  // outflow (uniform)
        const float l0 = l_n[f][p], l1 = l_n[f][NP + p], l2 = l_n[f][2 * NP + p];
        const float* u = l_u + f * kUF;
        const float* uan = u + 27;
#pragma unroll
        for (int q = 0; q < 9; q++) y[q] += u[q * 3] * l0 + u[q * 3 + 1] * l1 + u[q * 3 + 2] * l2;
#pragma unroll
        for (int q = 0; q < 6; q++) ua[q] += uan[q * 3] * l0 + uan[q * 3 + 1] * l1 + uan[q * 3 + 2] * l2;
      }
      if (j > 0) load_q(p);
#pragma unroll
      for (int q = 0; q < 9; q++) gq[q * NP + p] = qin[q] + y[q];
#pragma unroll
      for (int rm = 0; rm < NRM; rm++) {
        const float w = rfs[rm];
#pragma unroll
        for (int q = 0; q < 6; q++) gqa[(rm * 6 + q) * NP + p] = qain[rm][q] + w * ua[q];
      }
    }
  }
}

unsigned grid_of(int64_t n) { return (unsigned)(8 * ((n + 7) / 8)); }

template <int O, int NRM>
void launch_o(const SolverDev<float>& s, int64_t first, int64_t n, float dt, float* dofs, float* dofs_a,
              float* t_int, const float* tin, bool local, hipStream_t st) {
  if (local && s.hi_occ)
    hipLaunchKernelGGL((k_local_vm<O, NRM, true>), dim3(grid_of(n)), dim3(kLocalThreads), 0, st, s, first, n, dt,
                       dofs, dofs_a, t_int, nullptr);
  else if (local)
    hipLaunchKernelGGL((k_local_vm<O, NRM, false>), dim3(grid_of(n)), dim3(kLocalThreads), 0, st, s, first, n, dt,
                       dofs, dofs_a, t_int, nullptr);
  else if (s.proj)
    hipLaunchKernelGGL((k_neigh_vm<O, NRM, true>), dim3(grid_of(n)), dim3(kNeighThreads), 0, st, s, first, n, tin,
                       dofs, dofs_a);
  else
    hipLaunchKernelGGL((k_neigh_vm<O, NRM, false>), dim3(grid_of(n)), dim3(kNeighThreads), 0, st, s, first, n, tin,
                       dofs, dofs_a);
}

template <int O>
int launch_rm(const SolverDev<float>& s, int64_t first, int64_t n, float dt, float* dofs, float* dofs_a,
              float* t_int, const float* tin, bool local, hipStream_t st) {
  switch (s.n_rms) {
    case 1: launch_o<O, 1>(s, first, n, dt, dofs, dofs_a, t_int, tin, local, st); break;
    case 2: launch_o<O, 2>(s, first, n, dt, dofs, dofs_a, t_int, tin, local, st); break;
    case 3: launch_o<O, 3>(s, first, n, dt, dofs, dofs_a, t_int, tin, local, st); break;
    default: return fail(EDGE_ERR_UNSUPPORTED, "MFMA viscoelastic kernels: 1 to 3 mechanisms");
  }
  EDGE_HIP_TRY(hipGetLastError());
  return EDGE_OK;
}

int launch_vm(const SolverDev<float>& s, int64_t first, int64_t n, float dt, float* dofs, float* dofs_a,
              float* t_int, const float* tin, bool local, hipStream_t st) {
  if (n <= 0) return EDGE_OK;
  if (!s.fsf || !s.fsfa || !s.fs_rank3 || !s.sparse_star)
    return fail(EDGE_ERR_UNSUPPORTED, "MFMA viscoelastic kernels need the factored flux solvers");
  switch (s.order) {
    case 3: return launch_rm<3>(s, first, n, dt, dofs, dofs_a, t_int, tin, local, st);
    case 4: return launch_rm<4>(s, first, n, dt, dofs, dofs_a, t_int, tin, local, st);
    case 5: return launch_rm<5>(s, first, n, dt, dofs, dofs_a, t_int, tin, local, st);
  }
  return fail(EDGE_ERR_UNSUPPORTED, "MFMA viscoelastic kernels: orders 3 to 5");
}

// The fused launch (k_local_vm with NB) over all elements.
template <int O, int NRM>
void launch_fused_o(const SolverDev<float>& s, float dt, const float* proj_in, float* dofs, float* dofs_a,
                    float* t_int, hipStream_t st) {
  hipLaunchKernelGGL((k_local_vm<O, NRM, false, true>), dim3(grid_of(s.n_el)), dim3(kLocalThreads), 0, st, s,
                     (int64_t)0, s.n_el, dt, dofs, dofs_a, t_int, proj_in);
}

template <int O>
int launch_fused_rm(const SolverDev<float>& s, float dt, const float* proj_in, float* dofs, float* dofs_a,
                    float* t_int, hipStream_t st) {
  switch (s.n_rms) {
    case 1: launch_fused_o<O, 1>(s, dt, proj_in, dofs, dofs_a, t_int, st); break;
    case 2: launch_fused_o<O, 2>(s, dt, proj_in, dofs, dofs_a, t_int, st); break;
    case 3: launch_fused_o<O, 3>(s, dt, proj_in, dofs, dofs_a, t_int, st); break;
    default: return fail(EDGE_ERR_UNSUPPORTED, "MFMA viscoelastic kernels: 1 to 3 mechanisms");
  }
  EDGE_HIP_TRY(hipGetLastError());
  return EDGE_OK;
}

}  // namespace

int launch_fused_visco_mfma(const SolverDev<float>& s, float dt, const float* proj_in, float* dofs, float* dofs_a,
                            float* t_int, hipStream_t st) {
  if (!s.fsf || !s.fsfa || !s.fs_rank3 || !s.sparse_star || !s.proj || !proj_in || !s.rcv)
    return fail(EDGE_ERR_UNSUPPORTED, "fused viscoelastic step: factored flux solvers and face projections");
  if (s.n_runs != R) return fail(EDGE_ERR_UNSUPPORTED, "fused viscoelastic step: 8 runs");
  if (s.n_el <= 0) return EDGE_OK;
  switch (s.order) {
    case 3: return launch_fused_rm<3>(s, dt, proj_in, dofs, dofs_a, t_int, st);
    case 4: return launch_fused_rm<4>(s, dt, proj_in, dofs, dofs_a, t_int, st);
    case 5: return launch_fused_rm<5>(s, dt, proj_in, dofs, dofs_a, t_int, st);
  }
  return fail(EDGE_ERR_UNSUPPORTED, "MFMA viscoelastic kernels: orders 3 to 5");
}

bool visco_mfma_supported(int order, int n_runs, int precision, int n_rms) {
  return precision == 32 && n_runs == R && order >= 3 && order <= 5 && n_rms >= 1 && n_rms <= 3;
}

int launch_local_visco_mfma(const SolverDev<float>& s, int64_t first, int64_t n, float dt, float* dofs,
                            float* dofs_a, float* t_int, hipStream_t st) {
  return launch_vm(s, first, n, dt, dofs, dofs_a, t_int, nullptr, true, st);
}

int launch_neigh_visco_mfma(const SolverDev<float>& s, int64_t first, int64_t n, const float* t_int, float* dofs,
                            float* dofs_a, hipStream_t st) {
  return launch_vm(s, first, n, 0.f, dofs, dofs_a, nullptr, t_int, false, st);
}

}  // namespace edge
