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
// One wave handles one element: the 9 quantities of an (n, r) position then
// live in the same lane, so the per-element 9 x 9 star and flux-solver
// products are lane-local VALU FMAs with wave-uniform coefficients (scalar
// loads).  The star matrices are applied to the B operands before the
// stiffness products (9 mixed values per k-step instead of whole tiles);
// with the elastic pattern only its 24 non-zeros are touched.
//
// Occupancy: two waves per SIMD (<= 256 registers, <= 20 KiB LDS per wave)
// so that one wave's VALU work runs while the other's MFMAs do.
//
// Operations (src/impl/seismic/solvers/AderDg.hpp:362-533, :553-722):
//   local: CK time prediction (TimePredVanilla.hpp:200-306), volume
//          integral (VolIntVanilla.hpp:189-241), local surface integral
//          (SurfIntVanilla.hpp:212-262);
//   neigh: neighbouring surface integral (SurfIntVanilla.hpp:274-357).
// Summation order differs from the reference's loops (tile order, the star
// applied before the stiffness product); results agree to rounding.
#include <hip/hip_runtime.h>

#include "dg_masks.hpp"
#include "kernels_mfma.hpp"
#include "mfma_pack.hpp"

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
  static constexpr int NVEC = SZ / 4;             // 16-byte vectors per element tile
  static constexpr int CHUNKS = (NVEC + 63) / 64;  // KiB-sized async copies per tile
};

// blockIdx -> element offset: each of the 8 XCDs (blocks dispatched round
// robin) gets a contiguous element range, so neighbours share its L2 (the
// autotuner may switch this schedule off: block i -> element i).
__device__ __forceinline__ int64_t xcd_element(int64_t n) {
  const int64_t chunk = (n + 7) / 8;
  const int64_t b = blockIdx.x;
  return (b & 7) * chunk + (b >> 3);
}

__device__ __forceinline__ v4f zero4() { return v4f{0.f, 0.f, 0.f, 0.f}; }

// Per-element tables that no kernel writes (star matrices, factored flux
// solvers, face ids) are read through the constant address space, so that
// wave-uniform values come in by scalar loads into SGPRs instead of vector
// loads into VGPRs.
typedef const __attribute__((address_space(4))) float cfloat;
typedef const __attribute__((address_space(4))) int32_t cint32;
__device__ __forceinline__ const cfloat* as_const(const float* p) { return (const cfloat*)p; }
typedef const __attribute__((address_space(4))) uint8_t cuint8;
__device__ __forceinline__ const cint32* as_const(const int32_t* p) { return (const cint32*)p; }
__device__ __forceinline__ const cuint8* as_const(const uint8_t* p) { return (const cuint8*)p; }

// The value x, opaque to the optimiser: index arithmetic derived from it is
// computed where it is used instead of being hoisted (and spilled) early.
__device__ __forceinline__ int opaque(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

// Guarded load without a branch: the index is clamped into the valid range
// (so the load is always in bounds) and the value masked afterwards.
__device__ __forceinline__ float ldg(const float* __restrict__ p, int idx, int clamp_idx, bool ok) {
  const float v = p[ok ? idx : clamp_idx];
  return ok ? v : 0.f;
}

typedef __attribute__((address_space(3))) void* lds_ptr;

// Asynchronous copy of an element tile (n_vec 16-byte vectors, contiguous)
// from global memory into LDS (global_load_lds_dwordx4: 1 KiB per wave and
// instruction).  The LDS buffer holds CHUNKS KiB; lanes past the tile re-read
// the tile's first vector into the buffer's padding, so every access stays in
// bounds.  Completion is awaited with vmcnt.
template <int CHUNKS>
__device__ __forceinline__ void tile_to_lds(const float* __restrict__ src, float* dst, int n_vec, int lane) {
#pragma unroll 2
  for (int i = 0; i < CHUNKS; i++) {
    const int idx = i * 64 + lane;
    const float* gp = src + 4 * (idx < n_vec ? idx : 0);
    __builtin_amdgcn_global_load_lds((const void*)gp, (lds_ptr)(dst + i * 256), 16, 0, 0);
  }
}

__device__ __forceinline__ void wait_vm_all() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

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

// LDS index of row k, column c of quantity q in an element tile, or of a
// dummy slot past the tile (kLdsPad floats) when k is not below `rows`:
// branch-free guarded LDS accesses.
constexpr int kLdsPad = 16;
template <int O>
__device__ __forceinline__ int lds_idx(int q, int k, int rows, int c) {
  return k < rows ? (q * Dims<O>::N + k) * R + c : Dims<O>::SZ + c;
}

// Coefficients of one star matrix kept in (scalar) registers across a k-loop:
// the 24 elastic non-zeros in row order, or all 81 entries.
template <bool SP>
struct StarCoef {
  static constexpr int NC = SP ? 24 : 81;
  float v[NC];
  __device__ __forceinline__ void load(const cfloat* __restrict__ sd) {
    if constexpr (SP) {
      constexpr int cols[24] = {6, 7, 8, 6, 7, 8, 6, 7, 8, 6, 7, 7, 8, 6, 8, 0, 3, 5, 1, 3, 4, 2, 4, 5};
      constexpr int rows[24] = {0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 6, 7, 7, 7, 8, 8, 8};
#pragma unroll
      for (int i = 0; i < 24; i++) v[i] = sd[rows[i] * 9 + cols[i]];
    } else {
#pragma unroll
      for (int i = 0; i < 81; i++) v[i] = sd[i];
    }
  }
  // z = sum_p S[q][p] b[p]; q is a compile-time constant after unrolling
  __device__ __forceinline__ float row(int q, const float (&b)[9]) const {
    if constexpr (SP) {
      constexpr int cols[24] = {6, 7, 8, 6, 7, 8, 6, 7, 8, 6, 7, 7, 8, 6, 8, 0, 3, 5, 1, 3, 4, 2, 4, 5};
      constexpr int first[10] = {0, 3, 6, 9, 11, 13, 15, 18, 21, 24};
      float z = 0.f;
#pragma unroll
      for (int i = first[q]; i < first[q + 1]; i++) z += v[i] * b[cols[i]];
      return z;
    } else {
      float z = 0.f;
#pragma unroll
      for (int p = 0; p < 9; p++) z += v[q * 9 + p] * b[p];
      return z;
    }
  }
};

// ---------------------------------------------------------------- local
// CK stage 1 together with tInt's first term, any star matrices.  The LDS
// tile holds Q on entry; B values are read in row order (k = 4 s + g) once
// per direction, and the last direction replaces each row by dt Q;
// der_1[q] = sum_d (-K_d^T M^-1)^T (sum_p S_d[q][p] Q_p).
template <int O, bool SP>
__device__ __forceinline__ void ck_first(const float* __restrict__ stiffTn, const cfloat* __restrict__ star,
                                         float dt, float* l_ti, v4f (&der)[9][tiles(n_modes(O - 1))], int g,
                                         int c) {
  constexpr int N = Dims<O>::N;
  constexpr int NK = N, NN = n_modes(O - 1), TN = tiles(NN), KS = ksteps(NK);
#pragma unroll
  for (int q = 0; q < 9; q++)
#pragma unroll
    for (int t = 0; t < TN; t++) der[q][t] = zero4();
  // operands of step (d, s) are loaded one step ahead of their use
  auto load = [&](int d, int s, float (&a)[TN], float (&b)[9]) {
    const int k = 4 * s + g;
#pragma unroll
    for (int t = 0; t < TN; t++) a[t] = ldg(stiffTn + d * N * N, k * N + 16 * t + c, 0, k < NK && 16 * t + c < NN);
#pragma unroll
    for (int p = 0; p < 9; p++) {
      if (4 * s + 3 < NK)  // constant after unrolling
        b[p] = l_ti[(p * N + k) * R + c];
      else
        b[p] = ldg(l_ti, (p * N + k) * R + c, c, k < NK);
    }
  };
  float an[TN], bn[9];
  load(0, 0, an, bn);
#pragma unroll
  for (int d = 0; d < 3; d++) {
    StarCoef<SP> sc;
    sc.load(star + d * 81);
#pragma unroll
    for (int s = 0; s < KS; s++) {
      float a[TN], b[9];
#pragma unroll
      for (int t = 0; t < TN; t++) a[t] = an[t];
#pragma unroll
      for (int p = 0; p < 9; p++) b[p] = bn[p];
      if (s + 1 < KS)
        load(d, s + 1, an, bn);
      else if (d + 1 < 3)
        load(d + 1, 0, an, bn);
      __builtin_amdgcn_sched_barrier(0);  // bounds register pressure: no hoisting across k-steps
      if (d == 2) {
        const int k = 4 * s + g;
#pragma unroll
        for (int p = 0; p < 9; p++) l_ti[lds_idx<O>(p, k, NK, c)] = dt * b[p];
      }
#pragma unroll
      for (int q = 0; q < 9; q++) {
        const float z = sc.row(q, b);
#pragma unroll
        for (int t = 0; t < TN; t++) der[q][t] = mfma(a[t], z, der[q][t]);
      }
    }
  }
}

// CK stage DE >= 2 from the previous derivative held in registers
// (permuted k: row 16 u + 4 g + v is register v of tile u).
template <int O, int DE, bool SP>
__device__ __forceinline__ void ck_next(const float* __restrict__ stiffTn, const cfloat* __restrict__ star,
                                        const v4f (&prev)[9][tiles(n_modes(O - DE + 1))],
                                        v4f (&cur)[9][tiles(n_modes(O - DE))], int g, int c) {
  constexpr int N = Dims<O>::N;
  constexpr int NK = n_modes(O - DE + 1), NN = n_modes(O - DE), TP = tiles(NK), TN = tiles(NN);
  // k-steps (d, u, v) with any row below NK, in order
  constexpr int NS = 3 * ((TP - 1) * 4 + (NK - 16 * (TP - 1) < 4 ? NK - 16 * (TP - 1) : 4));
#pragma unroll
  for (int q = 0; q < 9; q++)
#pragma unroll
    for (int t = 0; t < TN; t++) cur[q][t] = zero4();
  auto step_of = [](int i, int& d, int& u, int& v) {
    constexpr int per_d = NS / 3;
    d = i / per_d;
    const int j = i % per_d;
    u = j / 4;
    v = j % 4;
  };
  auto load = [&](int i, float (&a)[TN]) {
    int d, u, v;
    step_of(i, d, u, v);
    const float* m = stiffTn + d * N * N;
    const int k = 16 * u + 4 * g + v;
#pragma unroll
    for (int t = 0; t < TN; t++) a[t] = ldg(m, k * N + 16 * t + c, 0, k < NK && 16 * t + c < NN);
  };
  float an[TN];
  load(0, an);
  StarCoef<SP> sc;
#pragma unroll
  for (int i = 0; i < NS; i++) {
    int d, u, v;
    step_of(i, d, u, v);
    if (i % (NS / 3) == 0) sc.load(star + d * 81);  // constant after unrolling
    float a[TN];
#pragma unroll
    for (int t = 0; t < TN; t++) a[t] = an[t];
    if (i + 1 < NS) load(i + 1, an);
    __builtin_amdgcn_sched_barrier(0);
    float b[9];
#pragma unroll
    for (int p = 0; p < 9; p++) b[p] = prev[p][u][v];
#pragma unroll
    for (int q = 0; q < 9; q++) {
      const float z = sc.row(q, b);
#pragma unroll
      for (int t = 0; t < TN; t++) cur[q][t] = mfma(a[t], z, cur[q][t]);
    }
  }
}

// Elastic star pattern: stress rows (0..5) have non-zeros only in the
// velocity columns 6..8 and velocity rows (6..8) only in stress columns.  A
// stress output sum_p S_d[q][p] X_p is then a combination of the three raw
// velocity tiles, so its stiffness product is formed from the three products
// K_d X_{6+i} (per direction) and combined afterwards, lane-locally: 6
// instead of 9 products per direction and k-step.
constexpr int kVelCols[3][3] = {{0, 3, 5}, {1, 3, 4}, {2, 4, 5}};  // rows 6, 7, 8
constexpr int kStrNnz[6] = {3, 3, 3, 2, 2, 2};
constexpr int kStrVel[6][3] = {{0, 1, 2}, {0, 1, 2}, {0, 1, 2}, {0, 1, 0}, {1, 2, 0}, {0, 2, 0}};  // column 6 + i

struct VelRows {  // S_d[6 + r][kVelCols[r][j]]
  float v[9];
  __device__ __forceinline__ void load(const cfloat* __restrict__ sd) {
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int j = 0; j < 3; j++) v[r * 3 + j] = sd[(6 + r) * 9 + kVelCols[r][j]];
  }
  __device__ __forceinline__ float row(int r, const float (&b)[9]) const {
    return v[r * 3] * b[kVelCols[r][0]] + v[r * 3 + 1] * b[kVelCols[r][1]] + v[r * 3 + 2] * b[kVelCols[r][2]];
  }
};

// out_q (+)= sum_i S_d[q][6 + i] P_i for the stress rows q = 0..5 (D tiles)
template <int TN>
__device__ __forceinline__ void add_stress(const cfloat* __restrict__ sd, const v4f (&p)[3][TN], v4f (&out)[9][TN],
                                           bool first) {
#pragma unroll
  for (int q = 0; q < 6; q++)
#pragma unroll
    for (int t = 0; t < TN; t++) {
      v4f z = first ? zero4() : out[q][t];
#pragma unroll
      for (int j = 0; j < kStrNnz[q]; j++) z += sd[q * 9 + 6 + kStrVel[q][j]] * p[kStrVel[q][j]][t];
      out[q][t] = z;
    }
}

// CK stage 1 for the elastic pattern (see ck_first): tInt's first term, the
// velocity outputs from star-combined B values, the stress outputs from the
// raw velocity products P[d][i] combined after the k-loop.  A operands from
// the packed table (four k-steps per load, one group ahead).
template <int O>
__device__ __forceinline__ void ck_first_el(const float* __restrict__ pk, const cfloat* __restrict__ star, float dt,
                                            float* l_ti, v4f (&der)[9][tiles(n_modes(O - 1))], int g, int c,
                                            int lane) {
  constexpr int N = Dims<O>::N;
  constexpr int NK = N, NN = n_modes(O - 1), TN = tiles(NN), KS = ksteps(NK), NG = pk::groups(KS);
  constexpr pk::Layout L{O};
  const v4f* pa = reinterpret_cast<const v4f*>(pk + L.ck1());  // [d][t][group][lane]
  v4f pv[3][3][TN];
#pragma unroll
  for (int d = 0; d < 3; d++)
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int t = 0; t < TN; t++) pv[d][i][t] = zero4();
#pragma unroll
  for (int r = 6; r < 9; r++)
#pragma unroll
    for (int t = 0; t < TN; t++) der[r][t] = zero4();
  VelRows sv[3];
#pragma unroll
  for (int d = 0; d < 3; d++) sv[d].load(star + d * 81);
  auto load_a = [&](int gr, v4f (&a)[3][TN]) {
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
      for (int t = 0; t < TN; t++) a[d][t] = pa[((d * TN + t) * NG + gr) * 64 + lane];
  };
  auto load_b = [&](int s, float (&b)[9]) {
    const int k = 4 * s + g;
#pragma unroll
    for (int p = 0; p < 9; p++) {
      if (4 * s + 3 < NK)  // constant after unrolling
        b[p] = l_ti[(p * N + k) * R + c];
      else
        b[p] = ldg(l_ti, (p * N + k) * R + c, c, k < NK);
    }
  };
  v4f a4[3][TN], a4n[3][TN];
  float bn[9];
  load_a(0, a4);
  load_b(0, bn);
#pragma unroll
  for (int s = 0; s < KS; s++) {
    if (s % 4 == 0 && s / 4 + 1 < NG) load_a(s / 4 + 1, a4n);
    float b[9];
#pragma unroll
    for (int p = 0; p < 9; p++) b[p] = bn[p];
    if (s + 1 < KS) load_b(s + 1, bn);
    __builtin_amdgcn_sched_barrier(0);  // bounds register pressure: no hoisting across k-steps
    const int k = 4 * s + g;
#pragma unroll
    for (int p = 0; p < 9; p++) l_ti[lds_idx<O>(p, k, NK, c)] = dt * b[p];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      // A blocks that are zero (the stiffness matrices' sparsity) are skipped
      constexpr masks::Tab NZ = masks::Nz<O, true>::tab;
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int t = 0; t < TN; t++)
          if ((NZ.ck[1][d][t] >> s) & 1) pv[d][i][t] = mfma(a4[d][t][s % 4], b[6 + i], pv[d][i][t]);
#pragma unroll
      for (int r = 0; r < 3; r++) {
        const float z = sv[d].row(r, b);
#pragma unroll
        for (int t = 0; t < TN; t++)
          if ((NZ.ck[1][d][t] >> s) & 1) der[6 + r][t] = mfma(a4[d][t][s % 4], z, der[6 + r][t]);
      }
    }
    if (s % 4 == 3) {
#pragma unroll
      for (int d = 0; d < 3; d++)
#pragma unroll
        for (int t = 0; t < TN; t++) a4[d][t] = a4n[d][t];
    }
  }
#pragma unroll
  for (int d = 0; d < 3; d++) add_stress<TN>(star + d * 81, pv[d], der, d == 0);
}

// CK stage DE >= 2 for the elastic pattern (see ck_next), direction-major
// k-steps (u, v); A operands from the packed table, one 4-step group ahead.
template <int O, int DE>
__device__ __forceinline__ void ck_next_el(const float* __restrict__ pk, const cfloat* __restrict__ star,
                                           const v4f (&prev)[9][tiles(n_modes(O - DE + 1))],
                                           v4f (&cur)[9][tiles(n_modes(O - DE))], int lane) {
  constexpr int NN = n_modes(O - DE), TN = tiles(NN);
  static_assert(TN == 1, "derivatives of stage 2 and up fit one tile");
  constexpr int PER_D = pk::next_steps(O, DE), NG = pk::groups(PER_D);
  constexpr pk::Layout L{O};
  const v4f* pa = reinterpret_cast<const v4f*>(pk + L.ck(DE));  // [d][group][lane]
#pragma unroll
  for (int r = 6; r < 9; r++) cur[r][0] = zero4();
  v4f a4 = pa[lane], a4n = a4;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    VelRows sv;
    sv.load(star + d * 81);
    v4f pv[3][1];
#pragma unroll
    for (int i = 0; i < 3; i++) pv[i][0] = zero4();
#pragma unroll
    for (int j = 0; j < PER_D; j++) {
      const int u = j / 4, v = j % 4;
      if (v == 0) {  // next group: (d, u + 1) or (d + 1, 0)
        if (u + 1 < NG)
          a4n = pa[(d * NG + u + 1) * 64 + lane];
        else if (d + 1 < 3)
          a4n = pa[((d + 1) * NG) * 64 + lane];
      }
      __builtin_amdgcn_sched_barrier(0);
      const float a = a4[v];
      if ((masks::Nz<O, true>::tab.ck[DE][d][0] >> j) & 1) {  // a non-zero A block
        float b[9];
#pragma unroll
        for (int p = 0; p < 9; p++) b[p] = prev[p][u][v];
#pragma unroll
        for (int q = 0; q < 3; q++) pv[q][0] = mfma(a, b[6 + q], pv[q][0]);
#pragma unroll
        for (int r = 0; r < 3; r++) cur[6 + r][0] = mfma(a, sv.row(r, b), cur[6 + r][0]);
      }
      if (v == 3 || j == PER_D - 1) a4 = a4n;
    }
    add_stress<1>(star + d * 81, pv, cur, d == 0);
  }
}

// tInt (LDS) += x for tile t of a D-layout array (rows below `rows`).  All
// reads are issued before any write: rows past `rows` share the dummy slot,
// so the compiler cannot prove the 36 read-modify-writes independent and
// would otherwise wait out the LDS latency of each one in turn.
template <int O>
__device__ __forceinline__ void add_tint_tile(float* __restrict__ l_ti, int rows, int t, const v4f (&x)[9], int g,
                                              int c) {
  float o[9][4];
#pragma unroll
  for (int q = 0; q < 9; q++)
#pragma unroll
    for (int v = 0; v < 4; v++) o[q][v] = l_ti[lds_idx<O>(q, 16 * t + 4 * g + v, rows, c)];
#pragma unroll
  for (int q = 0; q < 9; q++)
#pragma unroll
    for (int v = 0; v < 4; v++) l_ti[lds_idx<O>(q, 16 * t + 4 * g + v, rows, c)] = o[q][v] + x[q][v];
}

// CK stages DE >= 2.  Their derivatives have at most 16 rows, so their tInt
// terms sum into t0 (rows 0..15, registers) that is added to LDS at the end.
template <int O, int DE, bool SP>
__device__ __forceinline__ void ck_rest(const DgDev<float>& dg, const cfloat* star,
                                        const v4f (&prev)[9][tiles(n_modes(O - DE + 1))], v4f (&t0)[9], float dt,
                                        float sca, int g, int c, int lane) {
  if constexpr (DE < O) {
    constexpr int NN = n_modes(O - DE);
    static_assert(tiles(NN) == 1, "derivatives of stage 2 and up fit one tile");
    v4f cur[9][1];
    if constexpr (SP)
      ck_next_el<O, DE>(dg.pk, star, prev, cur, lane);
    else
      ck_next<O, DE, SP>(dg.stiffTn, star, prev, cur, g, c);
    sca *= dt / float(DE + 1);
#pragma unroll
    for (int q = 0; q < 9; q++) t0[q] += sca * cur[q][0];
    ck_rest<O, DE + 1, SP>(dg, star, cur, t0, dt, sca, g, c, lane);
  }
}

// acc_q += F^T_fa^T (fs (F_fa^T X_q)) for one face: the projection with F
// (fluxL or fluxN, [N][NF]) of the tile X in LDS ([q][N][R]), the 9 x 9 flux
// solver fs (wave-uniform), and the lift with F^T ([NF][N]) from registers.
template <int O>
__device__ __forceinline__ void face_update(const float* __restrict__ fl, const float* __restrict__ ft,
                                            const float* __restrict__ f, const float* __restrict__ x,
                                            v4f (&acc)[9][Dims<O>::NT], int g, int c) {
  constexpr int N = Dims<O>::N, NF = Dims<O>::NF, NT = Dims<O>::NT;
  constexpr int KS = ksteps(N);
  float at[4][NT];
#pragma unroll
  for (int v = 0; v < 4; v++)
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const int fr = 4 * g + v;  // permuted k: face mode held in register v
      at[v][t] = ldg(ft, fr * N + 16 * t + c, 0, fr < NF && 16 * t + c < N);
    }
  auto load = [&](int ks, float& al, float (&b)[9]) {
    const int k = 4 * ks + g;
    al = ldg(fl, k * NF + c, 0, k < N && c < NF);
#pragma unroll
    for (int q = 0; q < 9; q++) {
      if (4 * ks + 3 < N)  // constant after unrolling
        b[q] = x[(q * N + k) * R + c];
      else
        b[q] = ldg(x, (q * N + k) * R + c, c, k < N);
    }
  };
  v4f s0[9];
#pragma unroll
  for (int q = 0; q < 9; q++) s0[q] = zero4();
  float aln, bn[9];
  load(0, aln, bn);
#pragma unroll
  for (int ks = 0; ks < KS; ks++) {
    const float al = aln;
    float b[9];
#pragma unroll
    for (int q = 0; q < 9; q++) b[q] = bn[q];
    if (ks + 1 < KS) load(ks + 1, aln, bn);
    __builtin_amdgcn_sched_barrier(0);  // bounds register pressure: no hoisting across k-steps
#pragma unroll
    for (int q = 0; q < 9; q++) s0[q] = mfma(al, b[q], s0[q]);
  }
  // per register column v (face modes 4 g + v): s1 = fs s0, then the lift
#pragma unroll
  for (int v = 0; v < 4; v++) {
    __builtin_amdgcn_sched_barrier(0);
    float s1[9];
#pragma unroll
    for (int q = 0; q < 9; q++) {
      float z = 0.f;
#pragma unroll
      for (int p = 0; p < 9; p++) z += f[q * 9 + p] * s0[p][v];
      s1[q] = z;
    }
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
      for (int q = 0; q < 9; q++) acc[q][t] = mfma(at[v][t], s1[q], acc[q][t]);
  }
}

// Rank-3 face routines.  A flux solver factored as U V^T (U [9][3], V^T
// [3][9], wave-uniform) turns the 9-quantity projection / solver / lift chain
// into three combined columns: y_j = sum_p V^T[j][p] X_p on the B operands,
// w_j = F^T y_j (3 products instead of 9), lifted tiles L_j = F^T_fa^T w_j
// (3 instead of 9), and acc_q += sum_j U[q][j] L_j lane-locally.

// w_j (D layout: face-mode rows 4 g + v, run c) = F^T (V^T X)_j for the tile
// X in LDS ([q][N][R]) and F (fluxL or fluxN, [N][NF]) as packed A operands
// pm ([group][lane], four k-steps per vector).
// a_first: pm[lane], the first group, loaded by the caller ahead of time.
template <int O>
__device__ __forceinline__ void project3(const v4f* __restrict__ pm, v4f a_first, const cfloat* __restrict__ vt,
                                         const float* __restrict__ x, v4f (&w)[3], int g, int c, int lane) {
  constexpr int N = Dims<O>::N;
  constexpr int KS = ksteps(N), NG = pk::groups(KS);
  float cv[27];
#pragma unroll
  for (int i = 0; i < 27; i++) cv[i] = vt[i];
  auto load_b = [&](int ks, float (&b)[9]) {
    const int k = 4 * ks + g;
#pragma unroll
    for (int q = 0; q < 9; q++) {
      if (4 * ks + 3 < N)  // constant after unrolling
        b[q] = x[(q * N + k) * R + c];
      else
        b[q] = ldg(x, (q * N + k) * R + c, c, k < N);
    }
  };
#pragma unroll
  for (int j = 0; j < 3; j++) w[j] = zero4();
  v4f a4 = a_first, a4n = a4;
  float bn[9];
  load_b(0, bn);
#pragma unroll
  for (int ks = 0; ks < KS; ks++) {
    if (ks % 4 == 0 && ks / 4 + 1 < NG) a4n = pm[(ks / 4 + 1) * 64 + lane];
    float b[9];
#pragma unroll
    for (int q = 0; q < 9; q++) b[q] = bn[q];
    if (ks + 1 < KS) load_b(ks + 1, bn);
    __builtin_amdgcn_sched_barrier(0);
    const float al = a4[ks % 4];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      float y = 0.f;
#pragma unroll
      for (int p = 0; p < 9; p++) y += cv[j * 9 + p] * b[p];
      w[j] = mfma(al, y, w[j]);
    }
    if (ks % 4 == 3) a4 = a4n;
  }
}

// acc_q += sum_j U[q][j] (F^T_fa^T w_j) with the lift F^T_fa ([NF][N]) as
// packed A operands pl ([t][lane], the four register columns v per vector) on
// the three projected columns (B operands straight from the w tiles, k = 4 g + v).
template <int O>
__device__ __forceinline__ void load_lift(const v4f* __restrict__ pl, v4f (&at)[Dims<O>::NT], int lane) {
#pragma unroll
  for (int t = 0; t < Dims<O>::NT; t++) at[t] = pl[t * 64 + lane];
}

// at: the face's lift operands (load_lift), loaded by the caller ahead of time.
template <int O>
__device__ __forceinline__ void lift3(const v4f (&at)[Dims<O>::NT], const cfloat* __restrict__ u, const v4f (&w)[3],
                                      v4f (&acc)[9][Dims<O>::NT]) {
  constexpr int NT = Dims<O>::NT;
  v4f l[3][NT];
#pragma unroll
  for (int j = 0; j < 3; j++)
#pragma unroll
    for (int t = 0; t < NT; t++) l[j][t] = zero4();
#pragma unroll
  for (int v = 0; v < 4; v++)
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int t = 0; t < NT; t++) l[j][t] = mfma(at[t][v], w[j][v], l[j][t]);
  float cu[27];
#pragma unroll
  for (int i = 0; i < 27; i++) cu[i] = u[i];
#pragma unroll
  for (int q = 0; q < 9; q++)
#pragma unroll
    for (int t = 0; t < NT; t++) acc[q][t] += cu[q * 3] * l[0][t] + cu[q * 3 + 1] * l[1][t] + cu[q * 3 + 2] * l[2][t];
}

// Face projections exchanged between the local and the neighbour step: per
// element face 3 tiles of 16 x 16 (w_j, face modes x runs), lane-major
// (lane l holds the 4 values of its D-layout registers at [j][l][0..3]).
constexpr int kProjFace = 3 * 256;

__device__ __forceinline__ void store_proj(float* __restrict__ p, const v4f (&w)[3], int lane) {
#pragma unroll
  for (int j = 0; j < 3; j++) reinterpret_cast<v4f*>(p + j * 256)[lane] = w[j];
}

__device__ __forceinline__ void load_proj(const float* __restrict__ p, v4f (&w)[3], int lane) {
#pragma unroll
  for (int j = 0; j < 3; j++) w[j] = reinterpret_cast<const v4f*>(p + j * 256)[lane];
}

// The neighbouring surface integral of one element from the face
// projections its neighbours wrote (proj, [slot][3][256] lane-major):
// acc_q += sum_fa U_fa (F^T_fa^T w_fa), acc in the D layout.  At order 5
// every face's projection is requested up front and each face's lift
// operands one face ahead; below, the registers this takes would halve the
// occupancy.
// init(acc) fills the accumulators (the element's DOFs) once the projections
// are requested.
template <int O, typename Init>
__device__ __forceinline__ void neigh_lifts(const SolverDev<float>& s, const float* __restrict__ proj, int64_t el,
                                            v4f (&acc)[9][Dims<O>::NT], int lane, Init init) {
  constexpr int NT = Dims<O>::NT;
  constexpr bool AHEAD = O >= 5;
  const v4f* plft = reinterpret_cast<const v4f*>(s.dg.pk + pk::Layout{O}.lft());
  int32_t rcv[4];
  v4f w[AHEAD ? 4 : 1][3];
#pragma unroll
  for (int fa = 0; fa < 4; fa++) {
    rcv[fa] = as_const(s.rcv)[el * 4 + fa];
    if (AHEAD && rcv[fa] >= 0) load_proj(proj + (int64_t)rcv[fa] * kProjFace, w[AHEAD ? fa : 0], lane);
  }
  v4f at[NT];
  load_lift<O>(plft, at, lane);
  init(acc);
#pragma unroll
  for (int fa = 0; fa < 4; fa++) {
    v4f atn[NT];
    if (AHEAD && fa < 3) load_lift<O>(plft + (fa + 1) * NT * 64, atn, lane);
    if (!AHEAD && fa > 0) load_lift<O>(plft + fa * NT * 64, at, lane);
    if (rcv[fa] >= 0) {  // else outflow: no projection (uniform across the wave)
      if (!AHEAD) load_proj(proj + (int64_t)rcv[fa] * kProjFace, w[0], lane);
      lift3<O>(at, as_const(s.fsf + ((s.n_el + el) * 4 + fa) * 54), w[AHEAD ? fa : 0], acc);
    }
    if (AHEAD && fa < 3) {
#pragma unroll
      for (int t = 0; t < NT; t++) at[t] = atn[t];
    }
  }
}

// WP: also write the face projections for the neighbour step (s.proj,
// rank-3 solvers only).
// NB: first the previous time step's neighbouring surface integral from the
// projections in proj_in (written by the previous launch; s.proj receives
// this step's, so the two alternate between two buffers).  The updated DOFs
// go to LDS for the CK and back to global memory, from where the output
// stage reads them as it reads Q otherwise (and to s.tap for the elements
// s.tap_on marks).  One launch then carries the neighbour step of step
// k - 1 and the local step of step k.
template <int O, bool SP, bool R3, bool WP = false, bool NB = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
void k_local(SolverDev<float> s, int64_t first, int64_t n, float dt, float* __restrict__ dofs,
             float* __restrict__ t_int, const float* __restrict__ proj_in) {
  static_assert(!NB || (R3 && WP), "the fused step exchanges face projections");
  using D = Dims<O>;
  constexpr int N = D::N, NF = D::NF, NT = D::NT;
  constexpr int NV = n_modes(O - 1);  // rows of tInt the volume integral reads
  // Q, then the time-integrated DOFs, then the output
  __shared__ __align__(16) float l_ti[D::SZ + kLdsPad > D::CHUNKS * 256 ? D::SZ + kLdsPad : D::CHUNKS * 256];
  const int64_t off = s.xcd_map ? xcd_element(n) : (int64_t)blockIdx.x;
  if (off >= n) return;
  const int64_t el = first + off;
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  float* q0 = dofs + el * D::SZ;
  const cfloat* star = as_const(s.star + el * 243);
  if constexpr (NB) {
    {
      v4f nq[9][NT];
      neigh_lifts<O>(s, proj_in, el, nq, lane, [&](v4f (&a)[9][NT]) {
#pragma unroll
        for (int q = 0; q < 9; q++)
#pragma unroll
          for (int t = 0; t < NT; t++) a[q][t] = load_d(q0 + q * N * R, N, t, g, c);
      });
      const int ln = opaque(lane), go = ln >> 4, co = ln & 15;
#pragma unroll
      for (int q = 0; q < 9; q++)
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
          for (int v = 0; v < 4; v++) l_ti[lds_idx<O>(q, 16 * t + 4 * go + v, N, co)] = nq[q][t][v];
    }
    __syncthreads();
    float4* gq = reinterpret_cast<float4*>(q0);
    const float4* lq = reinterpret_cast<const float4*>(l_ti);
    for (int i = lane; i < D::NVEC; i += 64) gq[i] = lq[i];
    if (s.tap && as_const(s.tap_on)[el]) {  // uniform across the wave
      float4* tq = reinterpret_cast<float4*>(s.tap + el * D::SZ);
      for (int i = lane; i < D::NVEC; i += 64) tq[i] = lq[i];
    }
  } else {
    tile_to_lds<D::CHUNKS>(q0, l_ti, D::NVEC, lane);
  }

  // ---- CK: tInt = dt Q + sum_de dt^{de+1}/(de+1)! der_de, tile in LDS
  {
    v4f der[9][tiles(NV)];
    wait_vm_all();
    if constexpr (SP)
      ck_first_el<O>(s.dg.pk, star, dt, l_ti, der, g, c, lane);
    else
      ck_first<O, SP>(s.dg.stiffTn, star, dt, l_ti, der, g, c);
    const float sca = dt * dt / 2.f;
    v4f t0[9];
#pragma unroll
    for (int q = 0; q < 9; q++) t0[q] = sca * der[q][0];
#pragma unroll
    for (int t = 1; t < tiles(NV); t++) {
      v4f x[9];
#pragma unroll
      for (int q = 0; q < 9; q++) x[q] = sca * der[q][t];
      add_tint_tile<O>(l_ti, NV, t, x, g, c);
    }
    ck_rest<O, 2, SP>(s.dg, star, der, t0, dt, sca, g, c, lane);
    add_tint_tile<O>(l_ti, NV, 0, t0, g, c);
  }
  __syncthreads();

  // ---- volume accumulators start from Q (from zero with rank-3 faces, where
  // Q comes back through LDS at the end); tInt goes to global memory meanwhile
  v4f acc[9][NT];
#pragma unroll
  for (int q = 0; q < 9; q++)
#pragma unroll
    for (int t = 0; t < NT; t++) acc[q][t] = R3 ? zero4() : load_d(q0 + q * N * R, N, t, g, c);
  {
    float4* gti = reinterpret_cast<float4*>(t_int + el * D::SZ);
    const float4* lti = reinterpret_cast<const float4*>(l_ti);
    for (int i = lane; i < D::NVEC; i += 64) gti[i] = lti[i];
  }

  // ---- volume: acc_q += sum_d stiffV_d^T (sum_p S_d[q][p] tInt_p)
  constexpr int KV = ksteps(NV);
  if constexpr (SP) {
    // elastic pattern: per direction the velocity outputs from star-combined
    // B values, the stress outputs from the three raw velocity products;
    // A operands from the packed table, one 4-step group ahead
    constexpr int NG = pk::groups(KV);
    constexpr pk::Layout L{O};
    const v4f* pa = reinterpret_cast<const v4f*>(s.dg.pk + L.vol());  // [d][t][group][lane]
    auto load_a = [&](int d, int gr, v4f (&a)[NT]) {
#pragma unroll
      for (int t = 0; t < NT; t++) a[t] = pa[((d * NT + t) * NG + gr) * 64 + lane];
    };
    auto load_b = [&](int ks, float (&b)[9]) {
      const int k = 4 * ks + g;
#pragma unroll
      for (int p = 0; p < 9; p++) {
        if (4 * ks + 3 < NV)  // constant after unrolling
          b[p] = l_ti[(p * N + k) * R + c];
        else
          b[p] = ldg(l_ti, (p * N + k) * R + c, c, k < NV);
      }
    };
    v4f a4[NT], a4n[NT];
    float bn[9];
    load_a(0, 0, a4);
    load_b(0, bn);
#pragma unroll
    for (int d = 0; d < 3; d++) {
      VelRows sv;
      sv.load(star + d * 81);
      v4f pv[3][NT];
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int t = 0; t < NT; t++) pv[i][t] = zero4();
#pragma unroll
      for (int ks = 0; ks < KV; ks++) {
        if (ks % 4 == 0) {  // next group: (d, ks / 4 + 1) or (d + 1, 0)
          if (ks / 4 + 1 < NG)
            load_a(d, ks / 4 + 1, a4n);
          else if (d + 1 < 3)
            load_a(d + 1, 0, a4n);
        }
        float b[9];
#pragma unroll
        for (int p = 0; p < 9; p++) b[p] = bn[p];
        if (ks + 1 < KV)
          load_b(ks + 1, bn);
        else if (d + 1 < 3)
          load_b(0, bn);
        __builtin_amdgcn_sched_barrier(0);
        constexpr masks::Tab NZ = masks::Nz<O, true>::tab;  // zero A blocks are skipped
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
          for (int t = 0; t < NT; t++)
            if ((NZ.vol[d][t] >> ks) & 1) pv[i][t] = mfma(a4[t][ks % 4], b[6 + i], pv[i][t]);
#pragma unroll
        for (int r = 0; r < 3; r++) {
          const float z = sv.row(r, b);
#pragma unroll
          for (int t = 0; t < NT; t++)
            if ((NZ.vol[d][t] >> ks) & 1) acc[6 + r][t] = mfma(a4[t][ks % 4], z, acc[6 + r][t]);
        }
        if (ks % 4 == 3 || ks == KV - 1) {
#pragma unroll
          for (int t = 0; t < NT; t++) a4[t] = a4n[t];
        }
      }
      add_stress<NT>(star + d * 81, pv, acc, false);
    }
  } else {
    auto load = [&](int d, int ks, float (&a)[NT], float (&b)[9]) {
      const int k = 4 * ks + g;
#pragma unroll
      for (int t = 0; t < NT; t++) a[t] = ldg(s.dg.stiffV + d * N * N, k * N + 16 * t + c, 0, k < NV && 16 * t + c < N);
#pragma unroll
      for (int p = 0; p < 9; p++) {
        if (4 * ks + 3 < NV)  // constant after unrolling
          b[p] = l_ti[(p * N + k) * R + c];
        else
          b[p] = ldg(l_ti, (p * N + k) * R + c, c, k < NV);
      }
    };
    float an[NT], bn[9];
    load(0, 0, an, bn);
#pragma unroll
    for (int d = 0; d < 3; d++) {
      StarCoef<SP> sc;
      sc.load(star + d * 81);
#pragma unroll
      for (int ks = 0; ks < KV; ks++) {
        float a[NT], b[9];
#pragma unroll
        for (int t = 0; t < NT; t++) a[t] = an[t];
#pragma unroll
        for (int p = 0; p < 9; p++) b[p] = bn[p];
        if (ks + 1 < KV)
          load(d, ks + 1, an, bn);
        else if (d + 1 < 3)
          load(d + 1, 0, an, bn);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int q = 0; q < 9; q++) {
          const float z = sc.row(q, b);
#pragma unroll
          for (int t = 0; t < NT; t++) acc[q][t] = mfma(a[t], z, acc[q][t]);
        }
      }
    }
  }

  // ---- local surface integral
  if constexpr (R3) {
    constexpr pk::Layout L{O};
    constexpr int NGP = pk::groups(ksteps(N));
    const v4f* pprj = reinterpret_cast<const v4f*>(s.dg.pk + L.prj());
    const v4f* plft = reinterpret_cast<const v4f*>(s.dg.pk + L.lft());
    // Every projection first (the last reads of tInt), then Q streams into
    // LDS while the lifts run.  The first A group of each projection is
    // loaded one projection ahead.
    // ---- this element's tInt as each face-adjacent element sees it (the
    // neighbour's flux matrix and solver side), for the neighbour step
    v4f a0 = pprj[(WP ? as_const(s.snd_fm)[el * 4] : 0) * NGP * 64 + lane];
    if constexpr (WP) {
#pragma unroll 1
      for (int f = 0; f < 4; f++) {
        const int32_t rv = as_const(s.snd_v)[el * 4 + f];
        const int fm = as_const(s.snd_fm)[el * 4 + f];
        const int fmn = f < 3 ? as_const(s.snd_fm)[el * 4 + f + 1] : 0;  // then local face 0
        const v4f a0n = pprj[fmn * NGP * 64 + lane];
        if (rv >= 0) {  // somebody receives through this face (uniform across the wave)
          v4f w[3];
          project3<O>(pprj + fm * NGP * 64, a0, as_const(s.fsf + (s.n_el * 4 + rv) * 54 + 27), l_ti, w, g, c, lane);
          store_proj(s.proj + (el * 4 + f) * kProjFace, w, lane);
        }
        a0 = a0n;
      }
    } else {
      a0 = pprj[lane];
    }
    // ---- the local faces' projections, kept for the lifts
    v4f wl[4][3];
#pragma unroll
    for (int fa = 0; fa < 4; fa++) {
      __builtin_amdgcn_sched_barrier(0);  // one face at a time (register pressure)
      const v4f a0n = pprj[(fa < 3 ? fa + 1 : fa) * NGP * 64 + lane];
      project3<O>(pprj + fa * NGP * 64, a0, as_const(s.fsf + (el * 4 + fa) * 54 + 27), l_ti, wl[fa], g, c, lane);
      a0 = a0n;
    }
    __syncthreads();  // tInt's last reads are done: Q may overwrite it
    tile_to_lds<D::CHUNKS>(q0, l_ti, D::NVEC, lane);
#pragma unroll
    for (int fa = 0; fa < 4; fa++) {
      __builtin_amdgcn_sched_barrier(0);
      v4f at[NT];
      load_lift<O>(plft + fa * NT * 64, at, lane);
      lift3<O>(at, as_const(s.fsf + (el * 4 + fa) * 54), wl[fa], acc);
    }
  } else {
    const float* fs0 = s.fs + el * 324;
#pragma unroll 1
    for (int fa = 0; fa < 4; fa++)
      face_update<O>(s.dg.fluxLN + fa * N * NF, s.dg.fluxT + fa * NF * N, fs0 + fa * 81, l_ti, acc, g, c);
  }

  // ---- DOFs out through LDS (16-byte global stores)
  if constexpr (R3) {
    wait_vm_all();  // Q is in LDS
    // DOFs = Q + acc, three quantities at a time with every read before the
    // writes (rows past N share the dummy slot)
    const int ln = opaque(lane), go = ln >> 4, co = ln & 15;
#pragma unroll
    for (int qb = 0; qb < 9; qb += 3) {
      float o[3][NT][4];
#pragma unroll
      for (int q = 0; q < 3; q++)
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
          for (int v = 0; v < 4; v++) o[q][t][v] = l_ti[lds_idx<O>(qb + q, 16 * t + 4 * go + v, N, co)];
#pragma unroll
      for (int q = 0; q < 3; q++)
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
          for (int v = 0; v < 4; v++)
            l_ti[lds_idx<O>(qb + q, 16 * t + 4 * go + v, N, co)] = o[q][t][v] + acc[qb + q][t][v];
    }
  } else {
    __syncthreads();  // the surface integral's reads of tInt are done
    const int ln = opaque(lane), go = ln >> 4, co = ln & 15;
#pragma unroll
    for (int q = 0; q < 9; q++)
#pragma unroll
      for (int t = 0; t < NT; t++)
#pragma unroll
        for (int v = 0; v < 4; v++) l_ti[lds_idx<O>(q, 16 * t + 4 * go + v, N, co)] = acc[q][t][v];
  }
  __syncthreads();
  {
    float4* gq = reinterpret_cast<float4*>(dofs + el * D::SZ);
    const float4* lq = reinterpret_cast<const float4*>(l_ti);
    for (int i = lane; i < D::NVEC; i += 64) gq[i] = lq[i];
  }
}

// ---------------------------------------------------------------- neigh
template <int O, bool R3>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
void k_neigh(SolverDev<float> s, int64_t first, int64_t n, const float* __restrict__ t_int,
             float* __restrict__ dofs) {
  using D = Dims<O>;
  constexpr int N = D::N, NF = D::NF, NT = D::NT;
  __shared__ float l_nb[D::CHUNKS * 256];  // the neighbour's tInt tile
  const int64_t off = s.xcd_map ? xcd_element(n) : (int64_t)blockIdx.x;
  if (off >= n) return;
  const int64_t el = first + off;
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  float* q0 = dofs + el * D::SZ;
  v4f acc[9][NT];
#pragma unroll
  for (int q = 0; q < 9; q++)
#pragma unroll
    for (int t = 0; t < NT; t++) acc[q][t] = load_d(q0 + q * N * R, N, t, g, c);
#pragma unroll 1
  for (int fa = 0; fa < 4; fa++) {
    const int32_t ne = s.nb.ne[el * 4 + fa];
    if (ne < 0) continue;  // outflow (uniform across the wave)
    __syncthreads();       // the previous face's reads of l_nb are done
    tile_to_lds<D::CHUNKS>(t_int + (int64_t)ne * D::SZ, l_nb, D::NVEC, lane);
    wait_vm_all();
    __syncthreads();
    const int fm = s.nb.fmat[el * 4 + fa];
    if constexpr (R3) {
      constexpr pk::Layout L{O};
      constexpr int NGP = pk::groups(ksteps(N));
      const cfloat* f = as_const(s.fsf + ((s.n_el + el) * 4 + fa) * 54);
      const v4f* pm = reinterpret_cast<const v4f*>(s.dg.pk + L.prj()) + fm * NGP * 64;
      v4f w[3], at[NT];
      load_lift<O>(reinterpret_cast<const v4f*>(s.dg.pk + L.lft()) + fa * NT * 64, at, lane);
      project3<O>(pm, pm[lane], f + 27, l_nb, w, g, c, lane);
      lift3<O>(at, f, w, acc);
    } else {
      face_update<O>(s.dg.fluxLN + fm * N * NF, s.dg.fluxT + fa * NF * N, s.fs + ((s.n_el + el) * 4 + fa) * 81,
                     l_nb, acc, g, c);
    }
  }
#pragma unroll
  for (int q = 0; q < 9; q++)
#pragma unroll
    for (int t = 0; t < NT; t++) store_d(q0 + q * N * R, N, t, g, c, acc[q][t]);
}

// Neighbour step from the face projections the local step wrote: per face
// only the lift of three tiles (the neighbour's tInt is not read).
template <int O>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
void k_neigh_proj(SolverDev<float> s, int64_t first, int64_t n, float* __restrict__ dofs) {
  using D = Dims<O>;
  constexpr int N = D::N, NT = D::NT;
  __shared__ __align__(16) float l_out[D::SZ + kLdsPad];
  const int64_t off = s.xcd_map ? xcd_element(n) : (int64_t)blockIdx.x;
  if (off >= n) return;
  const int64_t el = first + off;
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  float* q0 = dofs + el * D::SZ;
  v4f acc[9][NT];
  neigh_lifts<O>(s, s.proj, el, acc, lane, [&](v4f (&a)[9][NT]) {
#pragma unroll
    for (int q = 0; q < 9; q++)
#pragma unroll
      for (int t = 0; t < NT; t++) a[q][t] = load_d(q0 + q * N * R, N, t, g, c);
  });
  {
    const int ln = opaque(lane), go = ln >> 4, co = ln & 15;
#pragma unroll
    for (int q = 0; q < 9; q++)
#pragma unroll
      for (int t = 0; t < NT; t++)
#pragma unroll
        for (int v = 0; v < 4; v++) l_out[lds_idx<O>(q, 16 * t + 4 * go + v, N, co)] = acc[q][t][v];
  }
  __syncthreads();
  float4* gq = reinterpret_cast<float4*>(q0);
  const float4* lq = reinterpret_cast<const float4*>(l_out);
  for (int i = lane; i < D::NVEC; i += 64) gq[i] = lq[i];
}

unsigned grid_of(int64_t n) { return (unsigned)(8 * ((n + 7) / 8)); }

template <int O>
void launch_local_o(const SolverDev<float>& s, int64_t first, int64_t n, float dt, float* dofs, float* t_int,
                    hipStream_t st) {
  const dim3 grid(grid_of(n)), block(64);
  const float* np = nullptr;
  if (s.sparse_star && s.fs_rank3 && s.proj)
    hipLaunchKernelGGL((k_local<O, true, true, true>), grid, block, 0, st, s, first, n, dt, dofs, t_int, np);
  else if (s.sparse_star && s.fs_rank3)
    hipLaunchKernelGGL((k_local<O, true, true>), grid, block, 0, st, s, first, n, dt, dofs, t_int, np);
  else if (s.fs_rank3 && s.proj)
    hipLaunchKernelGGL((k_local<O, false, true, true>), grid, block, 0, st, s, first, n, dt, dofs, t_int, np);
  else if (s.sparse_star)
    hipLaunchKernelGGL((k_local<O, true, false>), grid, block, 0, st, s, first, n, dt, dofs, t_int, np);
  else if (s.fs_rank3)
    hipLaunchKernelGGL((k_local<O, false, true>), grid, block, 0, st, s, first, n, dt, dofs, t_int, np);
  else
    hipLaunchKernelGGL((k_local<O, false, false>), grid, block, 0, st, s, first, n, dt, dofs, t_int, np);
}

// The neighbour step of the previous time step (projections proj_in) fused
// with this step's local step (projections to s.proj), over all elements.
template <int O>
void launch_fused_o(const SolverDev<float>& s, float dt, const float* proj_in, float* dofs, float* t_int,
                    hipStream_t st) {
  const dim3 grid(grid_of(s.n_el)), block(64);
  if (s.sparse_star)
    hipLaunchKernelGGL((k_local<O, true, true, true, true>), grid, block, 0, st, s, (int64_t)0, s.n_el, dt, dofs,
                       t_int, proj_in);
  else
    hipLaunchKernelGGL((k_local<O, false, true, true, true>), grid, block, 0, st, s, (int64_t)0, s.n_el, dt, dofs,
                       t_int, proj_in);
}

template <int O>
void launch_neigh_o(const SolverDev<float>& s, int64_t first, int64_t n, const float* t_int, float* dofs,
                    hipStream_t st) {
  const dim3 grid(grid_of(n)), block(64);
  if (s.proj)
    hipLaunchKernelGGL(k_neigh_proj<O>, grid, block, 0, st, s, first, n, dofs);
  else if (s.fs_rank3)
    hipLaunchKernelGGL((k_neigh<O, true>), grid, block, 0, st, s, first, n, t_int, dofs);
  else
    hipLaunchKernelGGL((k_neigh<O, false>), grid, block, 0, st, s, first, n, t_int, dofs);
}

}  // namespace

std::vector<float> mfma_pack_operands(const DgTables& tb) {
  const int O = tb.order, N = tb.n, NF = tb.nf;
  const pk::Layout L{O};
  std::vector<float> out((size_t)L.size(), 0.f);
  // value of block `blk` (ng groups) at k-step s for lane l
  auto put = [&](int64_t off, int blk, int ng, int s, int l, double v) {
    out[(size_t)(off + (((int64_t)blk * ng + s / 4) * 64 + l) * 4 + s % 4)] = (float)v;
  };
  const int NT = pk::tiles(N);
  for (int l = 0; l < 64; l++) {
    const int g = l / 16, c = l % 16;
    {  // CK stage 1: -K^T M^-1 (stiffT negated), rows k = 4 s + g < N, columns < n_modes(O - 1)
      const int nn = n_modes(O - 1), tn = pk::tiles(nn), ks = pk::ksteps(N), ng = pk::groups(ks);
      for (int d = 0; d < 3; d++)
        for (int t = 0; t < tn; t++)
          for (int s = 0; s < ks; s++) {
            const int k = 4 * s + g, col = 16 * t + c;
            if (k < N && col < nn) put(L.ck1(), d * tn + t, ng, s, l, -tb.stiffT[((size_t)d * N + k) * N + col]);
          }
    }
    for (int de = 2; de < O; de++) {  // CK stages >= 2: rows 16 u + 4 g + v < n_modes(O - de + 1)
      const int nk = n_modes(O - de + 1), nn = n_modes(O - de), st = pk::next_steps(O, de), ng = pk::groups(st);
      for (int d = 0; d < 3; d++)
        for (int s = 0; s < st; s++) {
          const int k = 16 * (s / 4) + 4 * g + s % 4;
          if (k < nk && c < nn) put(L.ck(de), d, ng, s, l, -tb.stiffT[((size_t)d * N + k) * N + c]);
        }
    }
    {  // volume: K M^-1, rows k = 4 s + g < n_modes(O - 1), all N columns
      const int nv = n_modes(O - 1), ks = pk::ksteps(nv), ng = pk::groups(ks);
      for (int d = 0; d < 3; d++)
        for (int t = 0; t < NT; t++)
          for (int s = 0; s < ks; s++) {
            const int k = 4 * s + g, col = 16 * t + c;
            if (k < nv && col < N) put(L.vol(), d * NT + t, ng, s, l, tb.stiffV[((size_t)d * N + k) * N + col]);
          }
    }
    {  // projections: fluxL (m < 4) or fluxN (m - 4), [N][NF], rows k = 4 s + g
      const int ks = pk::ksteps(N), ng = pk::groups(ks);
      for (int m = 0; m < 16; m++) {
        const double* f = m < 4 ? tb.fluxL + (size_t)m * N * NF : tb.fluxN + (size_t)(m - 4) * N * NF;
        for (int s = 0; s < ks; s++) {
          const int k = 4 * s + g;
          if (k < N && c < NF) put(L.prj(), m, ng, s, l, f[(size_t)k * NF + c]);
        }
      }
    }
    for (int fa = 0; fa < 4; fa++)  // lifts: fluxT [NF][N], rows 4 g + v
      for (int t = 0; t < NT; t++)
        for (int v = 0; v < 4; v++) {
          const int k = 4 * g + v, col = 16 * t + c;
          if (k < NF && col < N) put(L.lft(), fa * NT + t, 1, v, l, tb.fluxT[((size_t)fa * NF + k) * N + col]);
        }
  }
  return out;
}

bool mfma_supported(int order, int n_runs, int precision) {
  return precision == 32 && n_runs == 16 && order >= 2 && order <= 5;
}

template <>
int launch_local_mfma<float>(const SolverDev<float>& s, int64_t first, int64_t n, float dt, float* dofs,
                             float* t_int, hipStream_t st) {
  if (s.n_runs != 16) return fail(EDGE_ERR_UNSUPPORTED, "MFMA-tiled kernels: 16 runs");
  if (n <= 0) return EDGE_OK;
  switch (s.order) {
    case 2: launch_local_o<2>(s, first, n, dt, dofs, t_int, st); break;
    case 3: launch_local_o<3>(s, first, n, dt, dofs, t_int, st); break;
    case 4: launch_local_o<4>(s, first, n, dt, dofs, t_int, st); break;
    case 5: launch_local_o<5>(s, first, n, dt, dofs, t_int, st); break;
    default: return fail(EDGE_ERR_UNSUPPORTED, "order");
  }
  EDGE_HIP_TRY(hipGetLastError());
  return EDGE_OK;
}

template <>
int launch_neigh_mfma<float>(const SolverDev<float>& s, int64_t first, int64_t n, const float* t_int, float* dofs,
                             hipStream_t st) {
  if (s.n_runs != 16) return fail(EDGE_ERR_UNSUPPORTED, "MFMA-tiled kernels: 16 runs");
  if (n <= 0) return EDGE_OK;
  switch (s.order) {
    case 2: launch_neigh_o<2>(s, first, n, t_int, dofs, st); break;
    case 3: launch_neigh_o<3>(s, first, n, t_int, dofs, st); break;
    case 4: launch_neigh_o<4>(s, first, n, t_int, dofs, st); break;
    case 5: launch_neigh_o<5>(s, first, n, t_int, dofs, st); break;
    default: return fail(EDGE_ERR_UNSUPPORTED, "order");
  }
  EDGE_HIP_TRY(hipGetLastError());
  return EDGE_OK;
}

int launch_fused_mfma(const SolverDev<float>& s, float dt, const float* proj_in, float* dofs, float* t_int,
                      hipStream_t st) {
  if (s.n_runs != 16) return fail(EDGE_ERR_UNSUPPORTED, "fused step: 16 runs");
  if (!s.sparse_star && !s.fs_rank3) return fail(EDGE_ERR_UNSUPPORTED, "fused step: rank-3 flux solvers");
  if (!s.fs_rank3 || !s.proj || !proj_in) return fail(EDGE_ERR_UNSUPPORTED, "fused step: face projections");
  if (s.n_el <= 0) return EDGE_OK;
  switch (s.order) {
    case 2: launch_fused_o<2>(s, dt, proj_in, dofs, t_int, st); break;
    case 3: launch_fused_o<3>(s, dt, proj_in, dofs, t_int, st); break;
    case 4: launch_fused_o<4>(s, dt, proj_in, dofs, t_int, st); break;
    case 5: launch_fused_o<5>(s, dt, proj_in, dofs, t_int, st); break;
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
