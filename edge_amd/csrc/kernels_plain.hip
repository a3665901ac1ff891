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
// Per-element tables that no kernel writes (star matrices, flux solvers,
// source matrices) are read through the constant address space: wave-uniform
// values by scalar loads, which the scalar cache serves with a short latency
// and which land in scalar registers, not in vector ones.
typedef const __attribute__((address_space(4))) double cdouble;
__device__ __forceinline__ const cdouble* as_const(const double* p) { return (const cdouble*)p; }

__device__ __forceinline__ v4d mfma64(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

constexpr int tiles16(int n) { return (n + 15) / 16; }
constexpr int ksteps4(int n) { return (n + 3) / 4; }

// LDS row stride of an A-operand tile whose k-steps cover P4(n) columns.
// Lanes (g, c) of an MFMA read row c, column 4 s + g; with a stride that is
// a multiple of 4 doubles, rows c and c + 8 (c + 4, ...) fall on the same
// banks, so the tiles are padded past P4(n).
#ifndef EDGE_LDS_PAD
#define EDGE_LDS_PAD 1
#endif
constexpr int lds_ld(int n) { return PadLayout::p4(n) + EDGE_LDS_PAD; }

__device__ __forceinline__ v4d zero4d() { return v4d{0.0, 0.0, 0.0, 0.0}; }

// blockIdx -> element offset: each of the 8 XCDs (workgroups dispatched round
// robin) gets a contiguous element range, so neighbours share its L2 (the
// autotuner may switch this schedule off: workgroup i -> element i).  The
// grid then holds 8 ceil(n / 8) workgroups.
__device__ __forceinline__ int64_t xcd_element(int64_t n) {
  const int64_t chunk = (n + 7) / 8;
  const int64_t b = blockIdx.x;
  return (b & 7) * chunk + (b >> 3);
}

// Guarded loads without branches: clamped index, value masked afterwards.
__device__ __forceinline__ double ld(const double* __restrict__ p, int idx, bool ok) {
  const double v = p[ok ? idx : 0];
  return ok ? v : 0.0;
}

// X[q][n] = a (rows q < 9, columns n < nn) for one tile into LDS
__device__ __forceinline__ void store_tile(double* __restrict__ x, int ldx, int nn, const v4d& a, int g, int c) {
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int q = g + 4 * v;
    if (q < 9 && c < nn) x[q * ldx + c] = a[v];
  }
}

// X[q][n] = acc (rows q0 <= q < MR, columns n < nn) into LDS
template <int NT, int MR = 9, int Q0 = 0>
__device__ __forceinline__ void store_tiles(double* __restrict__ x, int ldx, int nn, const v4d (&acc)[NT], int g,
                                            int c) {
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int q = g + 4 * v, n = 16 * t + c;
      if (q >= Q0 && q < MR && n < nn) x[q * ldx + n] = acc[t][v];
    }
}

template <int O>
struct PlainDims {
  static constexpr int N = n_modes(O), NF = n_face_modes(O), NT = tiles16(N), SZ = 9 * N;
};

// ---------------------------------------------------------------- elastic
// The kernels keep their LDS tiles with padded leading dimensions (P4(N) for
// mode rows, P4(NF) for face-mode rows) and read B operands from the
// zero-padded matrix copies (PadLayout), so no operand load needs a guard:
// A rows past the operand's last (lanes c >= 9, or 15 with attenuation)
// repeat that row, and the D rows they produce are never stored; the k-step tail of every LDS A operand (columns
// NK to P4(NK) - 1) holds zeros, written by the operand's producer.

// acc[t] += A B_t for A [MR][4 KS] in LDS (leading dimension LDA) and the
// B_t [4 KS][16] (leading dimension LDB, padded copy) TS elements apart: the
// 16-column tiles of one matrix (TS = 16) or separate matrices; operands
// loaded one k-step ahead
template <int NT, int KS, int LDA, int LDB, int MR = 9, int TS = 16>
__device__ __forceinline__ void mmp(const double* __restrict__ a, const double* __restrict__ b, v4d (&acc)[NT],
                                    int g, int c) {
  const double* ap = a + (c < MR ? c : MR - 1) * LDA + g;
  const double* bp = b + g * LDB + c;
  double an, bn[NT];
  auto load = [&](int s) {
    an = ap[4 * s];
#pragma unroll
    for (int t = 0; t < NT; t++) bn[t] = bp[4 * s * LDB + TS * t];
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

// The B operands of mmp for every k-step at once, bq[s][t] = B_t[4 s + g][c]:
// requested before the A tile's producer runs, their latency (the shared
// matrices come from L2) overlaps it instead of stalling each k-step.
template <int NT, int KS, int LDB, int TS = 16>
__device__ __forceinline__ void load_b(const double* __restrict__ b, double (&bq)[KS][NT], int g, int c) {
  const double* bp = b + g * LDB + c;
#pragma unroll
  for (int s = 0; s < KS; s++)
#pragma unroll
    for (int t = 0; t < NT; t++) bq[s][t] = bp[4 * s * LDB + TS * t];
}

// mmp with the B operands of load_b
template <int NT, int KS, int LDA, int MR = 9>
__device__ __forceinline__ void mmq(const double* __restrict__ a, const double (&bq)[KS][NT], v4d (&acc)[NT], int g,
                                    int c) {
  const double* ap = a + (c < MR ? c : MR - 1) * LDA + g;
  double an = ap[0];
#pragma unroll
  for (int s = 0; s < KS; s++) {
    const double av = an;
    if (s + 1 < KS) an = ap[4 * (s + 1)];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = mfma64(av, bq[s][t], acc[t]);
  }
}

// acc += A B for A [9][NK] in global memory (the reference layout, leading
// dimension NK) and B [P4(NK)][16] (leading dimension LDB, padded copy); every
// k-step's operands requested at once (one memory round trip); only the last
// k-step's A load is guarded
template <int NK, int LDB>
__device__ __forceinline__ void mmg(const double* __restrict__ a, const double* __restrict__ b, v4d& acc, int g,
                                    int c) {
  constexpr int KS = ksteps4(NK);
  const double* ap = a + (c < 9 ? c : 8) * NK + g;
  const double* bp = b + g * LDB + c;
  double aq[KS], bq[KS];
#pragma unroll
  for (int s = 0; s < KS; s++) {
    aq[s] = (4 * s + 3 < NK) ? ap[4 * s] : ld(ap, 4 * s, 4 * s + g < NK);
    bq[s] = bp[4 * s * LDB];
  }
#pragma unroll
  for (int s = 0; s < KS; s++) acc = mfma64(aq[s], bq[s], acc);
}

// Y_f = F_f X_f in place for the four faces' tiles X_f [9][P4(NF)] at
// x + 9 P4(NF) f: columns k < NF, zeros for NF <= k < P4(NF); one lane per
// (face, column), lane = P4(NF) f + k, with the flux solver F_f at f + 81 f
// (per-lane coefficients, four distinct per wave)
template <int NF>
__device__ __forceinline__ void mixf4(const double* __restrict__ f, double* __restrict__ x, int lane) {
  constexpr int LDF = PadLayout::p4(NF);
  static_assert(4 * LDF <= 64, "one lane per face and column");
  if (lane < 4 * LDF) {
    const int fa = lane / LDF, k = lane - fa * LDF;
    double* col = x + fa * 9 * LDF + k;
    const double* fm = f + fa * 81;
    const bool in = k < NF;
    double xv[9];
#pragma unroll
    for (int p = 0; p < 9; p++) xv[p] = in ? col[p * LDF] : 0.0;
#pragma unroll
    for (int q = 0; q < 9; q++) {
      double z = 0.0;
#pragma unroll
      for (int p = 0; p < 9; p++) z += fm[q * 9 + p] * xv[p];
      col[q * LDF] = z;
    }
  }
}

// Y[q][k] = sum_p S[q][p] X[p][k] for k < NK, and 0 for NK <= k < P4(NK) (one
// lane per k): the elastic star's non-zeros (SP) or all 81 entries; S
// wave-uniform; X and Y leading dimensions at least P4(NK)
template <bool SP, int NK>
__device__ __forceinline__ void mixp(const cdouble* __restrict__ sd, const double* __restrict__ x, int ldx,
                                     double* __restrict__ y, int ldy, int lane) {
  for (int k = lane; k < PadLayout::p4(NK); k += 64) {
    const bool in = k < NK;
    double xv[9];
#pragma unroll
    for (int p = 0; p < 9; p++) xv[p] = in ? x[p * ldx + k] : 0.0;
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

// der_DE = acc into x (columns n < NN, leading dimension LD) and tInt += sca
// der_DE into the tInt tiles ti held in registers (the same D layout)
template <int NT, int NN, int LD, int NTN>
__device__ __forceinline__ void store_der(double* __restrict__ x, v4d (&ti)[NTN], double sca, const v4d (&acc)[NT],
                                          int g, int c) {
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int q = g + 4 * v, n = 16 * t + c;
      if (q < 9 && n < NN) x[q * LD + n] = acc[t][v];
      ti[t][v] += n < NN ? sca * acc[t][v] : 0.0;
    }
}

// CK stage DE (and the following ones): l_der holds der_{DE-1}; the stage
// overwrites it with der_DE once the last direction has read it, and adds
// sca dt / (DE + 1) der_DE to the tInt tiles ti (TimePredVanilla.hpp:242-306).
// The directions' A operands alternate between two LDS buffers, so that one
// direction's mix runs while the previous direction's products are in flight.
template <int O, int DE, bool SP, int NTN>
__device__ __forceinline__ void ck_stage(const double* __restrict__ stiff, const cdouble* __restrict__ star,
                                         double dt, double sca, double* l_der, double* l_a0, double* l_a1,
                                         v4d (&ti)[NTN], int lane, int g, int c) {
  if constexpr (DE < O) {
    using P = PadLayout;
    constexpr int N = n_modes(O), LD = P::p4(N), LQ = P::p16(N);
    constexpr int NK = n_modes(O - DE + 1), NN = n_modes(O - DE), NT = tiles16(NN), KS = ksteps4(NK);
    v4d acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) acc[t] = zero4d();
    mixp<SP, NK>(star, l_der, LD, l_a0, LD, lane);
    __syncthreads();
    mmp<NT, KS, LD, LQ>(l_a0, stiff, acc, g, c);
    mixp<SP, NK>(star + 81, l_der, LD, l_a1, LD, lane);
    __syncthreads();
    mmp<NT, KS, LD, LQ>(l_a1, stiff + LD * LQ, acc, g, c);
    mixp<SP, NK>(star + 162, l_der, LD, l_a0, LD, lane);
    __syncthreads();
    mmp<NT, KS, LD, LQ>(l_a0, stiff + 2 * LD * LQ, acc, g, c);
    sca *= dt / double(DE + 1);
    store_der<NT, NN, LD>(l_der, ti, sca, acc, g, c);
    __syncthreads();
    ck_stage<O, DE + 1, SP>(stiff, star, dt, sca, l_der, l_a0, l_a1, ti, lane, g, c);
  }
}

template <int O, bool SP>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4)))
void k_local_p64(SolverDev<double> s, int64_t first, int64_t n_el, double dt, double* __restrict__ dofs,
                 double* __restrict__ t_int) {
  using D = PlainDims<O>;
  using P = PadLayout;
  constexpr int N = D::N, NF = D::NF, NT = D::NT, SZ = D::SZ;
  constexpr int LD = P::p4(N), LQ = P::p16(N), LDF = P::p4(NF), LQF = P::p16(NF);
  constexpr int LDA = 2 * 9 * LD > 4 * 9 * LDF ? 2 * 9 * LD : 4 * 9 * LDF;
  // l_da: der and the A operands (CK, volume), then the four faces' tiles
  __shared__ double l_ti[9 * LD], l_da[LDA];
  double* l_der = l_da;
  double* l_a = l_da + 9 * LD;
  const int64_t off = s.xcd_map ? xcd_element(n_el) : (int64_t)blockIdx.x;
  if (off >= n_el) return;  // one wave per workgroup: no barrier is left waiting
  const int64_t el = first + off;
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const cdouble* star = as_const(s.star + el * 243);
  const double* pad = s.dg.pad;
  {
    constexpr int IT = (SZ + 63) / 64;
    double v[IT];
#pragma unroll
    for (int j = 0; j < IT; j++) v[j] = ld(dofs + el * SZ, lane + 64 * j, lane + 64 * j < SZ);
#pragma unroll
    for (int j = 0; j < IT; j++) {
      const int i = lane + 64 * j, q = i / N, n = i - q * N;
      if (i < SZ) l_der[q * LD + n] = v[j];
    }
  }
  // tInt in registers during the CK (D layout), dt Q to start with
  v4d ti[NT];
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int q = g + 4 * v, n = 16 * t + c;
      ti[t][v] = dt * ld(dofs + el * SZ, q * N + n, q < 9 && n < N);
    }
  __syncthreads();
  // ---- CK: der_de = sum_d (S_d der_{de-1}) (-K_d^T M^-1), tInt += dt^{de+1}/(de+1)! der_de;
  // l_ti serves as the second A buffer until tInt is stored
  ck_stage<O, 1, SP>(pad, star, dt, dt, l_der, l_a, l_ti, ti, lane, g, c);
  // tInt out, and into LDS with its k-step tail (zeros: the face projections
  // read it as an A operand)
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int q = g + 4 * v, n = 16 * t + c;
      if (q < 9 && n < N) t_int[el * SZ + q * N + n] = ti[t][v];
      if (q < 9 && n < LD) l_ti[q * LD + n] = ti[t][v];
    }
  __syncthreads();

  // ---- volume: dofs += sum_d (S_d tInt) (K_d M^-1), from Q read again
  v4d acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int q = g + 4 * v, n = 16 * t + c;
      acc[t][v] = ld(dofs + el * SZ, q * N + n, q < 9 && n < N);
    }
  // (the directions' A operands alternate between l_a and the free der tile)
  constexpr int NV = n_modes(O - 1);
#pragma unroll
  for (int d = 0; d < 3; d++) {
    double* a = d == 1 ? l_der : l_a;
    mixp<SP, NV>(star + d * 81, l_ti, LD, a, LD, lane);
    __syncthreads();
    mmp<NT, ksteps4(NV), LD, LQ>(a, pad + P::stiff_v(N) + d * LD * LQ, acc, g, c);
  }
  __syncthreads();
  // ---- local surface: dofs += sum_fa ((fs0_fa (tInt F^L_fa)) F^T_fa): the
  // four projections as independent chains on the shared A operand, one
  // flux-solver mix for all faces, then the four lifts
  {
    v4d s0[4];
#pragma unroll
    for (int f = 0; f < 4; f++) s0[f] = zero4d();
    mmp<4, ksteps4(N), LD, LQF, 9, LD * LQF>(l_ti, pad + P::flux_ln(N), s0, g, c);
#pragma unroll
    for (int f = 0; f < 4; f++) store_tile(l_da + f * 9 * LDF, LDF, NF, s0[f], g, c);
  }
  __syncthreads();
  mixf4<NF>(s.fs + el * 324, l_da, lane);
  __syncthreads();
#pragma unroll
  for (int fa = 0; fa < 4; fa++)
    mmp<NT, ksteps4(NF), LDF, LQ>(l_da + fa * 9 * LDF, pad + P::flux_t(N, NF) + fa * LDF * LQ, acc, g, c);
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int q = g + 4 * v, n = 16 * t + c;
      if (q < 9 && n < N) dofs[el * SZ + q * N + n] = acc[t][v];
    }
}

template <int O>
__global__ __launch_bounds__(64) void k_neigh_p64(SolverDev<double> s, int64_t first, int64_t n_el,
                                                  const double* __restrict__ t_int, double* __restrict__ dofs) {
  using D = PlainDims<O>;
  using P = PadLayout;
  constexpr int N = D::N, NF = D::NF, NT = D::NT, SZ = D::SZ;
  constexpr int LD = P::p4(N), LQ = P::p16(N), LDF = P::p4(NF), LQF = P::p16(NF);
  __shared__ double l_s0[9 * LDF], l_s1[9 * LDF];
  const int64_t off = s.xcd_map ? xcd_element(n_el) : (int64_t)blockIdx.x;
  if (off >= n_el) return;  // one wave per workgroup: no barrier is left waiting
  const int64_t el = first + off;
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const double* pad = s.dg.pad;
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
    mmg<N, LQF>(t_int + (int64_t)ne * SZ, pad + P::flux_ln(N) + fm * LD * LQF, s0[0], g, c);
    store_tiles<1>(l_s0, LDF, NF, s0, g, c);
    __syncthreads();
    mixp<false, NF>(as_const(s.fs + ((s.n_el + el) * 4 + fa) * 81), l_s0, LDF, l_s1, LDF, lane);
    __syncthreads();
    mmp<NT, ksteps4(NF), LDF, LQ>(l_s1, pad + P::flux_t(N, NF) + fa * LDF * LQ, acc, g, c);
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

// ---------------------------------------------------------------- viscoelastic
// With N_RMS relaxation mechanisms the anelastic star S^A_d (6 x 3, on the
// velocities) and the anelastic flux solvers fs_A (6 x 9) act on the same
// products as the elastic ones, so they take rows 9..14 of the A operand and
// one 16-row tile carries the elastic outputs and the anelastic combinations
// together (src/impl/seismic/kernels/TimePredVanilla.hpp:200-306,
// VolIntVanilla.hpp:189-241, SurfIntVanilla.hpp:212-357, SurfInt.hpp:139-156):
//   CK:      rows 0..8  (S_d der) (-K_d^T M^-1),   rows 9..14 (S^A_d der_vel) (-K_d^T M^-1)
//            (all N modes per derivative with attenuation)
//   volume:  rows 0..8 += (S_d tInt) (K_d M^-1),    rows 9..14 A = (S^A_d tInt_vel) (K_d M^-1)
//   surface: rows 0..8 += (fs (tInt F)) F^T,        rows 9..14 += (fs_A (tInt F)) F^T
// The relaxation terms are elementwise per mode, one lane per mode with the
// anelastic state in registers: E_rm derA_rm into the stress rows,
// derA_rm <- w_rm (A + derA_rm), tIntA_rm += c_de derA_rm; then
// Q[stress] += sum_rm E_rm tIntA_rm and QA_rm += w_rm (A + upA - tIntA_rm);
// the neighbour step adds w_rm upA_nb.

// Y rows 0..8 = S X (elastic star pattern or all 81 entries), rows 9..14 =
// S^A X_vel, for columns k < NK, and 0 for NK <= k < P4(NK) (one lane per k)
template <bool SP, int NK>
__device__ __forceinline__ void mix15(const cdouble* __restrict__ sd, const cdouble* __restrict__ sa,
                                      const double* __restrict__ x, int ldx, double* __restrict__ y, int ldy,
                                      int lane) {
  for (int k = lane; k < PadLayout::p4(NK); k += 64) {
    const bool in = k < NK;
    double xv[9];
#pragma unroll
    for (int p = 0; p < 9; p++) xv[p] = in ? x[p * ldx + k] : 0.0;
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
#pragma unroll
    for (int j = 0; j < 6; j++) y[(9 + j) * ldy + k] = sa[j * 3] * xv[6] + sa[j * 3 + 1] * xv[7] + sa[j * 3 + 2] * xv[8];
  }
}

// Y rows 0..8 = F X, rows 9..14 = F_A X for a flux solver F (9 x 9) and its
// anelastic part F_A (6 x 9), columns k < NK, and 0 for NK <= k < P4(NK)
// (coefficients row by row through scalar loads: all 135 held at once
// would not fit the registers)
template <int NK>
__device__ __forceinline__ void mixf15(const double* __restrict__ f, const double* __restrict__ fa,
                                       const double* __restrict__ x, int ldx, double* __restrict__ y, int ldy,
                                       int lane) {
  for (int k = lane; k < PadLayout::p4(NK); k += 64) {
    const bool in = k < NK;
    double xv[9];
#pragma unroll
    for (int p = 0; p < 9; p++) xv[p] = in ? x[p * ldx + k] : 0.0;
#pragma unroll 1
    for (int q = 0; q < 15; q++) {
      const cdouble* row = (const cdouble*)(q < 9 ? f + q * 9 : fa + (q - 9) * 9);
      double z = 0.0;
#pragma unroll
      for (int p = 0; p < 9; p++) z += row[p] * xv[p];
      y[q * ldy + k] = z;
    }
  }
}

// mixf15 with the coefficients gathered in LDS: cf [15][9], the flux
// solver's 9 rows then its anelastic part's 6 (gather_fs)
template <int NK>
__device__ __forceinline__ void mixf15l(const double* __restrict__ cf, const double* __restrict__ x, int ldx,
                                        double* __restrict__ y, int ldy, int lane) {
  for (int k = lane; k < PadLayout::p4(NK); k += 64) {
    const bool in = k < NK;
    double xv[9];
#pragma unroll
    for (int p = 0; p < 9; p++) xv[p] = in ? x[p * ldx + k] : 0.0;
#pragma unroll 3
    for (int q = 0; q < 15; q++) {
      const double* row = cf + q * 9;
      double z = 0.0;
#pragma unroll
      for (int p = 0; p < 9; p++) z += row[p] * xv[p];
      y[q * ldy + k] = z;
    }
  }
}
constexpr int kFsCoef = 135;
__device__ __forceinline__ void gather_fs(double* __restrict__ cf, const double* __restrict__ f,
                                          const double* __restrict__ fa, int lane) {
  for (int i = lane; i < kFsCoef; i += 64) cf[i] = i < 81 ? f[i] : fa[i - 81];
}

// The wave-uniform value x, opaque to the optimiser: per-element
// coefficients addressed with it are loaded where they are used instead of
// being hoisted out of the time-derivative loop (and held in registers).
__device__ __forceinline__ int opaque_u(int x) {
  asm volatile("" : "+s"(x));
  return x;
}

// With the elastic star pattern the per-element coefficients of the
// elementwise work are gathered in LDS once per element, in the order they
// are used: per direction the star's 24 non-zeros (rows in order, columns
// kStarCol) and the anelastic star (6 x 3), then per mechanism the source
// matrix's 12 pattern entries (normal block, then the shear diagonal).
// Through scalar loads the scattered entries were a chain of small loads,
// each waited for in full; from LDS every lane reads the same address (a
// broadcast) and the reads pipeline.
constexpr int kCfDir = 42, kCfSrc = 3 * kCfDir;
struct StarNnz {
  int idx[24] = {};
  constexpr StarNnz() {
    int i = 0;
    for (int q = 0; q < 9; q++)
      for (int j = 0; j < kStarNnzRow[q]; j++) idx[i++] = q * 9 + kStarCol[q][j];
  }
};
constexpr StarNnz kStarNnz{};
__device__ __forceinline__ double gather_cf(const double* __restrict__ star, const double* __restrict__ star_a,
                                            const double* __restrict__ src, int i) {
  if (i < kCfSrc) {
    const int d = i / kCfDir, k = i - d * kCfDir;
    return k < 24 ? star[d * 81 + kStarNnz.idx[k]] : star_a[d * 18 + k - 24];
  }
  const int rm = (i - kCfSrc) / 12, k = (i - kCfSrc) - 12 * rm;
  return src[rm * 36 + (k < 9 ? (k / 3) * 6 + k % 3 : 7 * (k - 6))];
}

// mix15 for the elastic star pattern with the direction's gathered
// coefficients cf (24 star non-zeros, then the anelastic star)
template <int NK>
__device__ __forceinline__ void mix15s(const double* __restrict__ cf, const double* __restrict__ x, int ldx,
                                       double* __restrict__ y, int ldy, int lane) {
  for (int k = lane; k < PadLayout::p4(NK); k += 64) {
    const bool in = k < NK;
    double xv[9];
#pragma unroll
    for (int p = 0; p < 9; p++) xv[p] = in ? x[p * ldx + k] : 0.0;
    int o = 0;
#pragma unroll
    for (int q = 0; q < 9; q++) {
      double z = 0.0;
#pragma unroll
      for (int j = 0; j < kStarNnzRow[q]; j++) z += cf[o + j] * xv[kStarCol[q][j]];
      o += kStarNnzRow[q];
      y[q * ldy + k] = z;
    }
#pragma unroll
    for (int j = 0; j < 6; j++)
      y[(9 + j) * ldy + k] = cf[24 + j * 3] * xv[6] + cf[24 + j * 3 + 1] * xv[7] + cf[24 + j * 3 + 2] * xv[8];
  }
}

// add_src with the 12 gathered entries of E
__device__ __forceinline__ void add_src_s(const double* __restrict__ e, const double (&x)[6], double (&y)[9]) {
#pragma unroll
  for (int q = 0; q < 3; q++) y[q] += e[q * 3] * x[0] + e[q * 3 + 1] * x[1] + e[q * 3 + 2] * x[2];
#pragma unroll
  for (int q = 3; q < 6; q++) y[q] += e[6 + q] * x[q];
}

// y[0..5] += E x for an anelastic source matrix E (6 x 6) of the source
// pattern (src/impl/seismic/setups/ViscoElasticity.h:379-458): the
// normal-stress block rows 0..2 x columns 0..2 and the shear diagonal
__device__ __forceinline__ void add_src(const cdouble* __restrict__ e, const double (&x)[6], double (&y)[9]) {
#pragma unroll
  for (int q = 0; q < 3; q++) y[q] += e[q * 6] * x[0] + e[q * 6 + 1] * x[1] + e[q * 6 + 2] * x[2];
#pragma unroll
  for (int q = 3; q < 6; q++) y[q] += e[q * 6 + q] * x[q];
}

template <int O, int NRM, bool SP>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
void k_local_pv64(SolverDev<double> s, int64_t first, int64_t n_el, double dt, double* __restrict__ dofs,
                  double* __restrict__ dofs_a, double* __restrict__ t_int) {
  using D = PlainDims<O>;
  using P = PadLayout;
  constexpr int N = D::N, NF = D::NF, NT = D::NT, SZ = D::SZ, SZA = 6 * N;
  constexpr int NV = n_modes(O - 1), TV = tiles16(NV);
  constexpr int LD = P::p4(N), LQ = P::p16(N), LDF = P::p4(NF), LQF = P::p16(NF);
  constexpr int SL = lds_ld(N), SLF = lds_ld(NF);  // LDS row strides
  static_assert(N <= 64, "one lane per mode");
  // l_x: der (rows 0..8) and the anelastic combination (rows 9..14); l_a: A
  // operands; l_ta: tIntA [rm][6][mode]
  // l_x and l_a also hold two faces' projections and mixed tiles
  constexpr int LXX = 15 * SL > 2 * 9 * SLF ? 15 * SL : 2 * 9 * SLF;
  constexpr int LXA = 15 * SL > 2 * 15 * SLF ? 15 * SL : 2 * 15 * SLF;
  __shared__ double l_ti[9 * SL], l_x[LXX], l_a[LXA], l_ta[NRM * 6 * N];
  // gathered coefficients (SP): the CK's and the volume's, then two faces'
  // flux solvers in the surface integral
  __shared__ double l_cf[SP ? (kCfSrc + 12 * NRM > 2 * kFsCoef ? kCfSrc + 12 * NRM : 2 * kFsCoef) : 1];
  const int64_t off = s.xcd_map ? xcd_element(n_el) : (int64_t)blockIdx.x;
  if (off >= n_el) return;  // one wave per workgroup: no barrier is left waiting
  const int64_t el = first + off;
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const cdouble* star = as_const(s.star + el * 243);
  const cdouble* star_a = as_const(s.star_a + el * 54);
  const cdouble* src = as_const(s.src_a + el * NRM * 36);
  const double* pad = s.dg.pad;
  double* qa = dofs_a + el * NRM * SZA;
  if constexpr (SP)
    for (int i = lane; i < kCfSrc + 12 * NRM; i += 64)
      l_cf[i] = gather_cf(s.star + el * 243, s.star_a + el * 54, s.src_a + el * NRM * 36, i);
  {
    constexpr int IT = (SZ + 63) / 64;
    double v[IT];
#pragma unroll
    for (int j = 0; j < IT; j++) v[j] = ld(dofs + el * SZ, lane + 64 * j, lane + 64 * j < SZ);
#pragma unroll
    for (int j = 0; j < IT; j++) {
      const int i = lane + 64 * j, q = i / N, n = i - q * N;
      if (i < SZ) {
        l_x[q * SL + n] = v[j];
        l_ti[q * SL + n] = dt * v[j];
      }
    }
    // tInt's k-step tail: the face projections read it as an A operand
    if constexpr (LD > N)
      for (int i = lane; i < 9 * (LD - N); i += 64) l_ti[(i / (LD - N)) * SL + N + i % (LD - N)] = 0.0;
  }
  // anelastic state of mode n = lane: derA in registers, tIntA in LDS
  double da[NRM][6];
#pragma unroll
  for (int rm = 0; rm < NRM; rm++)
#pragma unroll
    for (int j = 0; j < 6; j++) {
      da[rm][j] = ld(qa, (rm * 6 + j) * N + lane, lane < N);
      if (lane < N) l_ta[(rm * 6 + j) * N + lane] = dt * da[rm][j];
    }
  __syncthreads();

  // ---- CK
  double sca = dt;
#pragma unroll 1
  for (int de = 1; de < O; de++) {
    v4d acc[TV];
#pragma unroll
    for (int t = 0; t < TV; t++) acc[t] = zero4d();
#pragma unroll
    for (int d = 0; d < 3; d++) {
      double bq[ksteps4(N)][TV];
      load_b<TV, ksteps4(N), LQ>(pad + d * LD * LQ, bq, g, c);
      if constexpr (SP)
        mix15s<N>(l_cf + opaque_u(d * kCfDir), l_x, SL, l_a, SL, lane);
      else
        mix15<SP, N>(star + opaque_u(d * 81), star_a + opaque_u(d * 18), l_x, SL, l_a, SL, lane);
      __syncthreads();
      mmq<TV, ksteps4(N), SL, 15>(l_a, bq, acc, g, c);
      __syncthreads();
    }
    store_tiles<TV, 15>(l_x, SL, NV, acc, g, c);
    __syncthreads();
    sca *= dt / double(de + 1);
    if (lane < N) {
      const int n = lane;
      double y[9], a[6];
#pragma unroll
      for (int q = 0; q < 9; q++) y[q] = n < NV ? l_x[q * SL + n] : 0.0;
#pragma unroll
      for (int j = 0; j < 6; j++) a[j] = n < NV ? l_x[(9 + j) * SL + n] : 0.0;
#pragma unroll
      for (int rm = 0; rm < NRM; rm++) {
        if constexpr (SP)
          add_src_s(l_cf + opaque_u(kCfSrc + rm * 12), da[rm], y);
        else
          add_src(src + opaque_u(rm * 36), da[rm], y);
        const double w = as_const(s.rfs)[rm];
#pragma unroll
        for (int j = 0; j < 6; j++) {
          da[rm][j] = w * (a[j] + da[rm][j]);
          l_ta[(rm * 6 + j) * N + n] += sca * da[rm][j];
        }
      }
#pragma unroll
      for (int q = 0; q < 9; q++) {
        l_x[q * SL + n] = y[q];
        l_ti[q * SL + n] += sca * y[q];
      }
    }
    __syncthreads();
  }
  for (int i = lane; i < SZ; i += 64) t_int[el * SZ + i] = l_ti[(i / N) * SL + i % N];

  // ---- volume: rows 0..8 from Q, rows 9..14 the anelastic combination A
  v4d acc[NT];
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int q = g + 4 * v, n = 16 * t + c;
      acc[t][v] = ld(dofs + el * SZ, q * N + n, q < 9 && n < N);
    }
  for (int d = 0; d < 3; d++) {
    double bq[ksteps4(NV)][NT];
    load_b<NT, ksteps4(NV), LQ>(pad + P::stiff_v(N) + d * LD * LQ, bq, g, c);
    if constexpr (SP)
      mix15s<NV>(l_cf + d * kCfDir, l_ti, SL, l_a, SL, lane);
    else
      mix15<SP, NV>(star + d * 81, star_a + d * 18, l_ti, SL, l_a, SL, lane);
    __syncthreads();
    mmq<NT, ksteps4(NV), SL, 15>(l_a, bq, acc, g, c);
    __syncthreads();
  }
  // sum_rm E_rm tIntA_rm into the stress rows (through LDS)
  if (lane < N) {
    double y[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int rm = 0; rm < NRM; rm++) {
      double ta[6];
#pragma unroll
      for (int j = 0; j < 6; j++) ta[j] = l_ta[(rm * 6 + j) * N + lane];
      if constexpr (SP)
        add_src_s(l_cf + kCfSrc + rm * 12, ta, y);
      else
        add_src(src + rm * 36, ta, y);
    }
#pragma unroll
    for (int j = 0; j < 6; j++) l_x[j * SL + lane] = y[j];
  }
  __syncthreads();
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int q = g + 4 * v, n = 16 * t + c;
      if (q < 6 && n < N) acc[t][v] += l_x[q * SL + n];
    }
  __syncthreads();
  // ---- local surface
  // (faces in pairs: the two projections as independent chains on the
  // shared tInt operand, both flux-solver mixes, then both lifts)
  const double* fs0 = s.fs + el * 324;
  const double* fsa0 = s.fs_a + el * 4 * 54;
#pragma unroll 1
  for (int fp = 0; fp < 4; fp += 2) {
    v4d s0[2] = {zero4d(), zero4d()};
    if constexpr (SP) {  // l_cf's last reads are behind the previous barrier
      gather_fs(l_cf, fs0 + fp * 81, fsa0 + fp * 54, lane);
      gather_fs(l_cf + kFsCoef, fs0 + (fp + 1) * 81, fsa0 + (fp + 1) * 54, lane);
    }
    mmp<2, ksteps4(N), SL, LQF, 9, LD * LQF>(l_ti, pad + P::flux_ln(N) + fp * LD * LQF, s0, g, c);
    double bq[2][ksteps4(NF)][NT];
#pragma unroll
    for (int f = 0; f < 2; f++) load_b<NT, ksteps4(NF), LQ>(pad + P::flux_t(N, NF) + (fp + f) * LDF * LQ, bq[f], g, c);
    store_tile(l_x, SLF, NF, s0[0], g, c);
    store_tile(l_x + 9 * SLF, SLF, NF, s0[1], g, c);
    __syncthreads();
    if constexpr (SP) {
      mixf15l<NF>(l_cf, l_x, SLF, l_a, SLF, lane);
      mixf15l<NF>(l_cf + kFsCoef, l_x + 9 * SLF, SLF, l_a + 15 * SLF, SLF, lane);
    } else {
      mixf15<NF>(fs0 + fp * 81, fsa0 + fp * 54, l_x, SLF, l_a, SLF, lane);
      mixf15<NF>(fs0 + (fp + 1) * 81, fsa0 + (fp + 1) * 54, l_x + 9 * SLF, SLF, l_a + 15 * SLF, SLF, lane);
    }
    __syncthreads();
    mmq<NT, ksteps4(NF), SLF, 15>(l_a, bq[0], acc, g, c);
    mmq<NT, ksteps4(NF), SLF, 15>(l_a + 15 * SLF, bq[1], acc, g, c);
    __syncthreads();
  }
  // ---- DOFs out; rows 9..14 (A + upA) through LDS to the anelastic update
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int q = g + 4 * v, n = 16 * t + c;
      if (q < 9 && n < N) dofs[el * SZ + q * N + n] = acc[t][v];
    }
  store_tiles<NT, 15, 9>(l_x, SL, N, acc, g, c);
  __syncthreads();
  if (lane < N) {
#pragma unroll
    for (int rm = 0; rm < NRM; rm++) {
      const double w = as_const(s.rfs)[rm];
#pragma unroll
      for (int j = 0; j < 6; j++)
        qa[(rm * 6 + j) * N + lane] += w * (l_x[(9 + j) * SL + lane] - l_ta[(rm * 6 + j) * N + lane]);
    }
  }
}

template <int O, int NRM>
__global__ __launch_bounds__(64) void k_neigh_pv64(SolverDev<double> s, int64_t first, int64_t n_el,
                                                   const double* __restrict__ t_int, double* __restrict__ dofs,
                                                   double* __restrict__ dofs_a) {
  using D = PlainDims<O>;
  using P = PadLayout;
  constexpr int N = D::N, NF = D::NF, NT = D::NT, SZ = D::SZ, SZA = 6 * N;
  constexpr int LD = P::p4(N), LQ = P::p16(N), LDF = P::p4(NF), LQF = P::p16(NF);
  constexpr int SLF = lds_ld(NF);  // LDS row stride of the face tiles
  // l_a: the mixed face operand [15][SLF], then the output rows [15][N]
  __shared__ double l_s0[9 * SLF], l_a[15 * (SLF > N ? SLF : N)], l_fc[kFsCoef];
  const int64_t off = s.xcd_map ? xcd_element(n_el) : (int64_t)blockIdx.x;
  if (off >= n_el) return;  // one wave per workgroup: no barrier is left waiting
  const int64_t el = first + off;
  const int lane = threadIdx.x, g = lane >> 4, c = lane & 15;
  const double* pad = s.dg.pad;
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
    // the neighbour-side solver, gathered (l_fc's last reads are behind the previous barrier)
    gather_fs(l_fc, s.fs + ((s.n_el + el) * 4 + fa) * 81, s.fs_a + ((s.n_el + el) * 4 + fa) * 54, lane);
    v4d s0[1] = {zero4d()};
    mmg<N, LQF>(t_int + (int64_t)ne * SZ, pad + P::flux_ln(N) + fm * LD * LQF, s0[0], g, c);
    double bq[ksteps4(NF)][NT];
    load_b<NT, ksteps4(NF), LQ>(pad + P::flux_t(N, NF) + fa * LDF * LQ, bq, g, c);
    store_tiles<1>(l_s0, SLF, NF, s0, g, c);
    __syncthreads();
    mixf15l<NF>(l_fc, l_s0, SLF, l_a, SLF, lane);
    __syncthreads();
    mmq<NT, ksteps4(NF), SLF, 15>(l_a, bq, acc, g, c);
    __syncthreads();
  }
#pragma unroll
  for (int t = 0; t < NT; t++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int q = g + 4 * v, n = 16 * t + c;
      if (q < 9 && n < N) dofs[el * SZ + q * N + n] = acc[t][v];
    }
  store_tiles<NT, 15, 9>(l_a, N, N, acc, g, c);
  __syncthreads();
  if (lane < N) {
    double* qa = dofs_a + el * NRM * SZA;
#pragma unroll
    for (int rm = 0; rm < NRM; rm++) {
      const double w = as_const(s.rfs)[rm];
#pragma unroll
      for (int j = 0; j < 6; j++) qa[(rm * 6 + j) * N + lane] += w * l_a[(9 + j) * N + lane];
    }
  }
}

// one workgroup per element, 8 ceil(n / 8) under the XCD schedule
dim3 grid(const SolverDev<double>& s, int64_t n) { return dim3((unsigned)(s.xcd_map ? (n + 7) / 8 * 8 : n)); }

template <int O, int NRM>
int visco_o(const SolverDev<double>& s, int64_t first, int64_t n, double dt, double* dofs, double* dofs_a,
            double* t_int, const double* tin, bool local, hipStream_t st) {
  if (local) {
    if (s.sparse_star)
      hipLaunchKernelGGL((k_local_pv64<O, NRM, true>), grid(s, n), dim3(64), 0, st, s, first, n, dt, dofs, dofs_a,
                         t_int);
    else
      hipLaunchKernelGGL((k_local_pv64<O, NRM, false>), grid(s, n), dim3(64), 0, st, s, first, n, dt, dofs, dofs_a,
                         t_int);
  } else {
    hipLaunchKernelGGL((k_neigh_pv64<O, NRM>), grid(s, n), dim3(64), 0, st, s, first, n, tin, dofs, dofs_a);
  }
  EDGE_HIP_TRY(hipGetLastError());
  return EDGE_OK;
}

template <int O>
int visco_rm(const SolverDev<double>& s, int64_t first, int64_t n, double dt, double* dofs, double* dofs_a,
             double* t_int, const double* tin, bool local, hipStream_t st) {
  switch (s.n_rms) {
    case 1: return visco_o<O, 1>(s, first, n, dt, dofs, dofs_a, t_int, tin, local, st);
    case 2: return visco_o<O, 2>(s, first, n, dt, dofs, dofs_a, t_int, tin, local, st);
    case 3: return visco_o<O, 3>(s, first, n, dt, dofs, dofs_a, t_int, tin, local, st);
  }
  return fail(EDGE_ERR_UNSUPPORTED, "FP64 MFMA viscoelastic kernels: 1 to 3 mechanisms");
}

int visco_dispatch(const SolverDev<double>& s, int64_t first, int64_t n, double dt, double* dofs, double* dofs_a,
                   double* t_int, const double* tin, bool local, hipStream_t st) {
  if (n <= 0) return EDGE_OK;
  switch (s.order) {
    case 2: return visco_rm<2>(s, first, n, dt, dofs, dofs_a, t_int, tin, local, st);
    case 3: return visco_rm<3>(s, first, n, dt, dofs, dofs_a, t_int, tin, local, st);
    case 4: return visco_rm<4>(s, first, n, dt, dofs, dofs_a, t_int, tin, local, st);
    case 5: return visco_rm<5>(s, first, n, dt, dofs, dofs_a, t_int, tin, local, st);
  }
  return fail(EDGE_ERR_UNSUPPORTED, "order");
}

template <int O>
int local_o(const SolverDev<double>& s, int64_t first, int64_t n, double dt, double* dofs, double* t_int,
            hipStream_t st) {
  if (s.sparse_star)
    hipLaunchKernelGGL((k_local_p64<O, true>), grid(s, n), dim3(64), 0, st, s, first, n, dt, dofs, t_int);
  else
    hipLaunchKernelGGL((k_local_p64<O, false>), grid(s, n), dim3(64), 0, st, s, first, n, dt, dofs, t_int);
  EDGE_HIP_TRY(hipGetLastError());
  return EDGE_OK;
}

template <int O>
int neigh_o(const SolverDev<double>& s, int64_t first, int64_t n, const double* t_int, double* dofs,
            hipStream_t st) {
  hipLaunchKernelGGL((k_neigh_p64<O>), grid(s, n), dim3(64), 0, st, s, first, n, t_int, dofs);
  EDGE_HIP_TRY(hipGetLastError());
  return EDGE_OK;
}

}  // namespace

bool plain64_supported(int order, int n_runs, int prec_bytes) {
  return order >= 2 && order <= 5 && n_runs == 1 && prec_bytes == 8;
}

bool plain64_visco_supported(int order, int n_runs, int prec_bytes, int n_rms) {
  return plain64_supported(order, n_runs, prec_bytes) && n_rms >= 1 && n_rms <= 3;
}

int launch_local_plain64_visco(const SolverDev<double>& s, int64_t first, int64_t n, double dt, double* dofs,
                               double* dofs_a, double* t_int, hipStream_t st) {
  return visco_dispatch(s, first, n, dt, dofs, dofs_a, t_int, nullptr, true, st);
}

int launch_neigh_plain64_visco(const SolverDev<double>& s, int64_t first, int64_t n, const double* t_int,
                               double* dofs, double* dofs_a, hipStream_t st) {
  return visco_dispatch(s, first, n, 0.0, dofs, dofs_a, nullptr, t_int, false, st);
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
