// Device-side data of a solver and the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace edge {

// Reference-element matrices on the device in the solver's precision.
template <typename T>
struct DgDev {
  const T* stiffTn;  // [3][N][N]  -(K_d^T M^-1), negated as src/dg/TimePred.hpp:86 stores it
  const T* stiffV;   // [3][N][N]  K_d M^-1
  const T* fluxLN;   // [16][N][NF] local (0..3) then neighbouring (4 + v*4 + f)
  const T* fluxT;    // [4][NF][N]
  const T* pk = nullptr;  // MFMA A operands packed in lane order (mfma_pack.hpp), or null
  // FP64 single-run kernels: the same matrices zero-padded to whole MFMA
  // operand tiles (layout PadLayout below), or null
  const T* pad = nullptr;
};

// Zero-padded reference-element matrices of the FP64 single-run kernels:
// rows padded to whole k-steps of 4, columns to whole 16-wide tiles, so that
// no B operand load needs a bounds guard. Offsets in elements:
//   stiffTn [3][P4(N)][P16(N)] | stiffV [3][P4(N)][P16(N)] |
//   fluxLN [16][P4(N)][P16(NF)] | fluxT [4][P4(NF)][P16(N)]
struct PadLayout {
  static constexpr int p4(int n) { return (n + 3) / 4 * 4; }
  static constexpr int p16(int n) { return (n + 15) / 16 * 16; }
  static constexpr size_t stiff(int n) { return (size_t)3 * p4(n) * p16(n); }
  static constexpr size_t stiff_v(int n) { return stiff(n); }
  static constexpr size_t flux_ln(int n) { return 2 * stiff(n); }
  static constexpr size_t flux_t(int n, int nf) { return flux_ln(n) + (size_t)16 * p4(n) * p16(nf); }
  static constexpr size_t total(int n, int nf) { return flux_t(n, nf) + (size_t)4 * p4(nf) * p16(n); }
};

// Per element-face neighbour data, precomputed from the mesh.
struct NeighDev {
  const int32_t* ne;   // [nEl][4]: neighbour element (self on free surfaces), -1: skip (outflow)
  const uint8_t* fmat; // [nEl][4]: index into fluxLN for the neighbour's contribution
};

template <typename T>
struct SolverDev {
  int order, n_runs;
  int64_t n_el;
  DgDev<T> dg;
  NeighDev nb;
  const T* star;  // [nEl][3][81]
  const T* fs;    // [2][nEl][4][81]
  bool sparse_star;  // every star matrix has the elastic 24-entry pattern
  // MFMA kernels' workgroup -> element schedule: contiguous element ranges
  // per XCD (true) or workgroup i -> element i (false)
  bool xcd_map = true;
  // kernels with an occupancy variant: the one that keeps more waves per
  // SIMD resident (state moved from registers to LDS)
  bool hi_occ = false;
  // rank-3 factors of the flux solvers [2][nEl][4][54] (U [9][3], then V^T
  // [3][9]); null unless the MFMA-tiled kernels run
  const T* fsf = nullptr;
  bool fs_rank3 = false;  // every flux solver is rank 3 and fsf holds its factors
  // face projections written by the local step for the neighbour step
  // ([nEl][4][3][16][16]); null: the neighbour step reads tInt tiles
  T* proj = nullptr;
  const int32_t* snd_v = nullptr;  // [nEl][4]: receiving element-face (el * 4 + fa) of this face, -1: none
  const uint8_t* snd_fm = nullptr; // [nEl][4]: flux matrix (index into fluxLN) of that projection
  const int32_t* rcv = nullptr;    // [nEl][4]: projection slot (el * 4 + f) feeding this face, -1: none
  // fused neighbour + local launches: the DOFs each local step starts from
  // (after the neighbour step) also go to tap (the DOF layout) for the
  // elements with tap_on[el] != 0; null: none
  const uint8_t* tap_on = nullptr;
  T* tap = nullptr;
  // viscoelastic solvers (n_rms > 0)
  int n_rms = 0;
  const T* rfs = nullptr;     // [n_rms] relaxation frequencies
  const T* star_a = nullptr;  // [nEl][3][6][3] anelastic star matrices
  const T* src_a = nullptr;   // [nEl][n_rms][6][6] anelastic source matrices
  const T* fs_a = nullptr;    // [2][nEl][4][6][9] anelastic flux solvers
  // factors U_A [2][nEl][4][6][3] of the anelastic flux solvers over the
  // elastic factors' V^T (fs_a = U_A V^T); null unless the MFMA viscoelastic
  // kernels run
  const T* fsfa = nullptr;
};

// Non-zero pattern of the elastic star matrices: the union of the patterns
// of the Jacobians A, B, C (src/impl/seismic/setups/Elasticity.cpp:172-250),
// 24 of 81 entries; row q lists its columns.
constexpr int kStarNnzRow[9] = {3, 3, 3, 2, 2, 2, 3, 3, 3};
constexpr int kStarCol[9][3] = {{6, 7, 8}, {6, 7, 8}, {6, 7, 8}, {6, 7, 0}, {7, 8, 0},
                                {6, 8, 0}, {0, 3, 5}, {1, 3, 4}, {2, 4, 5}};

// Generic kernels: one workgroup per element, all tiles in LDS.
template <typename T>
int launch_local_generic(const SolverDev<T>& s, int64_t first, int64_t n, T dt, T* dofs, T* t_int, hipStream_t st);
template <typename T>
int launch_neigh_generic(const SolverDev<T>& s, int64_t first, int64_t n, const T* t_int, T* dofs, hipStream_t st);
bool generic_supported(int order, int n_runs, int prec_bytes);

// Viscoelastic kernels (kernels_visco.hip): one workgroup per element, 1 or 8 runs.
template <typename T>
int launch_local_visco(const SolverDev<T>& s, int64_t first, int64_t n, T dt, T* dofs, T* dofs_a, T* t_int,
                       hipStream_t st);
template <typename T>
int launch_neigh_visco(const SolverDev<T>& s, int64_t first, int64_t n, const T* t_int, T* dofs, T* dofs_a,
                       hipStream_t st);
bool visco_supported(int order, int n_runs, int prec_bytes, int n_rms);

// MFMA viscoelastic kernels (kernels_visco_mfma.hip): FP32, 8 fused runs,
// 1 to 3 mechanisms, elastic star pattern and rank-3 factored flux solvers.
bool visco_mfma_supported(int order, int n_runs, int precision, int n_rms);
int launch_local_visco_mfma(const SolverDev<float>& s, int64_t first, int64_t n, float dt, float* dofs,
                            float* dofs_a, float* t_int, hipStream_t st);
int launch_neigh_visco_mfma(const SolverDev<float>& s, int64_t first, int64_t n, const float* t_int, float* dofs,
                            float* dofs_a, hipStream_t st);
// The neighbour step of step k - 1 (from proj_in, the previous launch's
// projections) and the local step of step k (projections to s.proj) as one
// launch over all elements (FP32, 8 runs, rank-3 factored solvers).
int launch_fused_visco_mfma(const SolverDev<float>& s, float dt, const float* proj_in, float* dofs, float* dofs_a,
                            float* t_int, hipStream_t st);

// FP64 single-run kernels on FP64 MFMA (kernels_plain.hip).
bool plain64_supported(int order, int n_runs, int prec_bytes);
int launch_local_plain64(const SolverDev<double>& s, int64_t first, int64_t n, double dt, double* dofs,
                         double* t_int, hipStream_t st);
int launch_neigh_plain64(const SolverDev<double>& s, int64_t first, int64_t n, const double* t_int, double* dofs,
                         hipStream_t st);
// Single-run kernels with 16 elements per wave on FP64 / FP32 MFMA
// (kernels_ec64.hip): elastic star pattern and rank-3 factored flux solvers;
// the local step writes face projections [nEl][4][3][16] that the neighbour
// step lifts after the face-orientation map.
template <typename T>
struct EcDev {
  const T* pk = nullptr;   // packed A operands (ec64_pack.hpp, ec::Layout)
  const T* rot = nullptr;  // face-orientation maps [4][16][16]
  T* proj = nullptr;       // face projections [nEl][4][3][16]
};
bool ec_supported(int order, int n_runs);
template <typename T>
int launch_local_ec(const SolverDev<T>& s, const EcDev<T>& x, int64_t first, int64_t n, T dt, T* dofs, T* t_int,
                    hipStream_t st);
template <typename T>
int launch_neigh_ec(const SolverDev<T>& s, const EcDev<T>& x, int64_t first, int64_t n, T* dofs, hipStream_t st);
// ... and their viscoelastic form (1 to 3 mechanisms; anelastic rows 9..14 of the tiles)
bool plain64_visco_supported(int order, int n_runs, int prec_bytes, int n_rms);
int launch_local_plain64_visco(const SolverDev<double>& s, int64_t first, int64_t n, double dt, double* dofs,
                               double* dofs_a, double* t_int, hipStream_t st);
int launch_neigh_plain64_visco(const SolverDev<double>& s, int64_t first, int64_t n, const double* t_int,
                               double* dofs, double* dofs_a, hipStream_t st);

// What kernels outside solver.hip need of a solver handle.
struct SolverRef {
  int precision = 64, order = 0, n_runs = 0, n_modes = 0, n_rms = 0;
  int64_t n_el = 0;
  const double* jdet = nullptr;  // [nEl] Jacobian determinants (device), null for caller-given solvers
  const double* mass = nullptr;  // [N] diagonal reference mass matrix (device)
  SolverDev<float> f{};
  SolverDev<double> d{};
};
int solver_ref(const edge_el_solver* s, SolverRef* out);

// The fault phase of the coupled step (fault.hip): sub-face states from
// tInt, friction and net updates, DOF updates of both sides.
int fault_apply(edge_fault* f, double dt, const void* t_int, void* dofs, hipStream_t st);
const edge_el_solver* fault_solver(const edge_fault* f);

// Blocks of the L2-norm reduction (their partial sums are added in block
// order, on the host or on the device).
constexpr unsigned kL2Blocks = 1024;
// Per-block partial sums of the FP64 L2 norms of DOF arrays a (and a - b):
// part[nblk][2][9 R] = (sum |J| M_n (a - b)^2, sum |J| M_n b^2) over each
// block's element range; every entry is written.
int l2_partials(edge_el_solver* s, const void* a, int prec_a, const void* b, int prec_b, double* part, unsigned nblk,
                hipStream_t st);

// Receivers and point sources (points.hip).
template <typename T>
int launch_recv_eval(const SolverDev<T>& s, int64_t n, const int64_t* el, const double* basis, int n_tau,
                     const double* tau, const T* dofs, T* out, hipStream_t st);
bool recv_supported(int order, int n_runs, int prec_bytes);
// ... with attenuation (the anelastic DOFs of the receivers' elements)
template <typename T>
int launch_recv_eval_visco(const SolverDev<T>& s, int64_t n, const int64_t* el, const double* basis, int n_tau,
                           const double* tau, const T* dofs, const T* dofs_a, T* out, hipStream_t st);
bool recv_visco_supported(int order, int n_runs, int prec_bytes, int n_rms);

// Rate-2 local time stepping (lts.hip): tDofs[2] of the listed elements,
// tDofs[1] sums, and the interface faces of the listed elements.
template <typename T>
int launch_lts_half(const SolverDev<T>& s, const int64_t* list, int64_t n, T h, const T* dofs, T* t2,
                    hipStream_t st);
template <typename T>
int launch_lts_sum(const int64_t* list, int64_t n, int sz, int first_ts, const T* t0, T* t1, hipStream_t st);
template <typename T>
int launch_lts_faces(const SolverDev<T>& s, const uint8_t* fa_lts, const int64_t* list, int64_t n, int first_ts,
                     const T* t0, const T* t1, const T* t2, T* dofs, hipStream_t st);
bool lts_supported(int order, int n_runs, int prec_bytes);
// ... of viscoelastic solvers (lts_visco.hip): tDofs[2] from the CK with the
// anelastic terms, and interface faces with the anelastic update
template <typename T>
int launch_lts_half_visco(const SolverDev<T>& s, const int64_t* list, int64_t n, T h, const T* dofs,
                          const T* dofs_a, T* t2, hipStream_t st);
template <typename T>
int launch_lts_faces_visco(const SolverDev<T>& s, const uint8_t* fa_lts, const int64_t* list, int64_t n, int first_ts,
                           const T* t0, const T* t1, const T* t2, T* dofs, T* dofs_a, hipStream_t st);
bool lts_visco_supported(int order, int n_runs, int prec_bytes, int n_rms);

}  // namespace edge
