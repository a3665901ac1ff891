// A dynamic-rupture fault coupled to the element update: the fault phase of
// the time step, between the local and the neighbour step.
//
// The reference keeps the pieces apart and never calls them from its time
// loop (src/impl/seismic/steps.inc:24-77 runs local, sources, neigh): the
// internal-boundary solvers and net updates on a face's sub-faces
// (src/impl/seismic/solvers/InternalBoundary.hpp:184-280, :743-923), linear
// slip-weakening friction (src/impl/seismic/solvers/FrictionLaws.hpp:309-392)
// and the per-face friction data (src/impl/seismic/setups/RuptureInit.hpp:
// 170-206).  This file defines the call site.  For every fault face (face
// type EDGE_FA_RUPTURE), with L = the face's first element and R the second:
//
//   1. sub-face states: the time average tInt / dt of both elements' traces,
//      averaged over each of the face's n_sf = (2 O - 1)^2 sub-triangles (in
//      L's face coordinates; R's trace through its face and vertex
//      orientation, the neighbour flux matrices' convention,
//      SurfInt.hpp:127-130);
//   2. the net updates of the sub-faces (middle state, friction, fluxes,
//      scaled by dt and by (area / n_sf) / (volume / n_sc));
//   3. the DOF updates: the face polynomial of degree O - 1 whose sub-face
//      means are closest to the net updates (least squares), lifted like a
//      surface flux, into L and R.
//
// The element update gives fault faces no Godunov flux (zero solvers, no
// neighbour term), so this phase supplies their whole flux.  Step 3 inverts
// step 1 on face polynomials, so a fault that never slips gives the welded
// interface's Godunov flux exactly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"
#include "rupture_core.hpp"

struct edge_fault {
  int order = 0, n_runs = 0, precision = 64, n_modes = 0, n_sf = 0;
  int64_t n_el = 0, n_fa = 0;
  const edge_el_solver* solver = nullptr;
  edge_rupture* rup = nullptr;
  int32_t* fel = nullptr;     // [nF][2] elements L, R
  uint8_t* ftab = nullptr;    // [nF][2] table of L (its face, 0..3) and of R (4 + 4 vId + fId)
  int32_t* list = nullptr;    // fault faces in colour order
  std::vector<int64_t> col;   // colour ranges in list (no element twice in one range)
  double* scat = nullptr;     // [16][N][n_sf]: element modes -> sub-face means of the trace
  double* lift = nullptr;     // [16][n_sf][N]: sub-face net updates -> DOF updates
  float* scat32 = nullptr;    // FP32 copies for the MFMA phase (FP32, 16 runs)
  float* lift32 = nullptr;
  size_t lds_mfma = 0;
  std::vector<int64_t> faces; // mesh face ids of the fault faces, in fault order
  std::vector<int32_t> h_fel;
  std::vector<uint8_t> h_ftab;
  size_t lds = 0;
};

namespace edge {
namespace {

// One workgroup per fault face of one colour.  LDS: both elements' tInt
// tiles [9][N][R], then both sides' net updates [9][n_sf][R] (type T).
template <typename T>
__global__ __launch_bounds__(256) void k_fault(const int32_t* __restrict__ list, int64_t first, int N, int NSF,
                                                int R, double dt, const int32_t* __restrict__ fel,
                                                const uint8_t* __restrict__ ftab, const double* __restrict__ scat,
                                                const double* __restrict__ lift, RuptureDev rd,
                                                const T* __restrict__ t_int, T* __restrict__ dofs) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int64_t fi = list[first + blockIdx.x];
  const int64_t eL = fel[2 * fi], eR = fel[2 * fi + 1];
  const int tl = ftab[2 * fi], tr = ftab[2 * fi + 1];
  const int sz = 9 * N * R, nsz = 9 * NSF * R;
  T* sL = reinterpret_cast<T*>(lds);
  T* sR = sL + sz;
  T* nL = sR + sz;
  T* nR = nL + nsz;
  for (int i = threadIdx.x; i < sz; i += blockDim.x) {
    sL[i] = t_int[eL * sz + i];
    sR[i] = t_int[eR * sz + i];
  }
  __syncthreads();
  const double* scL = scat + (size_t)tl * N * NSF;
  const double* scR = scat + (size_t)tr * N * NSF;
  const double* s = rd.solv + fi * 324;
  const double* face = rd.face + fi * 3;
  const double idt = 1.0 / dt;
  for (int i = threadIdx.x; i < NSF * R; i += blockDim.x) {
    const int sf = i / R, r = i % R;
    double qL[9], qR[9];
    for (int q = 0; q < 9; q++) qL[q] = qR[q] = 0;
    for (int n = 0; n < N; n++) {
      const double a = scL[n * NSF + sf], b = scR[n * NSF + sf];
      for (int q = 0; q < 9; q++) {
        qL[q] += a * (double)sL[(q * N + n) * R + r];
        qR[q] += b * (double)sR[(q * N + n) * R + r];
      }
    }
    for (int q = 0; q < 9; q++) {
      qL[q] *= idt;
      qR[q] *= idt;
    }
    double* st = rd.state + (fi * kRuptureStateVals * NSF + sf) * R + r;
    double a[9], b[9];
    const bool failing = lsw_net_updates(s, face, rd.glob[r], rd.glob[R + r], rd.glob[2 * R + r], dt, qL, qR, st,
                                         (int64_t)NSF * R, a, b);
    rd.per[(fi * NSF + sf) * R + r] = failing ? 1 : 0;
    for (int q = 0; q < 9; q++) {
      nL[(q * NSF + sf) * R + r] = (T)a[q];
      nR[(q * NSF + sf) * R + r] = (T)b[q];
    }
  }
  __syncthreads();
  const double* lfL = lift + (size_t)tl * NSF * N;
  const double* lfR = lift + (size_t)tr * NSF * N;
  for (int i = threadIdx.x; i < sz; i += blockDim.x) {
    const int q = i / (N * R), n = (i / R) % N, r = i % R;
    double aL = 0, aR = 0;
    for (int sf = 0; sf < NSF; sf++) {
      aL += lfL[sf * N + n] * (double)nL[(q * NSF + sf) * R + r];
      aR += lfR[sf * N + n] * (double)nR[(q * NSF + sf) * R + r];
    }
    dofs[eL * sz + i] = (T)((double)dofs[eL * sz + i] + aL);
    dofs[eR * sz + i] = (T)((double)dofs[eR * sz + i] + aR);
  }
}

size_t fault_lds(int N, int NSF, int R, int prec_bytes) { return (size_t)prec_bytes * 2 * 9 * R * (N + NSF); }

// FP32 with 16 runs: the sub-face projections and the lifts are products of
// the face's tables with the element tiles, as v_mfma_f32_16x16x4f32 with
// the runs on the N dimension (A lane (g, c): table row 16 t + c, k-step
// row 4 s + g; B lane (g, c): tile row 4 s + g, run c; D register v: row
// 16 t + 4 g + v, run c); the friction in FP64 per (sub-face, run) as above.
// LDS (dynamic): both tInt tiles [2][9][N][16], then both sides' sub-face
// traces, overwritten by their net updates, [2][9][n_sf][16]; the last
// k-step of a product reads rows past the tile as zeros (at order 4, 79,488
// bytes: two workgroups per CU).
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr size_t fault_lds_mfma(int N, int NSF) { return sizeof(float) * 2 * 9 * 16 * (N + NSF); }

// B operand of k-step row k of a [rows][16] tile (zero past the tile)
template <int ROWS>
__device__ __forceinline__ float tile_b(const float* x, int s, int k, int c) {
  if (4 * s + 3 < ROWS) return x[k * 16 + c];  // constant after unrolling
  const float v = x[(k < ROWS ? k : 0) * 16 + c];
  return k < ROWS ? v : 0.f;
}

template <int O>
__global__ __launch_bounds__(256) void k_fault_mfma(const int32_t* __restrict__ list, int64_t first, double dt,
                                                     const int32_t* __restrict__ fel,
                                                     const uint8_t* __restrict__ ftab,
                                                     const float* __restrict__ scat32,
                                                     const float* __restrict__ lift32, RuptureDev rd,
                                                     const float* __restrict__ t_int, float* __restrict__ dofs) {
  constexpr int N = n_modes(O), NSF = n_sub_faces(O), R = 16;
  constexpr int KN = (N + 3) / 4, KSF = (NSF + 3) / 4, RN = N, RSF = NSF;
  constexpr int TN = (N + 15) / 16, TSF = (NSF + 15) / 16;
  extern __shared__ __align__(16) unsigned char lds[];
  float* sX = reinterpret_cast<float*>(lds);  // [2][9][N][16]
  float* nX = sX + 2 * 9 * RN * 16;           // [2][9][NSF][16]
  const int64_t fi = list[first + blockIdx.x];
  const int64_t e[2] = {fel[2 * fi], fel[2 * fi + 1]};
  const int tb[2] = {ftab[2 * fi], ftab[2 * fi + 1]};
  const int tid = threadIdx.x, lane = tid & 63, g = lane >> 4, c = lane & 15;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  for (int i = tid; i < 2 * 9 * N * R; i += 256) {
    const int side = i / (9 * N * R), j = i % (9 * N * R);
    sX[i] = t_int[e[side] * 9 * N * R + j];
  }
  __syncthreads();
  // sub-face traces (time averages): D[sf][r] = sum_n scat[n][sf] tInt[q][n][r] / dt
  const float idt = (float)(1.0 / dt);
  for (int task = wv; task < 2 * 9 * TSF; task += 4) {
    const int side = task / (9 * TSF), q = (task / TSF) % 9, t = task % TSF;
    const float* sc = scat32 + (size_t)tb[side] * N * NSF;
    const float* x = sX + (side * 9 + q) * RN * R;
    const int m = 16 * t + c;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < KN; s++) {
      const int k = 4 * s + g;
      const bool in = k < N && m < NSF;
      const float a = sc[in ? k * NSF + m : 0];
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(in ? a : 0.f, tile_b<N>(x, s, k, c), acc, 0, 0, 0);
    }
    float* o = nX + (side * 9 + q) * RSF * R;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int sf = 16 * t + 4 * g + v;
      if (sf < NSF) o[sf * R + c] = acc[v] * idt;
    }
  }
  __syncthreads();
  const double* s = rd.solv + fi * 324;
  const double* face = rd.face + fi * 3;
  for (int i = tid; i < NSF * R; i += 256) {
    const int sf = i / R, r = i % R;
    double qL[9], qR[9], a[9], b[9];
#pragma unroll
    for (int q = 0; q < 9; q++) {
      qL[q] = nX[(q * RSF + sf) * R + r];
      qR[q] = nX[((9 + q) * RSF + sf) * R + r];
    }
    double* st = rd.state + (fi * kRuptureStateVals * NSF + sf) * R + r;
    const bool failing = lsw_net_updates(s, face, rd.glob[r], rd.glob[R + r], rd.glob[2 * R + r], dt, qL, qR, st,
                                         (int64_t)NSF * R, a, b);
    rd.per[(fi * NSF + sf) * R + r] = failing ? 1 : 0;
#pragma unroll
    for (int q = 0; q < 9; q++) {
      nX[(q * RSF + sf) * R + r] = (float)a[q];
      nX[((9 + q) * RSF + sf) * R + r] = (float)b[q];
    }
  }
  __syncthreads();
  // DOF updates: D[n][r] = sum_sf lift[sf][n] net[q][sf][r]
  for (int task = wv; task < 2 * 9 * TN; task += 4) {
    const int side = task / (9 * TN), q = (task / TN) % 9, t = task % TN;
    const float* lf = lift32 + (size_t)tb[side] * NSF * N;
    const float* x = nX + (side * 9 + q) * RSF * R;
    const int m = 16 * t + c;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < KSF; s++) {
      const int k = 4 * s + g;
      const bool in = k < NSF && m < N;
      const float a = lf[in ? k * N + m : 0];
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(in ? a : 0.f, tile_b<NSF>(x, s, k, c), acc, 0, 0, 0);
    }
    float* d = dofs + (e[side] * 9 + q) * N * R;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int n = 16 * t + 4 * g + v;
      if (n < N) d[n * R + c] += acc[v];
    }
  }
}

template <int O>
void launch_fault_mfma_o(const edge_fault* f, int64_t a, int64_t n, double dt, const RuptureDev& rd,
                         const float* t_int, float* dofs, hipStream_t st) {
  hipLaunchKernelGGL((k_fault_mfma<O>), dim3((unsigned)n), dim3(256), f->lds_mfma, st, f->list, a, dt, f->fel,
                     f->ftab, f->scat32, f->lift32, rd, t_int, dofs);
}

int launch_fault_mfma(const edge_fault* f, int64_t a, int64_t n, double dt, const RuptureDev& rd,
                      const float* t_int, float* dofs, hipStream_t st) {
  switch (f->order) {
    case 2: launch_fault_mfma_o<2>(f, a, n, dt, rd, t_int, dofs, st); break;
    case 3: launch_fault_mfma_o<3>(f, a, n, dt, rd, t_int, dofs, st); break;
    case 4: launch_fault_mfma_o<4>(f, a, n, dt, rd, t_int, dofs, st); break;
    case 5: launch_fault_mfma_o<5>(f, a, n, dt, rd, t_int, dofs, st); break;
    default: return fail(EDGE_ERR_UNSUPPORTED, "fault: order");
  }
  EDGE_HIP_TRY(hipGetLastError());
  return EDGE_OK;
}

hipError_t fault_mfma_allow_lds(int order, int lds) {
  const hipFuncAttribute at = hipFuncAttributeMaxDynamicSharedMemorySize;
  switch (order) {
    case 2: return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fault_mfma<2>), at, lds);
    case 3: return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fault_mfma<3>), at, lds);
    case 4: return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fault_mfma<4>), at, lds);
    case 5: return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fault_mfma<5>), at, lds);
  }
  return hipErrorInvalidValue;
}

// scat[t] = F_t M_F^-1 S^T and lift[t] = c M^-1 F_t G with F_t the face
// trace integrals of table t (fluxL for t < 4, fluxN[t - 4] after), S the
// sub-face means of the face basis and G its least-squares inverse.
void fault_tables(const DgTables& t, int nsf, double lift_scale, std::vector<double>* scat, std::vector<double>* lift) {
  const int N = t.n, NF = t.nf;
  scat->assign((size_t)16 * N * nsf, 0.0);
  lift->assign((size_t)16 * nsf * N, 0.0);
  for (int tb = 0; tb < 16; tb++) {
    const double* F = tb < 4 ? t.fluxL + (size_t)tb * N * NF : t.fluxN + (size_t)(tb - 4) * N * NF;
    for (int n = 0; n < N; n++)
      for (int sf = 0; sf < nsf; sf++) {
        double a = 0, b = 0;
        for (int f = 0; f < NF; f++) {
          a += F[n * NF + f] / t.massF[f] * t.sfMean[sf * NF + f];
          b += F[n * NF + f] * t.sfGather[f * nsf + sf];
        }
        (*scat)[((size_t)tb * N + n) * nsf + sf] = a;
        (*lift)[((size_t)tb * nsf + sf) * N + n] = lift_scale * b / t.mass[n];
      }
  }
}

}  // namespace

int fault_apply(edge_fault* f, double dt, const void* t_int, void* dofs, hipStream_t st) {
  if (!(dt > 0)) return fail(EDGE_ERR_ARG, "fault: dt must be positive");
  RuptureDev rd;
  int rc = rupture_dev(f->rup, &rd);
  if (rc) return rc;
  const int N = f->n_modes, R = f->n_runs;
  for (size_t c = 0; c + 1 < f->col.size(); c++) {
    const int64_t a = f->col[c], n = f->col[c + 1] - a;
    if (n <= 0) continue;
    if (f->scat32) {
      rc = launch_fault_mfma(f, a, n, dt, rd, (const float*)t_int, (float*)dofs, st);
      if (rc) return rc;
      continue;
    }
    if (f->precision == 64)
      hipLaunchKernelGGL(k_fault<double>, dim3((unsigned)n), dim3(256), f->lds, st, f->list, a, N, f->n_sf, R, dt,
                         f->fel, f->ftab, f->scat, f->lift, rd, (const double*)t_int, (double*)dofs);
    else
      hipLaunchKernelGGL(k_fault<float>, dim3((unsigned)n), dim3(256), f->lds, st, f->list, a, N, f->n_sf, R, dt,
                         f->fel, f->ftab, f->scat, f->lift, rd, (const float*)t_int, (float*)dofs);
    EDGE_HIP_TRY(hipGetLastError());
  }
  return EDGE_OK;
}

const edge_el_solver* fault_solver(const edge_fault* f) { return f ? f->solver : nullptr; }

}  // namespace edge

extern "C" int edge_dg_fault_tables(int order, int32_t* n_sub_faces, double* scat, double* lift) {
  edge::DgTables t;
  if (!edge::dg_tables(order, &t)) return edge::fail(EDGE_ERR_UNSUPPORTED, "order not generated");
  if (n_sub_faces) *n_sub_faces = t.nsf;
  if (!scat && !lift) return EDGE_OK;
  const int nsc = (2 * order - 1) * (2 * order - 1) * (2 * order - 1);
  std::vector<double> a, b;
  edge::fault_tables(t, t.nsf, (double)t.nsf / (3.0 * nsc), &a, &b);
  if (scat) std::copy(a.begin(), a.end(), scat);
  if (lift) std::copy(b.begin(), b.end(), lift);
  return EDGE_OK;
}

extern "C" int edge_fault_create(const edge_el_solver* s, const edge_mesh_in* m, const double* rho, const double* lam,
                                 const double* mu, const double* fault_frame, const double* glob,
                                 const double* state0, edge_fault** out) {
  using edge::fail;
  if (out) *out = nullptr;
  edge::SolverRef sr;
  int rc = edge::solver_ref(s, &sr);
  if (rc) return rc;
  if (!out || !m || !m->ve_crds || !m->el_ve || !m->el_fa || !m->fa_el || !m->f_id || !m->v_id || !m->fa_type ||
      !m->normals || !m->tangents || !m->areas || !rho || !lam || !mu || !fault_frame || !glob || !state0)
    return fail(EDGE_ERR_ARG, "fault: argument missing");
  if (m->n_el != sr.n_el) return fail(EDGE_ERR_ARG, "fault: the mesh is not the solver's");
  if (sr.n_rms > 0) return fail(EDGE_ERR_UNSUPPORTED, "fault: elastic solvers only");
  edge::DgTables t;
  if (!edge::dg_tables(sr.order, &t)) return fail(EDGE_ERR_UNSUPPORTED, "fault: order not generated");
  const int N = t.n, nsf = t.nsf, R = sr.n_runs, pb = sr.precision / 8;
  const size_t lds = edge::fault_lds(N, nsf, R, pb);
  if (lds > 160 * 1024) return fail(EDGE_ERR_UNSUPPORTED, "fault: order, runs and precision exceed the LDS");
  // fault faces, their elements, local faces and the right side's orientation
  std::vector<int64_t> faces;
  std::vector<int32_t> fel;
  std::vector<uint8_t> ftab;
  for (int64_t f = 0; f < m->n_fa; f++) {
    if ((m->fa_type[f] & EDGE_FA_RUPTURE) != EDGE_FA_RUPTURE) continue;
    const int64_t eL = m->fa_el[2 * f], eR = m->fa_el[2 * f + 1];
    if (eL < 0 || eR < 0 || eL >= m->n_el || eR >= m->n_el || eL == eR)
      return fail(EDGE_ERR_MESH, "fault: a fault face needs two distinct elements");
    int fl = -1;
    for (int k = 0; k < 4; k++)
      if (m->el_fa[eL * 4 + k] == f) fl = k;
    if (fl < 0) return fail(EDGE_ERR_MESH, "fault: face not among its element's faces");
    const unsigned v = m->v_id[eL * 4 + fl], fi = m->f_id[eL * 4 + fl];
    if (v > 2 || fi > 3 || m->el_fa_el[eL * 4 + fl] != eR) return fail(EDGE_ERR_MESH, "fault: neighbour data");
    faces.push_back(f);
    fel.push_back((int32_t)eL);
    fel.push_back((int32_t)eR);
    ftab.push_back((uint8_t)fl);
    ftab.push_back((uint8_t)(4 + 4 * v + fi));
  }
  const int64_t nF = (int64_t)faces.size();
  if (nF == 0) return fail(EDGE_ERR_MESH, "fault: the mesh has no face of type EDGE_FA_RUPTURE");
  // per-face data of the internal-boundary solvers (InternalBoundary.hpp:743-923)
  std::vector<double> nrm(3 * nF), tan(6 * nF), area(nF), ml(3 * nF), mr(3 * nF), vl(nF), vr(nF);
  auto vol = [&](int64_t e) {
    const double* p[4];
    for (int k = 0; k < 4; k++) p[k] = m->ve_crds + 3 * m->el_ve[e * 4 + k];
    double a[3], b[3], c[3];
    for (int d = 0; d < 3; d++) {
      a[d] = p[1][d] - p[0][d];
      b[d] = p[2][d] - p[0][d];
      c[d] = p[3][d] - p[0][d];
    }
    return std::abs(a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                    a[2] * (b[0] * c[1] - b[1] * c[0])) / 6.0;
  };
  for (int64_t i = 0; i < nF; i++) {
    const int64_t f = faces[i], eL = fel[2 * i], eR = fel[2 * i + 1];
    for (int d = 0; d < 3; d++) nrm[3 * i + d] = m->normals[3 * f + d];
    for (int d = 0; d < 6; d++) tan[6 * i + d] = m->tangents[6 * f + d];
    area[i] = m->areas[f];
    ml[3 * i] = rho[eL];
    ml[3 * i + 1] = lam[eL];
    ml[3 * i + 2] = mu[eL];
    mr[3 * i] = rho[eR];
    mr[3 * i + 1] = lam[eR];
    mr[3 * i + 2] = mu[eR];
    vl[i] = vol(eL);
    vr[i] = vol(eR);
  }
  // sub-cells per element: (2 O - 1)^3 (CE_N_SUB_CELLS, src/sc/cexprs.inc:66-77)
  const int nsc = (2 * sr.order - 1) * (2 * sr.order - 1) * (2 * sr.order - 1);
  edge_rupture_config rc_cfg{nsf, nsc, R};
  edge_fault* F = new (std::nothrow) edge_fault();
  if (!F) return fail(EDGE_ERR_NOMEM, "fault");
  F->order = sr.order;
  F->n_runs = R;
  F->precision = sr.precision;
  F->n_modes = N;
  F->n_sf = nsf;
  F->n_el = sr.n_el;
  F->n_fa = nF;
  F->solver = s;
  F->faces = faces;
  F->h_fel = fel;
  F->h_ftab = ftab;
  F->lds = lds;
  rc = edge_rupture_create(&rc_cfg, nF, nrm.data(), tan.data(), area.data(), fault_frame, ml.data(), mr.data(),
                           vl.data(), vr.data(), glob, state0, &F->rup);
  if (rc) {
    edge_fault_destroy(F);
    return rc;
  }
  // colours: no element in two faces of one launch
  std::vector<int> fcol(nF);
  std::vector<uint64_t> used(sr.n_el, 0);
  int ncol = 0;
  for (int64_t i = 0; i < nF; i++) {
    const uint64_t u = used[fel[2 * i]] | used[fel[2 * i + 1]];
    int c = 0;
    while (c < 64 && (u >> c & 1)) c++;
    if (c == 64) {
      edge_fault_destroy(F);
      return fail(EDGE_ERR_MESH, "fault: face colouring failed (an element with too many fault faces)");
    }
    fcol[i] = c;
    used[fel[2 * i]] |= 1ull << c;
    used[fel[2 * i + 1]] |= 1ull << c;
    ncol = std::max(ncol, c + 1);
  }
  std::vector<int32_t> list;
  F->col.push_back(0);
  for (int c = 0; c < ncol; c++) {
    for (int64_t i = 0; i < nF; i++)
      if (fcol[i] == c) list.push_back((int32_t)i);
    F->col.push_back((int64_t)list.size());
  }
  // the lift's scale: net updates are sub-cell-average increments, scaled by
  // (area / n_sf) / (volume / n_sc); the DG update of a face flux f is
  // (area / (3 volume)) M^-1 int_ref-face phi f (|J| = 6 volume, face
  // Jacobian 2 area)
  std::vector<double> scat, lift;
  edge::fault_tables(t, nsf, (double)nsf / (3.0 * nsc), &scat, &lift);
  hipError_t e = hipSuccess;
  auto up = [&](void** d, const void* h, size_t b) {
    if (e == hipSuccess) e = hipMalloc(d, b);
    if (e == hipSuccess) e = hipMemcpy(*d, h, b, hipMemcpyHostToDevice);
  };
  up((void**)&F->fel, fel.data(), sizeof(int32_t) * fel.size());
  up((void**)&F->ftab, ftab.data(), ftab.size());
  up((void**)&F->list, list.data(), sizeof(int32_t) * list.size());
  up((void**)&F->scat, scat.data(), sizeof(double) * scat.size());
  up((void**)&F->lift, lift.data(), sizeof(double) * lift.size());
  if (F->precision == 32 && R == 16) {  // the MFMA fault phase
    const std::vector<float> s32(scat.begin(), scat.end()), l32(lift.begin(), lift.end());
    up((void**)&F->scat32, s32.data(), sizeof(float) * s32.size());
    up((void**)&F->lift32, l32.data(), sizeof(float) * l32.size());
    F->lds_mfma = edge::fault_lds_mfma(N, nsf);
    if (e == hipSuccess && F->lds_mfma > 64 * 1024) e = edge::fault_mfma_allow_lds(F->order, (int)F->lds_mfma);
  }
  if (e == hipSuccess && lds > 64 * 1024) {
    e = F->precision == 64 ? hipFuncSetAttribute(reinterpret_cast<const void*>(&edge::k_fault<double>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)
                           : hipFuncSetAttribute(reinterpret_cast<const void*>(&edge::k_fault<float>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  }
  if (e != hipSuccess) {
    edge_fault_destroy(F);
    return fail(EDGE_ERR_HIP, std::string("fault set-up: ") + hipGetErrorString(e));
  }
  *out = F;
  return EDGE_OK;
}

extern "C" int edge_fault_destroy(edge_fault* f) {
  if (!f) return EDGE_OK;
  edge_rupture_destroy(f->rup);
  for (void* p : {(void*)f->fel, (void*)f->ftab, (void*)f->list, (void*)f->scat, (void*)f->lift, (void*)f->scat32,
                  (void*)f->lift32})
    (void)hipFree(p);
  delete f;
  return EDGE_OK;
}

extern "C" int edge_fault_apply(edge_fault* f, const edge_el_solver* s, double dt, const void* t_int, void* dofs,
                                void* stream) {
  using edge::fail;
  if (!f || !s) return fail(EDGE_ERR_ARG, "fault or solver handle missing");
  if (s != f->solver) return fail(EDGE_ERR_ARG, "fault: built for another solver");
  if (!t_int || !dofs) return fail(EDGE_ERR_ARG, "device arrays missing");
  return edge::fault_apply(f, dt, t_int, dofs, (hipStream_t)stream);
}

extern "C" int edge_fault_info(const edge_fault* f, int64_t* n_faces, int32_t* n_sub_faces, int64_t* faces,
                               int64_t* elements, int32_t* local_faces) {
  if (!f) return edge::fail(EDGE_ERR_ARG, "fault handle missing");
  if (n_faces) *n_faces = f->n_fa;
  if (n_sub_faces) *n_sub_faces = f->n_sf;
  for (int64_t i = 0; i < f->n_fa; i++) {
    if (faces) faces[i] = f->faces[i];
    if (elements) {
      elements[2 * i] = f->h_fel[2 * i];
      elements[2 * i + 1] = f->h_fel[2 * i + 1];
    }
    if (local_faces) local_faces[i] = f->h_ftab[2 * i];
  }
  return EDGE_OK;
}

extern "C" int edge_fault_get(const edge_fault* f, double* state, uint8_t* slipping) {
  if (!f) return edge::fail(EDGE_ERR_ARG, "fault handle missing");
  return edge_rupture_get(f->rup, nullptr, state, slipping);
}
