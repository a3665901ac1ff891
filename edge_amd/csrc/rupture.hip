// Dynamic rupture on the GPU (FP64): per fault face the internal-boundary
// solvers, and per sub-face and fused run the middle state of the welded
// interface, its perturbation by linear slip-weakening friction and the net
// updates of the two adjacent sub-cells.
//
//   face solvers:  InternalBoundarySolvers<TET4,O,3>::init
//                  (src/impl/seismic/solvers/InternalBoundary.hpp:743-923):
//                  T^-1 of the face frame (oriented by the fault frame), the
//                  jump-to-middle-state solver (generated/MiddleStateJump{L,R}3D.inc
//                  restated: ms = q + J (qR - qL) from the waves travelling
//                  into the respective medium), the flux solvers T A_n
//                  (generated/MiddleStateFlux3D.inc restated) scaled by
//                  (area / n_sf) / (volume / n_sc);
//   minus / plus:  RuptureInit.hpp:198-206 (left is the minus side if its
//                  outward normal opposes the fault normal; cs / mu per side);
//   net updates:   InternalBoundary::netUpdates (InternalBoundary.hpp:184-280)
//                  with FrictionLaws<3,R>::linSlipWeak (FrictionLaws.hpp:309-392).
#include <hip/hip_runtime.h>

#include <cmath>
#include <new>
#include <vector>

#include "common.hpp"
#include "face_frame.hpp"
#include "rupture_core.hpp"

struct edge_rupture {
  edge_rupture_config cfg;
  int64_t n_fa = 0;
  double* solv = nullptr;   // [nF][4][81]: T^-1, jump, flux left, flux right
  double* face = nullptr;   // [nF][3]: l_eq_m, cs/mu minus, cs/mu plus
  double* glob = nullptr;   // [3][R]: mus, mud, 1/Dc
  double* state = nullptr;  // [nF][14][n_sf][R] (the C ABI's [nF][n_sf][R][14], per face transposed)
  uint8_t* per = nullptr;   // [nF][n_sf][R]
};

namespace edge {
namespace {

constexpr int kStateVals = kRuptureStateVals;

__host__ __device__ void mm9(const double* a, const double* b, double* c) {
  for (int i = 0; i < 9; i++)
    for (int j = 0; j < 9; j++) {
      double s = 0;
      for (int k = 0; k < 9; k++) s += a[i * 9 + k] * b[k * 9 + j];
      c[i * 9 + j] = s;
    }
}

// middle-state jump solvers of the left (side 0) or right (side 1) medium
__device__ void ms_jump(int side, double lamL, double lamR, double muL, double muR, double rhoL, double rhoR,
                        double J[81]) {
  const double zpL = sqrt((lamL + 2 * muL) * rhoL), zpR = sqrt((lamR + 2 * muR) * rhoR);
  const double zsL = sqrt(muL * rhoL), zsR = sqrt(muR * rhoR);
  const double dp = zpL + zpR, ds = zsL + zsR;
  for (int i = 0; i < 81; i++) J[i] = 0;
  const double sgn = side == 0 ? 1.0 : -1.0;
  const double zpo = side == 0 ? zpL : zpR, zso = side == 0 ? zsL : zsR;  // own impedances
  const double zpx = side == 0 ? zpR : zpL, zsx = side == 0 ? zsR : zsL;  // the other side's
  const double lr = side == 0 ? lamL / (lamL + 2 * muL) : lamR / (lamR + 2 * muR);
  J[0 * 9 + 0] = sgn * zpo / dp;
  J[0 * 9 + 6] = zpL * zpR / dp;
  for (int r = 1; r < 3; r++) {
    J[r * 9 + 0] = lr * J[0];
    J[r * 9 + 6] = lr * J[6];
  }
  J[3 * 9 + 3] = J[5 * 9 + 5] = sgn * zso / ds;
  J[3 * 9 + 7] = J[5 * 9 + 8] = zsL * zsR / ds;
  J[6 * 9 + 0] = 1 / dp;
  J[6 * 9 + 6] = sgn * zpx / dp;
  J[7 * 9 + 3] = J[8 * 9 + 5] = 1 / ds;
  J[7 * 9 + 7] = J[8 * 9 + 8] = sgn * zsx / ds;
}

// flux of a face-aligned middle state in physical coordinates: T A_n
__device__ void ms_flux(const double T[81], double lam, double mu, double rho, double F[81]) {
  double A[81];
  for (int i = 0; i < 81; i++) A[i] = 0;
  A[0 * 9 + 6] = -(lam + 2 * mu);
  A[1 * 9 + 6] = A[2 * 9 + 6] = -lam;
  A[3 * 9 + 7] = A[5 * 9 + 8] = -mu;
  A[6 * 9 + 0] = A[7 * 9 + 3] = A[8 * 9 + 5] = -1 / rho;
  mm9(T, A, F);
}

__global__ void k_rupture_setup(int64_t n_fa, const double* normals, const double* tangents, const double* areas,
                                const double* bnd, const double* mat_l, const double* mat_r, const double* vol_l,
                                const double* vol_r, int n_sf, int n_sc, double* solv, double* face, int* bad) {
  const int64_t fa = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (fa >= n_fa) return;
  double f[3][3];
  for (int d = 0; d < 3; d++) {
    f[0][d] = normals[fa * 3 + d];
    f[1][d] = tangents[fa * 6 + d];
    f[2][d] = tangents[fa * 6 + 3 + d];
  }
  const double sp = bnd[0] * f[0][0] + bnd[1] * f[0][1] + bnd[2] * f[0][2];
  if (!(fabs(sp) > 1e-10)) atomicOr(bad, 1);  // face in the fault plane's normal direction
  const bool pref = sp > 0;
  if (!pref)
    for (int d = 0; d < 3; d++) f[0][d] = -f[0][d];
  // rotation taking the fault normal onto the (oriented) face normal (Rodrigues),
  // applied to the fault's tangents: src/linalg/Geom.hpp:939-989
  {
    double a[3], b[3], na = 0, nb = 0;
    for (int i = 0; i < 3; i++) {
      na += bnd[i] * bnd[i];
      nb += f[0][i] * f[0][i];
    }
    for (int i = 0; i < 3; i++) {
      a[i] = bnd[i] / sqrt(na);
      b[i] = f[0][i] / sqrt(nb);
    }
    const double c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const double sm[9] = {0, -c[2], c[1], c[2], 0, -c[0], -c[1], c[0], 0};
    const double sca = 1.0 / (1.0 + a[0] * b[0] + a[1] * b[1] + a[2] * b[2]);
    double R[9];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        double v = 0;
        for (int k = 0; k < 3; k++) v += sm[i * 3 + k] * sm[k * 3 + j];
        R[i * 3 + j] = v * sca + sm[i * 3 + j] + (i == j ? 1.0 : 0.0);
      }
    for (int k = 1; k < 3; k++)
      for (int i = 0; i < 3; i++) {
        double v = 0;
        for (int j = 0; j < 3; j++) v += R[i * 3 + j] * bnd[k * 3 + j];
        f[k][i] = v;
      }
  }
  double* o = solv + fa * 324;
  double T[81];
  face_trafo(f[0], f[1], f[2], T, o);
  const double rL = mat_l[fa * 3], lL = mat_l[fa * 3 + 1], mL = mat_l[fa * 3 + 2];
  const double rR = mat_r[fa * 3], lR = mat_r[fa * 3 + 1], mR = mat_r[fa * 3 + 2];
  if (pref) {
    ms_jump(0, lL, lR, mL, mR, rL, rR, o + 81);
  } else {  // left and right swap roles; the jump qR - qL changes its sign
    ms_jump(1, lR, lL, mR, mL, rR, rL, o + 81);
    for (int i = 0; i < 81; i++) o[81 + i] = -o[81 + i];
  }
  ms_flux(T, lL, mL, rL, o + 162);
  ms_flux(T, lR, mR, rR, o + 243);
  for (int sd = 0; sd < 2; sd++) {
    double sca = (areas[fa] / n_sf) / ((sd == 0 ? vol_l[fa] : vol_r[fa]) / n_sc);
    if (!pref) sca = -sca;
    for (int i = 0; i < 81; i++) o[(2 + sd) * 81 + i] *= sca;
  }
  const bool l_eq_m = !pref;
  const double cdL = sqrt(1.0 / (mL * rL)), cdR = sqrt(1.0 / (mR * rR));
  face[fa * 3 + 0] = l_eq_m ? 1.0 : 0.0;
  face[fa * 3 + 1] = l_eq_m ? cdL : cdR;  // minus side
  face[fa * 3 + 2] = l_eq_m ? cdR : cdL;  // plus side
}

typedef const __attribute__((address_space(4))) double cdouble;

// One workgroup per face, threads over its (sub-face, run) pairs: the face's
// solvers and data are wave-uniform and come in by scalar loads (the tables
// are read-only here).
__global__ __launch_bounds__(256) void k_rupture_net_updates(int64_t first, int n_sf, int R, double dt,
                                                             const double* solv_, const double* face_,
                                                             const double* glob, const double* dofs_l,
                                                             const double* dofs_r, double* net_l, double* net_r,
                                                             double* state, uint8_t* per_out) {
  const int64_t fa = first + blockIdx.x;
  const cdouble* s = (const cdouble*)(solv_ + fa * 324);
  const cdouble* face = (const cdouble*)(face_ + fa * 3);
  for (int i = threadIdx.x; i < n_sf * R; i += blockDim.x) {
    const int r = i % R, sf = i / R;
    const int64_t base = fa * 9 * n_sf * R + (int64_t)sf * R + r, qs = (int64_t)n_sf * R;
    double qL[9], qR[9], nL[9], nR[9];
    for (int q = 0; q < 9; q++) {
      qL[q] = dofs_l[base + q * qs];
      qR[q] = dofs_r[base + q * qs];
    }
    // the face's state is [14][n_sf][R] on the device: value k of this sub-face and run at st[k * sst]
    double* st = state + (fa * kStateVals * n_sf + sf) * R + r;
    const bool failing = lsw_net_updates(s, face, glob[r], glob[R + r], glob[2 * R + r], dt, qL, qR, st,
                                         (int64_t)n_sf * R, nL, nR);
    per_out[(fa * n_sf + sf) * R + r] = failing ? 1 : 0;
    for (int q = 0; q < 9; q++) {
      net_l[base + q * qs] = nL[q];
      net_r[base + q * qs] = nR[q];
    }
  }
}

}  // namespace
}  // namespace edge

namespace edge {
int rupture_dev(edge_rupture* r, RuptureDev* out) {
  if (!r) return fail(EDGE_ERR_ARG, "rupture handle missing");
  *out = RuptureDev{r->solv, r->face, r->glob, r->state, r->per, r->cfg.n_sub_faces, r->cfg.n_runs, r->n_fa};
  return EDGE_OK;
}
}  // namespace edge

extern "C" int edge_rupture_create(const edge_rupture_config* cfg, int64_t n_faces, const double* normals,
                                   const double* tangents, const double* areas, const double* fault_frame,
                                   const double* mat_l, const double* mat_r, const double* vol_l,
                                   const double* vol_r, const double* glob, const double* state0,
                                   edge_rupture** out) {
  using edge::fail;
  if (out) *out = nullptr;
  if (!cfg || !out || !normals || !tangents || !areas || !fault_frame || !mat_l || !mat_r || !vol_l || !vol_r ||
      !glob || !state0)
    return fail(EDGE_ERR_ARG, "rupture: argument missing");
  if (n_faces <= 0 || cfg->n_sub_faces <= 0 || cfg->n_sub_cells <= 0 || cfg->n_runs <= 0)
    return fail(EDGE_ERR_ARG, "rupture: sizes must be positive");
  for (int k = 0; k < 3; k++) {
    const double* v = fault_frame + 3 * k;
    if (std::abs(std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) - 1.0) > 1e-8)
      return fail(EDGE_ERR_ARG, "rupture: fault frame vectors must be unit vectors");
  }
  const int R = cfg->n_runs;
  for (int r = 0; r < R; r++)
    if (!(glob[r] > 0) || !(glob[R + r] > 0) || !(glob[2 * R + r] > 0))
      return fail(EDGE_ERR_ARG, "rupture: friction coefficients and 1/Dc must be positive");
  for (int64_t fa = 0; fa < n_faces; fa++)
    if (!(areas[fa] > 0) || !(vol_l[fa] > 0) || !(vol_r[fa] > 0) || !(mat_l[fa * 3] > 0) ||
        !(mat_l[fa * 3 + 2] > 0) || !(mat_r[fa * 3] > 0) || !(mat_r[fa * 3 + 2] > 0))
      return fail(EDGE_ERR_ARG, "rupture: non-positive area, volume, density or shear modulus");
  edge_rupture* s = new (std::nothrow) edge_rupture();
  if (!s) return fail(EDGE_ERR_NOMEM, "rupture");
  s->cfg = *cfg;
  s->n_fa = n_faces;
  const size_t ns = (size_t)n_faces * cfg->n_sub_faces * R;
  double *d_nrm = nullptr, *d_tan = nullptr, *d_area = nullptr, *d_bnd = nullptr, *d_ml = nullptr, *d_mr = nullptr,
         *d_vl = nullptr, *d_vr = nullptr;
  int* d_bad = nullptr;
  int bad = 0;
  hipError_t e = hipSuccess;
  auto al = [&](void** p, size_t b) {
    if (e == hipSuccess) e = hipMalloc(p, b);
  };
  auto cp = [&](void* d, const void* h, size_t b) {
    if (e == hipSuccess) e = hipMemcpy(d, h, b, hipMemcpyHostToDevice);
  };
  al((void**)&s->solv, sizeof(double) * 324 * n_faces);
  al((void**)&s->face, sizeof(double) * 3 * n_faces);
  al((void**)&s->glob, sizeof(double) * 3 * R);
  al((void**)&s->state, sizeof(double) * 14 * ns);
  al((void**)&s->per, ns);
  al((void**)&d_nrm, sizeof(double) * 3 * n_faces);
  al((void**)&d_tan, sizeof(double) * 6 * n_faces);
  al((void**)&d_area, sizeof(double) * n_faces);
  al((void**)&d_bnd, sizeof(double) * 9);
  al((void**)&d_ml, sizeof(double) * 3 * n_faces);
  al((void**)&d_mr, sizeof(double) * 3 * n_faces);
  al((void**)&d_vl, sizeof(double) * n_faces);
  al((void**)&d_vr, sizeof(double) * n_faces);
  al((void**)&d_bad, sizeof(int));
  cp(s->glob, glob, sizeof(double) * 3 * R);
  {  // [nF][n_sf][R][14] -> [nF][14][n_sf][R]
    const int64_t per = (int64_t)cfg->n_sub_faces * R;
    std::vector<double> t((size_t)14 * ns);
    for (int64_t f = 0; f < n_faces; f++)
      for (int64_t i = 0; i < per; i++)
        for (int k = 0; k < 14; k++) t[(f * 14 + k) * per + i] = state0[(f * per + i) * 14 + k];
    cp(s->state, t.data(), sizeof(double) * 14 * ns);
  }
  cp(d_nrm, normals, sizeof(double) * 3 * n_faces);
  cp(d_tan, tangents, sizeof(double) * 6 * n_faces);
  cp(d_area, areas, sizeof(double) * n_faces);
  cp(d_bnd, fault_frame, sizeof(double) * 9);
  cp(d_ml, mat_l, sizeof(double) * 3 * n_faces);
  cp(d_mr, mat_r, sizeof(double) * 3 * n_faces);
  cp(d_vl, vol_l, sizeof(double) * n_faces);
  cp(d_vr, vol_r, sizeof(double) * n_faces);
  if (e == hipSuccess) e = hipMemset(d_bad, 0, sizeof(int));
  if (e == hipSuccess) e = hipMemset(s->per, 0, ns);
  if (e == hipSuccess) {
    const int bs = 128;
    hipLaunchKernelGGL(edge::k_rupture_setup, dim3((unsigned)((n_faces + bs - 1) / bs)), dim3(bs), 0, 0, n_faces,
                       d_nrm, d_tan, d_area, d_bnd, d_ml, d_mr, d_vl, d_vr, cfg->n_sub_faces, cfg->n_sub_cells,
                       s->solv, s->face, d_bad);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost);
  for (void* p : {(void*)d_nrm, (void*)d_tan, (void*)d_area, (void*)d_bnd, (void*)d_ml, (void*)d_mr, (void*)d_vl,
                  (void*)d_vr, (void*)d_bad})
    (void)hipFree(p);
  int rc = EDGE_OK;
  if (e != hipSuccess) rc = fail(EDGE_ERR_HIP, std::string("rupture set-up: ") + hipGetErrorString(e));
  else if (bad) rc = fail(EDGE_ERR_MESH, "rupture: a fault face's normal is perpendicular to the fault normal");
  if (rc != EDGE_OK) {
    edge_rupture_destroy(s);
    return rc;
  }
  *out = s;
  return EDGE_OK;
}

extern "C" int edge_rupture_destroy(edge_rupture* s) {
  if (!s) return EDGE_OK;
  (void)hipFree(s->solv);
  (void)hipFree(s->face);
  (void)hipFree(s->glob);
  (void)hipFree(s->state);
  (void)hipFree(s->per);
  delete s;
  return EDGE_OK;
}

extern "C" int edge_rupture_net_updates(edge_rupture* s, int64_t first, int64_t n_faces, double dt,
                                        const double* dofs_l, const double* dofs_r, double* net_l, double* net_r,
                                        void* stream) {
  using edge::fail;
  if (!s) return fail(EDGE_ERR_ARG, "rupture handle missing");
  if (first < 0 || n_faces < 0 || first + n_faces > s->n_fa) return fail(EDGE_ERR_ARG, "face range out of bounds");
  if (!dofs_l || !dofs_r || !net_l || !net_r) return fail(EDGE_ERR_ARG, "device arrays missing");
  if (n_faces == 0) return EDGE_OK;
  hipLaunchKernelGGL(edge::k_rupture_net_updates, dim3((unsigned)n_faces), dim3(256), 0, (hipStream_t)stream, first,
                     s->cfg.n_sub_faces, s->cfg.n_runs, dt, s->solv, s->face, s->glob, dofs_l, dofs_r, net_l, net_r,
                     s->state, s->per);
  EDGE_HIP_TRY(hipGetLastError());
  return EDGE_OK;
}

extern "C" int edge_rupture_get(const edge_rupture* s, double* solvers, double* state, uint8_t* per) {
  using edge::fail;
  if (!s) return fail(EDGE_ERR_ARG, "rupture handle missing");
  const size_t ns = (size_t)s->n_fa * s->cfg.n_sub_faces * s->cfg.n_runs;
  EDGE_HIP_TRY(hipDeviceSynchronize());
  if (solvers) EDGE_HIP_TRY(hipMemcpy(solvers, s->solv, sizeof(double) * 324 * s->n_fa, hipMemcpyDeviceToHost));
  if (state) {  // [nF][14][n_sf][R] -> [nF][n_sf][R][14]
    std::vector<double> t((size_t)14 * ns);
    EDGE_HIP_TRY(hipMemcpy(t.data(), s->state, sizeof(double) * 14 * ns, hipMemcpyDeviceToHost));
    const int64_t per = (int64_t)s->cfg.n_sub_faces * s->cfg.n_runs;
    for (int64_t f = 0; f < s->n_fa; f++)
      for (int64_t i = 0; i < per; i++)
        for (int k = 0; k < 14; k++) state[(f * per + i) * 14 + k] = t[(f * 14 + k) * per + i];
  }
  if (per) EDGE_HIP_TRY(hipMemcpy(per, s->per, ns, hipMemcpyDeviceToHost));
  return EDGE_OK;
}
