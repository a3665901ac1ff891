// FP64 set-up on the GPU: element Jacobians and star matrices, and the
// per-face Godunov flux solvers of both adjacent elements.  Results are
// written in the solver's storage precision T.
//
// Physics follows the reference's setup:
//  * star matrices S_d = sum_k A_k (J^-1)[k][d]
//    (src/impl/seismic/solvers/AderDgInit.hpp:129-174,
//     src/impl/seismic/setups/Elasticity.cpp:167-252);
//  * flux solvers fsL = T FlMid T^-1, fsR = T FrMid [mirror] T^-1 in the
//    face frame (n, s, t), the right element with (-n, s, t), both scaled by
//    -2 area / |J| (AderDgInit.hpp:394-665, Elasticity.cpp:372-441).
// FlMid / FrMid are derived here from the welded-interface Riemann problem.
#include <hip/hip_runtime.h>

#include "face_frame.hpp"
#include "setup.hpp"

namespace edge {
namespace {

__device__ void jac_inv(const double* ve, double* det, double jinv[9]) {
  double a[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) a[i * 3 + j] = ve[(i + 1) * 3 + j] - ve[j];
  const double d = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                   a[2] * (a[3] * a[7] - a[4] * a[6]);
  *det = d;
  const double id = 1.0 / d;
  jinv[0] = (a[4] * a[8] - a[5] * a[7]) * id;
  jinv[1] = (a[2] * a[7] - a[1] * a[8]) * id;
  jinv[2] = (a[1] * a[5] - a[2] * a[4]) * id;
  jinv[3] = (a[5] * a[6] - a[3] * a[8]) * id;
  jinv[4] = (a[0] * a[8] - a[2] * a[6]) * id;
  jinv[5] = (a[2] * a[3] - a[0] * a[5]) * id;
  jinv[6] = (a[3] * a[7] - a[4] * a[6]) * id;
  jinv[7] = (a[1] * a[6] - a[0] * a[7]) * id;
  jinv[8] = (a[0] * a[4] - a[1] * a[3]) * id;
}

__device__ void load_ve(const double* crds, const int64_t* el_ve, int64_t el, double ve[12]) {
  for (int v = 0; v < 4; v++) {
    const int64_t id = el_ve[el * 4 + v];
    for (int d = 0; d < 3; d++) ve[v * 3 + d] = crds[id * 3 + d];
  }
}

template <typename T>
__global__ void k_star(int64_t n_el, const double* crds, const int64_t* el_ve, const double* rho,
                       const double* lam, const double* mu, T* star, double* jdet, T* star_a, int* bad) {
  const int64_t el = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (el >= n_el) return;
  double ve[12], det, ji[9];
  load_ve(crds, el_ve, el, ve);
  jac_inv(ve, &det, ji);
  jdet[el] = det;
  if (!(det > 0)) atomicOr(bad, 1);
  const double lp = -lam[el] - 2 * mu[el], l = -lam[el], m = -mu[el], ir = -1.0 / rho[el];
  double s[243];
  for (int i = 0; i < 243; i++) s[i] = 0;
  // nonzeros (row, col) of the Jacobians A (x), B (y), C (z) and their values
  const int rc[3][8][2] = {{{0, 6}, {1, 6}, {2, 6}, {3, 7}, {5, 8}, {6, 0}, {7, 3}, {8, 5}},
                           {{0, 7}, {1, 7}, {2, 7}, {3, 6}, {4, 8}, {6, 3}, {7, 1}, {8, 4}},
                           {{0, 8}, {1, 8}, {2, 8}, {4, 7}, {5, 6}, {6, 5}, {7, 4}, {8, 2}}};
  const double va[3][8] = {{lp, l, l, m, m, ir, ir, ir}, {l, lp, l, m, m, ir, ir, ir}, {l, l, lp, m, m, ir, ir, ir}};
  for (int d = 0; d < 3; d++)
    for (int k = 0; k < 3; k++)
      for (int e = 0; e < 8; e++) s[d * 81 + rc[k][e][0] * 9 + rc[k][e][1]] += va[k][e] * ji[k * 3 + d];
  for (int i = 0; i < 243; i++) star[el * 243 + i] = (T)s[i];
  if (star_a) {
    // anelastic star matrices [3][6][3]: memory variables from the velocity
    // gradients, the Jacobians of src/impl/seismic/setups/ViscoElasticity.cpp:233-316
    double a[54];
    for (int i = 0; i < 54; i++) a[i] = 0;
    for (int d = 0; d < 3; d++) {
      double* o = a + d * 18;
      const double j0 = ji[0 * 3 + d], j1 = ji[1 * 3 + d], j2 = ji[2 * 3 + d];
      o[0 * 3 + 0] -= j0;
      o[3 * 3 + 1] -= 0.5 * j0;
      o[5 * 3 + 2] -= 0.5 * j0;
      o[1 * 3 + 1] -= j1;
      o[3 * 3 + 0] -= 0.5 * j1;
      o[4 * 3 + 2] -= 0.5 * j1;
      o[2 * 3 + 2] -= j2;
      o[4 * 3 + 1] -= 0.5 * j2;
      o[5 * 3 + 0] -= 0.5 * j2;
    }
    for (int i = 0; i < 54; i++) star_a[el * 54 + i] = (T)a[i];
  }
}

// derivative of the left element's flux A_L q* w.r.t. the left (fl) and the
// right (fr) face-aligned states; q* is the welded-interface middle state
__device__ void fs_mid(double rhoL, double rhoR, double lamL, double lamR, double muL, double muR,
                       double fl[81], double fr[81]) {
  for (int i = 0; i < 81; i++) fl[i] = fr[i] = 0;
  const double zpL = sqrt((lamL + 2 * muL) * rhoL), zpR = sqrt((lamR + 2 * muR) * rhoR);
  const double zsL = sqrt(muL * rhoL), zsR = sqrt(muR * rhoR);
  const double ip = 1.0 / (zpL + zpR);
  const double cu[3] = {-(lamL + 2 * muL), -lamL, -lamL};
  for (int r = 0; r < 3; r++) {
    fl[r * 9 + 0] = -cu[r] * ip;
    fl[r * 9 + 6] = cu[r] * zpL * ip;
    fr[r * 9 + 0] = cu[r] * ip;
    fr[r * 9 + 6] = cu[r] * zpR * ip;
  }
  fl[6 * 9 + 0] = -zpR * ip / rhoL;
  fl[6 * 9 + 6] = zpL * zpR * ip / rhoL;
  fr[6 * 9 + 0] = -zpL * ip / rhoL;
  fr[6 * 9 + 6] = -zpL * zpR * ip / rhoL;
  const double is = 1.0 / (zsL + zsR);
  const int pr[2][2] = {{3, 7}, {5, 8}};
  for (int p = 0; p < 2; p++) {
    const int si = pr[p][0], vv = pr[p][1];
    fl[si * 9 + si] = muL * is;
    fl[si * 9 + vv] = -muL * zsL * is;
    fr[si * 9 + si] = -muL * is;
    fr[si * 9 + vv] = -muL * zsR * is;
    fl[vv * 9 + si] = -zsR * is / rhoL;
    fl[vv * 9 + vv] = zsL * zsR * is / rhoL;
    fr[vv * 9 + si] = -zsL * is / rhoL;
    fr[vv * 9 + vv] = -zsL * zsR * is / rhoL;
  }
}

__device__ void mm9(const double* a, const double* b, double* c) {
  for (int i = 0; i < 9; i++)
    for (int j = 0; j < 9; j++) {
      double s = 0;
      for (int k = 0; k < 9; k++) s += a[i * 9 + k] * b[k * 9 + j];
      c[i * 9 + j] = s;
    }
}

__device__ void face_solvers(double rL, double rR, double lL, double lR, double mL, double mR, const double n[3],
                             const double s[3], const double t[3], bool free_surface, double fsL[81],
                             double fsR[81]) {
  double T[81], Ti[81], fl[81], fr[81], tmp[81];
  face_trafo(n, s, t, T, Ti);
  fs_mid(rL, rR, lL, lR, mL, mR, fl, fr);
  mm9(T, fl, tmp);
  mm9(tmp, Ti, fsL);
  mm9(T, fr, tmp);
  if (free_surface)  // mirror the face-aligned normal tractions (columns 0, 3, 5)
    for (int r = 0; r < 9; r++) {
      tmp[r * 9 + 0] = -tmp[r * 9 + 0];
      tmp[r * 9 + 3] = -tmp[r * 9 + 3];
      tmp[r * 9 + 5] = -tmp[r * 9 + 5];
    }
  mm9(tmp, Ti, fsR);
}

// Anelastic flux solvers [6][9] of both sides: the middle state's anelastic
// part in face coordinates (memory variables of sigma_nn / sigma_ns /
// sigma_nt driven by the P and S waves, src/impl/seismic/setups/
// ViscoElasticity.cpp:376-441), rotated with the stress block T_A of T and
// T^-1, the right state mirrored on a free surface (ViscoElasticity.cpp:
// 505-570, solvers/AderDgInit.hpp:288-370).
__device__ void anelastic_solvers(double rL, double rR, double lL, double lR, double mL, double mR, const double n[3],
                                  const double s[3], const double t[3], bool free_surface, double fsAl[54],
                                  double fsAr[54]) {
  double T[81], Ti[81], mid[2][54], tmp[54];
  face_trafo(n, s, t, T, Ti);
  const double cpL = sqrt((lL + 2 * mL) / rL), cpR = sqrt((lR + 2 * mR) / rR);
  const double csL = sqrt(mL / rL), csR = sqrt(mR / rR);
  for (int i = 0; i < 54; i++) mid[0][i] = mid[1][i] = 0;
  const double dp = lL * cpR + lR * cpL + 2 * mL * cpR + 2 * mR * cpL;
  const double ds = 2 * mL * csR + 2 * mR * csL;
  mid[0][0 * 9 + 0] = cpL * cpR / dp;
  mid[0][0 * 9 + 6] = -cpR * (lL + 2 * mL) / dp;
  mid[0][3 * 9 + 3] = mid[0][5 * 9 + 5] = csL * csR / ds;
  mid[0][3 * 9 + 7] = mid[0][5 * 9 + 8] = -mL * csR / ds;
  mid[1][0 * 9 + 0] = -cpL * cpR / dp;
  mid[1][0 * 9 + 6] = -cpL * (lR + 2 * mR) / dp;
  mid[1][3 * 9 + 3] = mid[1][5 * 9 + 5] = -csL * csR / ds;
  mid[1][3 * 9 + 7] = mid[1][5 * 9 + 8] = -mR * csL / ds;
  for (int sd = 0; sd < 2; sd++) {
    for (int r = 0; r < 6; r++)
      for (int c = 0; c < 9; c++) {
        double v = 0;
        for (int k = 0; k < 6; k++) v += T[r * 9 + k] * mid[sd][k * 9 + c];
        tmp[r * 9 + c] = v;
      }
    if (sd == 1 && free_surface)
      for (int r = 0; r < 6; r++) {
        tmp[r * 9 + 0] = -tmp[r * 9 + 0];
        tmp[r * 9 + 3] = -tmp[r * 9 + 3];
        tmp[r * 9 + 5] = -tmp[r * 9 + 5];
      }
    double* o = sd == 0 ? fsAl : fsAr;
    for (int r = 0; r < 6; r++)
      for (int c = 0; c < 9; c++) {
        double v = 0;
        for (int k = 0; k < 9; k++) v += tmp[r * 9 + k] * Ti[k * 9 + c];
        o[r * 9 + c] = v;
      }
  }
}

// Rank-3 factors of a flux solver F (9 x 9): F = U V^T with U = [u_0 u_1 u_2]
// orthonormal (Gram-Schmidt with column pivoting; each new column is
// re-orthogonalised against the earlier ones) and V^T = U^T F.  Writes U
// row-major [9][3] then V^T row-major [3][9].  Returns false when the
// remainder F - U U^T F exceeds 1e-10 of F (Frobenius norms).  The elastic
// Godunov solvers have rank 3: the flux A_L q* depends on either side's state
// only through its three waves travelling towards the face.
template <typename T>
__device__ bool rank3_factors(const double* F, T* out) {
  double W[81], U[27];
  double fn2 = 0;
  for (int i = 0; i < 81; i++) {
    W[i] = F[i];
    fn2 += F[i] * F[i];
  }
  for (int i = 0; i < 27; i++) U[i] = 0;
  for (int j = 0; j < 3; j++) {
    int best = -1;
    double bn = 0;
    for (int c = 0; c < 9; c++) {
      double nr = 0;
      for (int r = 0; r < 9; r++) nr += W[r * 9 + c] * W[r * 9 + c];
      if (nr > bn) {
        bn = nr;
        best = c;
      }
    }
    if (best < 0 || !(bn > 1e-28 * fn2)) break;  // rank below 3: the remaining columns stay zero
    double u[9];
    for (int r = 0; r < 9; r++) u[r] = W[r * 9 + best];
    for (int i = 0; i < j; i++) {
      double d = 0;
      for (int r = 0; r < 9; r++) d += U[r * 3 + i] * u[r];
      for (int r = 0; r < 9; r++) u[r] -= d * U[r * 3 + i];
    }
    double nu = 0;
    for (int r = 0; r < 9; r++) nu += u[r] * u[r];
    const double inv = 1.0 / sqrt(nu);
    for (int r = 0; r < 9; r++) U[r * 3 + j] = u[r] * inv;
    for (int c = 0; c < 9; c++) {
      double d = 0;
      for (int r = 0; r < 9; r++) d += U[r * 3 + j] * W[r * 9 + c];
      for (int r = 0; r < 9; r++) W[r * 9 + c] -= d * U[r * 3 + j];
    }
  }
  double rem = 0;
  for (int i = 0; i < 81; i++) rem += W[i] * W[i];
  for (int i = 0; i < 27; i++) out[i] = (T)U[i];
  for (int j = 0; j < 3; j++)
    for (int p = 0; p < 9; p++) {
      double v = 0;
      for (int q = 0; q < 9; q++) v += U[q * 3 + j] * F[q * 9 + p];
      out[27 + j * 9 + p] = (T)v;
    }
  return rem <= 1e-20 * fn2;
}

// Factor U_A [6][3] of an anelastic flux solver FA [6][9] over the rows V^T
// [3][9] of the elastic solver's rank-3 factors: FA = U_A V^T, by least
// squares through the Gram matrix of V^T's rows.  The anelastic middle state
// reads each side's state through the same three wave combinations as the
// elastic one (sigma_nn - Z_p v_n, sigma_ns - Z_s v_s, sigma_nt - Z_s v_t in
// face coordinates, src/impl/seismic/setups/ViscoElasticity.cpp:376-441),
// so FA lies in that row space.  Returns false when the remainder exceeds
// 1e-10 of FA (Frobenius norms) or the Gram matrix is singular.
template <typename T>
__device__ bool coupled_factor(const double* FA, const double* vt, T* out) {
  double G[9];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double v = 0;
      for (int p = 0; p < 9; p++) v += vt[i * 9 + p] * vt[j * 9 + p];
      G[i * 3 + j] = v;
    }
  const double det = G[0] * (G[4] * G[8] - G[5] * G[7]) - G[1] * (G[3] * G[8] - G[5] * G[6]) +
                     G[2] * (G[3] * G[7] - G[4] * G[6]);
  const double tr = G[0] + G[4] + G[8];
  if (!(det > 1e-24 * tr * tr * tr)) return false;
  double Gi[9];
  Gi[0] = (G[4] * G[8] - G[5] * G[7]) / det;
  Gi[1] = (G[2] * G[7] - G[1] * G[8]) / det;
  Gi[2] = (G[1] * G[5] - G[2] * G[4]) / det;
  Gi[3] = (G[5] * G[6] - G[3] * G[8]) / det;
  Gi[4] = (G[0] * G[8] - G[2] * G[6]) / det;
  Gi[5] = (G[2] * G[3] - G[0] * G[5]) / det;
  Gi[6] = (G[3] * G[7] - G[4] * G[6]) / det;
  Gi[7] = (G[1] * G[6] - G[0] * G[7]) / det;
  Gi[8] = (G[0] * G[4] - G[1] * G[3]) / det;
  double U[18], fn2 = 0, rem = 0;
  for (int q = 0; q < 6; q++) {
    double h[3];
    for (int j = 0; j < 3; j++) {
      double v = 0;
      for (int p = 0; p < 9; p++) v += FA[q * 9 + p] * vt[j * 9 + p];
      h[j] = v;
    }
    for (int j = 0; j < 3; j++) U[q * 3 + j] = h[0] * Gi[0 * 3 + j] + h[1] * Gi[1 * 3 + j] + h[2] * Gi[2 * 3 + j];
    for (int p = 0; p < 9; p++) {
      const double d = FA[q * 9 + p] - (U[q * 3] * vt[p] + U[q * 3 + 1] * vt[9 + p] + U[q * 3 + 2] * vt[18 + p]);
      rem += d * d;
      fn2 += FA[q * 9 + p] * FA[q * 9 + p];
    }
  }
  for (int i = 0; i < 18; i++) out[i] = (T)U[i];
  return rem <= 1e-20 * fn2;
}

template <typename T>
__global__ void k_fs_factor(int64_t n_mats, const double* fs, T* fsf, int* bad) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_mats) return;
  if (!rank3_factors(fs + i * 81, fsf + i * 54)) atomicOr(bad, 8);
}

template <typename T>
__global__ void k_flux_solvers(int64_t n_fa, int64_t n_el, const int64_t* fa_el, const int64_t* el_fa,
                               const int32_t* fa_type, const double* normals, const double* tangents,
                               const double* areas, const double* jdet, const double* rho, const double* lam,
                               const double* mu, T* fs, T* fsf, T* fs_a, T* fsfa, int* bad) {
  const int64_t fa = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (fa >= n_fa) return;
  const int64_t eL = fa_el[fa * 2], eR = fa_el[fa * 2 + 1];
  if (eL < 0 || eL >= n_el || eR >= n_el) {
    atomicOr(bad, 2);
    return;
  }
  int fL = -1, fR = -1;
  for (int f = 0; f < 4; f++) {
    if (el_fa[eL * 4 + f] == fa) fL = f;
    if (eR >= 0 && el_fa[eR * 4 + f] == fa) fR = f;
  }
  if (fL < 0 || (eR >= 0 && fR < 0)) {
    atomicOr(bad, 2);
    return;
  }
  if (!(areas[fa] > 0)) atomicOr(bad, 4);
  const bool fsurf = (fa_type[fa] & EDGE_FA_FREE_SURFACE) == EDGE_FA_FREE_SURFACE;
  const double rL = rho[eL], lL = lam[eL], mL = mu[eL];
  const double rR = eR >= 0 ? rho[eR] : rL, lR = eR >= 0 ? lam[eR] : lL, mR = eR >= 0 ? mu[eR] : mL;
  const double* n = normals + fa * 3;
  const double* s = tangents + fa * 6;
  const double* t = tangents + fa * 6 + 3;
  double a[81], b[81], al[54], ar[54];
  // both sides' solvers of element e's face f, scaled by sc; with the
  // anelastic solvers al / ar (unscaled) when fs_a is given
  auto emit = [&](int64_t e, int f, double sc) {
    for (int i = 0; i < 81; i++) {
      a[i] *= sc;
      b[i] *= sc;
      fs[(e * 4 + f) * 81 + i] = (T)a[i];
      fs[((n_el + e) * 4 + f) * 81 + i] = (T)b[i];
    }
    double fac[2][54];
    if (fsf) {
      bool ok = rank3_factors(a, fac[0]);
      ok = rank3_factors(b, fac[1]) && ok;
      for (int i = 0; i < 54; i++) {
        fsf[(e * 4 + f) * 54 + i] = (T)fac[0][i];
        fsf[((n_el + e) * 4 + f) * 54 + i] = (T)fac[1][i];
      }
      if (!ok) atomicOr(bad, 8);
    }
    if (fs_a) {
      for (int i = 0; i < 54; i++) {
        al[i] *= sc;
        ar[i] *= sc;
        fs_a[(e * 4 + f) * 54 + i] = (T)al[i];
        fs_a[((n_el + e) * 4 + f) * 54 + i] = (T)ar[i];
      }
      if (fsf && fsfa) {
        bool ok = coupled_factor(al, fac[0] + 27, fsfa + (e * 4 + f) * 18);
        ok = coupled_factor(ar, fac[1] + 27, fsfa + ((n_el + e) * 4 + f) * 18) && ok;
        if (!ok) atomicOr(bad, 8);
      }
    }
  };
  face_solvers(rL, rR, lL, lR, mL, mR, n, s, t, fsurf, a, b);
  if (fs_a) anelastic_solvers(rL, rR, lL, lR, mL, mR, n, s, t, fsurf, al, ar);
  emit(eL, fL, -2.0 * areas[fa] / jdet[eL]);
  if (eR >= 0) {
    const double nm[3] = {-n[0], -n[1], -n[2]};
    face_solvers(rR, rL, lR, lL, mR, mL, nm, s, t, false, a, b);
    if (fs_a) anelastic_solvers(rR, rL, lR, lL, mR, mL, nm, s, t, false, al, ar);
    emit(eR, fR, -2.0 * areas[fa] / jdet[eR]);
  }
}

}  // namespace

template <typename T>
int build_setup(const SetupInput& in, T* d_star, T* d_fs, double* d_jdet, T* d_fsf, bool* rank3, T* d_star_a,
                T* d_fs_a, T* d_fsfa, hipStream_t st) {
  double *d_crds = nullptr, *d_rho = nullptr, *d_lam = nullptr, *d_mu = nullptr, *d_nrm = nullptr,
         *d_tan = nullptr, *d_area = nullptr;
  int64_t *d_elve = nullptr, *d_fael = nullptr, *d_elfa = nullptr;
  int32_t* d_type = nullptr;
  int* d_bad = nullptr;
  int rc = EDGE_OK;
  const int64_t ne = in.n_el, nf = in.n_fa;
  auto cleanup = [&]() {
    (void)hipFree(d_crds); (void)hipFree(d_rho); (void)hipFree(d_lam); (void)hipFree(d_mu); (void)hipFree(d_nrm); (void)hipFree(d_tan);
    (void)hipFree(d_area); (void)hipFree(d_elve); (void)hipFree(d_fael); (void)hipFree(d_elfa); (void)hipFree(d_type);
    (void)hipFree(d_bad);
  };
#define EDGE_SETUP_TRY(x)                                                                         \
  do {                                                                                            \
    hipError_t e_ = (x);                                                                          \
    if (e_ != hipSuccess) {                                                                       \
      cleanup();                                                                                  \
      return fail(EDGE_ERR_HIP, std::string("setup: ") + #x + ": " + hipGetErrorString(e_));      \
    }                                                                                             \
  } while (0)
  EDGE_SETUP_TRY(hipMalloc(&d_crds, sizeof(double) * 3 * in.n_ve));
  EDGE_SETUP_TRY(hipMalloc(&d_elve, sizeof(int64_t) * 4 * ne));
  EDGE_SETUP_TRY(hipMalloc(&d_rho, sizeof(double) * ne));
  EDGE_SETUP_TRY(hipMalloc(&d_lam, sizeof(double) * ne));
  EDGE_SETUP_TRY(hipMalloc(&d_mu, sizeof(double) * ne));
  EDGE_SETUP_TRY(hipMalloc(&d_fael, sizeof(int64_t) * 2 * nf));
  EDGE_SETUP_TRY(hipMalloc(&d_elfa, sizeof(int64_t) * 4 * ne));
  EDGE_SETUP_TRY(hipMalloc(&d_type, sizeof(int32_t) * nf));
  EDGE_SETUP_TRY(hipMalloc(&d_nrm, sizeof(double) * 3 * nf));
  EDGE_SETUP_TRY(hipMalloc(&d_tan, sizeof(double) * 6 * nf));
  EDGE_SETUP_TRY(hipMalloc(&d_area, sizeof(double) * nf));
  EDGE_SETUP_TRY(hipMalloc(&d_bad, sizeof(int)));
  EDGE_SETUP_TRY(hipMemcpyAsync(d_crds, in.ve_crds, sizeof(double) * 3 * in.n_ve, hipMemcpyHostToDevice, st));
  EDGE_SETUP_TRY(hipMemcpyAsync(d_elve, in.el_ve, sizeof(int64_t) * 4 * ne, hipMemcpyHostToDevice, st));
  EDGE_SETUP_TRY(hipMemcpyAsync(d_rho, in.rho, sizeof(double) * ne, hipMemcpyHostToDevice, st));
  EDGE_SETUP_TRY(hipMemcpyAsync(d_lam, in.lam, sizeof(double) * ne, hipMemcpyHostToDevice, st));
  EDGE_SETUP_TRY(hipMemcpyAsync(d_mu, in.mu, sizeof(double) * ne, hipMemcpyHostToDevice, st));
  EDGE_SETUP_TRY(hipMemcpyAsync(d_fael, in.fa_el, sizeof(int64_t) * 2 * nf, hipMemcpyHostToDevice, st));
  EDGE_SETUP_TRY(hipMemcpyAsync(d_elfa, in.el_fa, sizeof(int64_t) * 4 * ne, hipMemcpyHostToDevice, st));
  EDGE_SETUP_TRY(hipMemcpyAsync(d_type, in.fa_type, sizeof(int32_t) * nf, hipMemcpyHostToDevice, st));
  EDGE_SETUP_TRY(hipMemcpyAsync(d_nrm, in.normals, sizeof(double) * 3 * nf, hipMemcpyHostToDevice, st));
  EDGE_SETUP_TRY(hipMemcpyAsync(d_tan, in.tangents, sizeof(double) * 6 * nf, hipMemcpyHostToDevice, st));
  EDGE_SETUP_TRY(hipMemcpyAsync(d_area, in.areas, sizeof(double) * nf, hipMemcpyHostToDevice, st));
  EDGE_SETUP_TRY(hipMemsetAsync(d_bad, 0, sizeof(int), st));
  // sentinel values: every (element, face) slot must be written by its face
  EDGE_SETUP_TRY(hipMemsetAsync(d_fs, 0xFF, sizeof(T) * 2 * 4 * 81 * ne, st));
  if (d_fs_a) EDGE_SETUP_TRY(hipMemsetAsync(d_fs_a, 0xFF, sizeof(T) * 2 * 4 * 54 * ne, st));
  const int bs = 128;
  hipLaunchKernelGGL(k_star<T>, dim3((unsigned)((ne + bs - 1) / bs)), dim3(bs), 0, st, ne, d_crds, d_elve, d_rho,
                     d_lam, d_mu, d_star, d_jdet, d_star_a, d_bad);
  EDGE_SETUP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_flux_solvers<T>, dim3((unsigned)((nf + bs - 1) / bs)), dim3(bs), 0, st, nf, ne, d_fael,
                     d_elfa, d_type, d_nrm, d_tan, d_area, d_jdet, d_rho, d_lam, d_mu, d_fs, d_fsf, d_fs_a, d_fsfa,
                     d_bad);
  EDGE_SETUP_TRY(hipGetLastError());
  int bad = 0;
  EDGE_SETUP_TRY(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, st));
  EDGE_SETUP_TRY(hipStreamSynchronize(st));
  if (bad & 1) rc = fail(EDGE_ERR_MESH, "element with non-positive Jacobian determinant");
  else if (bad & 2) rc = fail(EDGE_ERR_MESH, "inconsistent face-element relations");
  else if (bad & 4) rc = fail(EDGE_ERR_MESH, "face with non-positive area");
  if (rank3) *rank3 = d_fsf && !(bad & 8);
  cleanup();
#undef EDGE_SETUP_TRY
  return rc;
}

template int build_setup<double>(const SetupInput&, double*, double*, double*, double*, bool*, double*, double*,
                                  double*, hipStream_t);
template int build_setup<float>(const SetupInput&, float*, float*, double*, float*, bool*, float*, float*, float*,
                                 hipStream_t);

template <typename T>
int factor_flux_solvers(const double* h_fs, int64_t n_mats, T* d_fsf, bool* rank3) {
  double* d_fs = nullptr;
  int* d_bad = nullptr;
  int bad = 0;
  hipError_t e = hipMalloc(&d_fs, sizeof(double) * 81 * n_mats);
  if (e == hipSuccess) e = hipMalloc(&d_bad, sizeof(int));
  if (e == hipSuccess) e = hipMemcpy(d_fs, h_fs, sizeof(double) * 81 * n_mats, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(d_bad, 0, sizeof(int));
  if (e == hipSuccess) {
    const int bs = 128;
    hipLaunchKernelGGL(k_fs_factor<T>, dim3((unsigned)((n_mats + bs - 1) / bs)), dim3(bs), 0, 0, n_mats, d_fs,
                       d_fsf, d_bad);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost);
  (void)hipFree(d_fs);
  (void)hipFree(d_bad);
  if (e != hipSuccess) return fail(EDGE_ERR_HIP, std::string("flux-solver factors: ") + hipGetErrorString(e));
  *rank3 = !(bad & 8);
  return EDGE_OK;
}

template int factor_flux_solvers<double>(const double*, int64_t, double*, bool*);
template int factor_flux_solvers<float>(const double*, int64_t, float*, bool*);

}  // namespace edge
