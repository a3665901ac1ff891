// Viscoelastic set-up on the host (FP64): relaxation frequencies, the
// constant-Q least-squares fit per element, the unrelaxed (elastic) Lame
// parameters and the anelastic source matrices.
//
// Follows src/impl/seismic/setups/ViscoElasticity.{h,cpp}:
//   frequencies          ViscoElasticity.h:169-182
//   anelasticCoeffPs     ViscoElasticity.cpp:69-128 (the reference solves the
//                        least-squares problem with an SVD; its full-rank
//                        solution is unique, here by Householder QR)
//   elasticLame          ViscoElasticity.cpp:30-67
//   anelasticCoeffLame   ViscoElasticity.cpp:130-176
//   src (3D)             ViscoElasticity.h:380-458
// and the relaxation frequencies handed to the kernels, every second of the
// 2n-1 frequencies (src/impl/seismic/solvers/AderDg.hpp:248-258).
#include "visco.hpp"

#include <cmath>

namespace edge {
namespace {

constexpr double kPi = 3.14159265358979323846;

// least-squares solution of A y = b, A [m][n] row-major (m >= n, full rank)
void lsq(int m, int n, double* A, double* b, double* y) {
  for (int k = 0; k < n; k++) {
    double nr = 0;
    for (int i = k; i < m; i++) nr += A[i * n + k] * A[i * n + k];
    nr = std::sqrt(nr);
    const double alpha = A[k * n + k] > 0 ? -nr : nr;
    double v[2 * kMaxMechs];
    for (int i = k; i < m; i++) v[i] = A[i * n + k];
    v[k] -= alpha;
    double vn = 0;
    for (int i = k; i < m; i++) vn += v[i] * v[i];
    if (vn == 0) continue;
    for (int j = k; j < n; j++) {
      double d = 0;
      for (int i = k; i < m; i++) d += v[i] * A[i * n + j];
      for (int i = k; i < m; i++) A[i * n + j] -= 2 * d / vn * v[i];
    }
    double d = 0;
    for (int i = k; i < m; i++) d += v[i] * b[i];
    for (int i = k; i < m; i++) b[i] -= 2 * d / vn * v[i];
  }
  for (int k = n - 1; k >= 0; k--) {
    double r = b[k];
    for (int j = k + 1; j < n; j++) r -= A[k * n + j] * y[j];
    y[k] = r / A[k * n + k];
  }
}

// anelastic coefficients of one wave type for a constant quality factor q
void coeff_ps(int n, double q, const double* freqs, double* out) {
  const int m = 2 * n - 1;
  const double qi = 1.0 / q;
  double A[2 * kMaxMechs * kMaxMechs], b[2 * kMaxMechs];
  for (int fr = 0; fr < m; fr++) {
    for (int me = 0; me < n; me++) {
      const double mf = freqs[me * 2];
      A[fr * n + me] = (mf * freqs[fr] + mf * mf * qi) / (freqs[fr] * freqs[fr] + mf * mf);
    }
    b[fr] = qi;
  }
  lsq(m, n, A, b, out);
}

}  // namespace

void visco_frequencies(int n, double f_cen, double f_rat, double* freqs) {
  freqs[0] = 2 * kPi * f_cen / std::sqrt(f_rat);
  for (int fr = 1; fr < 2 * n - 1; fr++)
    freqs[fr] = std::exp(std::log(freqs[0]) + (fr * std::log(f_rat)) / (2 * (n - 1)));
}

void visco_element(int n, const double* freqs, double f_cen, double qp, double qs, double lam, double mu,
                   double* lam_e, double* mu_e, double* src) {
  double ck[kMaxMechs], cm[kMaxMechs];
  coeff_ps(n, qp, freqs, ck);
  coeff_ps(n, qs, freqs, cm);
  // unrelaxed Lame parameters at the reference frequency
  const double f_ref = 2 * kPi * f_cen;
  double th1[2] = {1, 1}, th2[2] = {0, 0};
  for (int me = 0; me < n; me++) {
    const double rat = f_ref / freqs[me * 2], sca = 1.0 / (1 + rat * rat);
    th1[0] -= ck[me] * sca;
    th1[1] -= cm[me] * sca;
    th2[0] += ck[me] * rat * sca;
    th2[1] += cm[me] * rat * sca;
  }
  const double r0 = std::sqrt(th1[0] * th1[0] + th2[0] * th2[0]), r1 = std::sqrt(th1[1] * th1[1] + th2[1] * th2[1]);
  const double m_e = mu * (r1 + th1[1]) / (2 * r1 * r1);
  const double l_e = (lam + 2 * mu) * (r0 + th1[0]) / (2 * r0 * r0) - 2 * m_e;
  // Lame form of the coefficients, then the source matrices on the stresses
  for (int me = 0; me < n; me++) {
    const double sca = 2.0 * m_e / l_e;
    const double cl = (1.0 + sca) * ck[me] - sca * cm[me];
    double* o = src + me * 36;
    for (int i = 0; i < 36; i++) o[i] = 0;
    const double d = -l_e * cl - 2 * m_e * cm[me], sh = -2 * m_e * cm[me], off = -l_e * cl;
    for (int i = 0; i < 3; i++) {
      o[i * 6 + i] = d;
      o[(3 + i) * 6 + 3 + i] = sh;
      for (int j = 0; j < 3; j++)
        if (j != i) o[i * 6 + j] = off;
    }
  }
  *lam_e = l_e;
  *mu_e = m_e;
}

}  // namespace edge
