// Host side of the element-batched single-run kernels: packed A operands and
// the face-orientation maps (ec64_pack.hpp).
#include "ec64_pack.hpp"

#include <algorithm>
#include <cmath>

namespace edge {
namespace {

// Appends the packed A operand of Out[m][e] = sum_{k < nk} M[k][m] X_e[k],
// m < nn, for M [rows][ld] (row-major), over ks k-steps with the row map of
// B operands from LDS (chain = false) or from D tiles (chain = true, which
// differs for FP32).
void pack_product(const std::vector<double>& m, int ld, int nk, int nn, int ks, bool chain, bool f32,
                  std::vector<double>* out) {
  const int nt = ec::tiles(nn), np = ec::pairs(ks);
  for (int t = 0; t < nt; t++)
    for (int p = 0; p < np; p++)
      for (int lane = 0; lane < 64; lane++)
        for (int i = 0; i < 2; i++) {
          const int g = lane >> 4, c = lane & 15, s = 2 * p + i, col = 16 * t + c;
          const int k = chain && f32 ? 16 * (s / 4) + 4 * g + s % 4 : 4 * s + g;
          out->push_back(s < ks && k < nk && col < nn ? m[(size_t)k * ld + col] : 0.0);
        }
}

// Least squares X = argmin ||A X - B|| for A [n][k] (full column rank) and B
// [n][k], by Householder QR; X [k][k].
void lstsq(int n, int k, std::vector<double> a, std::vector<double> b, std::vector<double>* x) {
  for (int j = 0; j < k; j++) {
    double nrm = 0;
    for (int i = j; i < n; i++) nrm += a[i * k + j] * a[i * k + j];
    nrm = std::sqrt(nrm);
    const double alpha = a[j * k + j] > 0 ? -nrm : nrm;
    std::vector<double> v(n, 0.0);
    for (int i = j; i < n; i++) v[i] = a[i * k + j];
    v[j] -= alpha;
    double vv = 0;
    for (int i = j; i < n; i++) vv += v[i] * v[i];
    if (vv == 0) continue;
    auto reflect = [&](std::vector<double>& m, int cols, int c0) {
      for (int c = c0; c < cols; c++) {
        double d = 0;
        for (int i = j; i < n; i++) d += v[i] * m[i * cols + c];
        d *= 2.0 / vv;
        for (int i = j; i < n; i++) m[i * cols + c] -= d * v[i];
      }
    };
    reflect(a, k, j);
    reflect(b, k, 0);
  }
  x->assign((size_t)k * k, 0.0);
  for (int c = 0; c < k; c++)
    for (int i = k - 1; i >= 0; i--) {
      double s = b[i * k + c];
      for (int j = i + 1; j < k; j++) s -= a[i * k + j] * (*x)[j * k + c];
      (*x)[i * k + c] = s / a[i * k + i];
    }
}

}  // namespace

std::vector<double> ec_pack_operands(const DgTables& t, bool f32) {
  const int O = t.order, N = t.n, NF = t.nf;
  std::vector<double> out;
  const ec::Layout L{O, f32};
  out.reserve((size_t)L.size());
  std::vector<double> m;
  for (int de = 1; de < O; de++)
    for (int d = 0; d < 3; d++) {
      m.assign(t.stiffT + (size_t)d * N * N, t.stiffT + (size_t)(d + 1) * N * N);
      for (double& x : m) x = -x;  // -(K_d^T M^-1), as src/dg/TimePred.hpp:86 stores it
      const int nk = n_modes(O - de + 1);
      if (de == 1)
        pack_product(m, N, nk, n_modes(O - de), ec::ksteps(nk), false, f32, &out);
      else
        pack_product(m, N, nk, n_modes(O - de), L.chain(nk), true, f32, &out);
    }
  for (int d = 0; d < 3; d++) {
    m.assign(t.stiffV + (size_t)d * N * N, t.stiffV + (size_t)(d + 1) * N * N);
    pack_product(m, N, n_modes(O - 1), N, ec::ksteps(n_modes(O - 1)), false, f32, &out);
  }
  for (int f = 0; f < 4; f++) {
    m.assign(t.fluxL + (size_t)f * N * NF, t.fluxL + (size_t)(f + 1) * N * NF);
    pack_product(m, NF, N, NF, ec::ksteps(N), false, f32, &out);
  }
  for (int f = 0; f < 4; f++) {
    m.assign(t.fluxT + (size_t)f * NF * N, t.fluxT + (size_t)(f + 1) * NF * N);
    pack_product(m, N, NF, N, L.chain(NF), true, f32, &out);
  }
  for (int f = 0; f < 4; f++) {
    m.assign(t.fluxT + (size_t)f * NF * N, t.fluxT + (size_t)(f + 1) * NF * N);
    pack_product(m, N, NF, N, ec::ksteps(NF), false, f32, &out);
  }
  return out;
}

bool ec_rotations(const DgTables& t, std::vector<double>* rot) {
  const int N = t.n, NF = t.nf;
  if (NF > 16) return false;
  rot->assign(ec::kRotSize, 0.0);
  for (int k = 0; k < NF; k++) (*rot)[k * 16 + k] = 1.0;
  for (int v = 0; v < 3; v++) {
    std::vector<double> r0;
    for (int f = 0; f < 4; f++) {
      const std::vector<double> a(t.fluxL + (size_t)f * N * NF, t.fluxL + (size_t)(f + 1) * N * NF);
      const std::vector<double> b(t.fluxN + (size_t)(4 * v + f) * N * NF, t.fluxN + (size_t)(4 * v + f + 1) * N * NF);
      std::vector<double> x;
      lstsq(N, NF, a, b, &x);
      // residual of fluxL[f] R - fluxN[4 v + f]
      double res = 0, ref = 0;
      for (int i = 0; i < N; i++)
        for (int c = 0; c < NF; c++) {
          double s = -b[i * NF + c];
          for (int j = 0; j < NF; j++) s += a[i * NF + j] * x[j * NF + c];
          res = std::max(res, std::fabs(s));
          ref = std::max(ref, std::fabs(b[i * NF + c]));
        }
      if (!(res <= 1e-12 * ref)) return false;
      if (f == 0) {
        r0 = x;
      } else {
        double dif = 0, mag = 0;
        for (size_t i = 0; i < x.size(); i++) {
          dif = std::max(dif, std::fabs(x[i] - r0[i]));
          mag = std::max(mag, std::fabs(r0[i]));
        }
        if (!(dif <= 1e-12 * mag)) return false;
      }
    }
    // block-diagonal by degree (the kernels only apply the blocks): face
    // mode j has degree p with p (p + 1) / 2 <= j < (p + 1) (p + 2) / 2
    auto degree = [](int j) {
      int p = 0;
      while ((p + 1) * (p + 2) / 2 <= j) p++;
      return p;
    };
    double mag = 0;
    for (double x : r0) mag = std::max(mag, std::fabs(x));
    for (int j = 0; j < NF; j++)
      for (int c = 0; c < NF; c++) {
        const double x = r0[j * NF + c];
        if (degree(j) != degree(c)) {
          if (!(std::fabs(x) <= 1e-12 * mag)) return false;
          continue;
        }
        (*rot)[((1 + v) * 16 + j) * 16 + c] = x;
      }
  }
  return true;
}

}  // namespace edge
