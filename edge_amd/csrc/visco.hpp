// Viscoelastic set-up on the host (see visco.cpp).
#pragma once

namespace edge {

constexpr int kMaxMechs = 5;  // relaxation mechanisms supported

// 2n-1 relaxation frequencies log-spaced around 2 pi f_cen over f_rat
void visco_frequencies(int n, double f_cen, double f_rat, double* freqs);

// Constant-Q fit of one element: the unrelaxed Lame parameters (from the
// phase ones at the central frequency) and the anelastic source matrices
// src [n][6][6] acting on the stresses.
void visco_element(int n, const double* freqs, double f_cen, double qp, double qs, double lam, double mu,
                   double* lam_e, double* mu_e, double* src);

}  // namespace edge
