// Test-only harness around the reference's own set-up code, compiled from
// the reference tree in place (see oracle/Makefile).  It exposes, with a
// C ABI, the pieces the oracle's flux-solver and star-matrix restatement is
// checked against:
//   edge::seismic::setups::Elasticity::star  (Elasticity.cpp:167-252)
//   edge::seismic::setups::Elasticity::fs    (Elasticity.cpp:372-441)
//   edge::seismic::common::setupTrafo3d / setupTrafoInv3d (common.hpp:314-344)
//   edge::seismic::common::getTimeStepCFL (common.hpp:84-98)
//   the generated middle-state jump / flux solvers of the internal boundary
// composed the way AderDgInit::setUpFs composes them (AderDgInit.hpp:300-370).
#include <algorithm>
#include <cmath>
#include <limits>

#include "impl/seismic/common.hpp"
#include "impl/seismic/setups/Elasticity.h"

extern "C" void ref_star(double rho, double lam, double mu, const double* jinv, double* out) {
  double ji[3][3];
  for (int i = 0; i < 9; i++) ji[i / 3][i % 3] = jinv[i];
  double s[3][81];
  edge::seismic::setups::Elasticity::star(rho, lam, mu, ji, s);
  for (int i = 0; i < 243; i++) out[i] = s[i / 81][i % 81];
}

extern "C" void ref_trafo(const double* n, const double* s, const double* t, double* T, double* Ti) {
  double a[9][9], b[9][9];
  edge::seismic::common::setupTrafo3d(n[0], n[1], n[2], s[0], s[1], s[2], t[0], t[1], t[2], a);
  edge::seismic::common::setupTrafoInv3d(n[0], n[1], n[2], s[0], s[1], s[2], t[0], t[1], t[2], b);
  for (int i = 0; i < 81; i++) {
    T[i] = a[i / 9][i % 9];
    Ti[i] = b[i / 9][i % 9];
  }
}

extern "C" void ref_fs(double rhoL, double rhoR, double lamL, double lamR, double muL, double muR, const double* n,
                       const double* s, const double* t, int free_surface, double* fsL, double* fsR) {
  double a[9][9], b[9][9];
  edge::seismic::common::setupTrafo3d(n[0], n[1], n[2], s[0], s[1], s[2], t[0], t[1], t[2], a);
  edge::seismic::common::setupTrafoInv3d(n[0], n[1], n[2], s[0], s[1], s[2], t[0], t[1], t[2], b);
  edge::seismic::setups::Elasticity::fs(rhoL, rhoR, lamL, lamR, muL, muR, a, b, fsL, fsR, free_surface != 0);
}

// The reference's generated middle-state solvers of the internal boundary
// (src/impl/seismic/generated/MiddleState{JumpL,JumpR,Flux}3D.inc), included
// into functions with the parameters InternalBoundarySolvers<TET4> gives them
// (src/impl/seismic/solvers/InternalBoundary.hpp:650-740).
extern "C" void ref_ms_jump_l(double i_lamL, double i_lamR, double i_muL, double i_muR, double i_rhoL,
                              double i_rhoR, double* out) {
  double o_msJump[9][9];
#include "impl/seismic/generated/MiddleStateJumpL3D.inc"
  for (int i = 0; i < 81; i++) out[i] = o_msJump[i / 9][i % 9];
}

extern "C" void ref_ms_jump_r(double i_lamL, double i_lamR, double i_muL, double i_muR, double i_rhoL,
                              double i_rhoR, double* out) {
  double o_msJump[9][9];
#include "impl/seismic/generated/MiddleStateJumpR3D.inc"
  for (int i = 0; i < 81; i++) out[i] = o_msJump[i / 9][i % 9];
}

extern "C" void ref_ms_flux(const double* n, const double* s, const double* t, double i_lam, double i_mu,
                            double i_rho, double* out) {
  const double i_nx = n[0], i_ny = n[1], i_nz = n[2], i_sx = s[0], i_sy = s[1], i_sz = s[2], i_tx = t[0],
               i_ty = t[1], i_tz = t[2];
  double o_msFlux[9][9];
#include "impl/seismic/generated/MiddleStateFlux3D.inc"
  for (int i = 0; i < 81; i++) out[i] = o_msFlux[i / 9][i % 9];
}

// The reference's CFL time step (common.hpp:84-98) for the order this
// harness is compiled with (PP_ORDER); i_scale is SCALE_CFL = 1/2
// (const.inc:135-137) in the reference's set-up.
extern "C" double ref_dt_cfl(double rho, double lam, double mu, double in_dia, double scale) {
  return edge::seismic::common::getTimeStepCFL(rho, lam, mu, in_dia, scale);
}
