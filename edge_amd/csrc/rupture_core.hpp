// The per-sub-face, per-run core of the internal-boundary net updates with
// linear slip-weakening friction, FP64 (InternalBoundary::netUpdates,
// src/impl/seismic/solvers/InternalBoundary.hpp:184-280, with
// FrictionLaws<3,R>::linSlipWeak, src/impl/seismic/solvers/FrictionLaws.hpp:
// 309-392).  Shared by the stand-alone net-update kernel and the fault phase
// of the coupled time step.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace edge {

constexpr int kRuptureStateVals = 14;  // sn0, co0, ss0[2], ss0A, muf, st, tr[2], sr[2], dd[2]

// Device view of a rupture handle (rupture.hip).
struct RuptureDev {
  const double* solv;  // [nF][4][81]
  const double* face;  // [nF][3]
  const double* glob;  // [3][R]
  double* state;       // [nF][14][n_sf][R]
  uint8_t* per;        // [nF][n_sf][R]
  int n_sf, n_runs;
  int64_t n_fa;
};
int rupture_dev(edge_rupture* r, RuptureDev* out);

// s: the face's solvers [4][81] (T^-1, jump, left flux, right flux, the flux
// solvers scaled by (area / n_sf) / (volume / n_sc)); face: l_eq_m, cs / mu
// of the minus and plus side; st: the sub-face's friction state, value k at
// st[k * sst].  qL, qR: both sides' quantities at the sub-face (physical
// frame).  Returns whether the friction law perturbed the middle state and
// sets the net updates (left-going negative, both scaled by dt).
template <typename P>
__device__ inline bool lsw_net_updates(const P* s, const P* face, double mus, double mud, double dcInv, double dt,
                                       const double (&qLp)[9], const double (&qRp)[9], double* st, int64_t sst,
                                       double (&netL)[9], double (&netR)[9]) {
  double qL[9], qR[9];
  for (int q = 0; q < 9; q++) {  // rotate to face coordinates
    double a = 0, b = 0;
    for (int p = 0; p < 9; p++) {
      a += s[q * 9 + p] * qLp[p];
      b += s[q * 9 + p] * qRp[p];
    }
    qL[q] = a;
    qR[q] = b;
  }
  double ms[9];
  for (int q = 0; q < 9; q++) {
    double v = qL[q];
    for (int p = 0; p < 9; p++) v += s[81 + q * 9 + p] * (qR[p] - qL[p]);
    ms[q] = v;
  }
  const bool l_eq_m = face[0] != 0.0;
  const double cdM = face[1], cdP = face[2];
  double msM[9], msP[9];
  for (int q = 0; q < 9; q++) msM[q] = msP[q] = ms[q];
  double strength = st[5 * sst] * (st[0 * sst] + ms[0]) + st[1 * sst];
  strength = strength < 0 ? -strength : 0.0;
  const double s1 = st[2 * sst] + ms[3], s2 = st[3 * sst] + ms[5];
  const double shear = sqrt(s1 * s1 + s2 * s2);
  const bool failing = shear > strength;
  const double scale = strength / shear;
  const double tr0 = failing ? s1 * scale - st[2 * sst] : ms[3];
  const double tr1 = failing ? s2 * scale - st[3 * sst] : ms[5];
  msM[3] = msP[3] = tr0;
  msM[5] = msP[5] = tr1;
  const double d1 = tr0 - ms[3], d2 = tr1 - ms[5];
  msM[7] -= cdM * d1;
  msM[8] -= cdM * d2;
  msP[7] += cdP * d1;
  msP[8] += cdP * d2;
  const double sr0 = msP[7] - msM[7], sr1 = msP[8] - msM[8];
  const double dd0 = st[11 * sst] + fabs(sr0) * dt, dd1 = st[12 * sst] + fabs(sr1) * dt;
  const double arg = mus - (mus - mud) * dcInv * sqrt(dd0 * dd0 + dd1 * dd1);
  st[5 * sst] = arg > mud ? arg : mud;
  st[6 * sst] = strength;
  st[7 * sst] = tr0 + st[2 * sst];
  st[8 * sst] = tr1 + st[3 * sst];
  st[9 * sst] = sr0;
  st[10 * sst] = sr1;
  st[11 * sst] = dd0;
  st[12 * sst] = dd1;
  const double* mL = l_eq_m ? msM : msP;
  const double* mR = l_eq_m ? msP : msM;
  for (int q = 0; q < 9; q++) {  // left-going updates subtract, right-going add
    double a = 0, b = 0;
    for (int p = 0; p < 9; p++) {
      a += s[162 + q * 9 + p] * mL[p];
      b += s[243 + q * 9 + p] * mR[p];
    }
    netL[q] = -a * dt;
    netR[q] = b * dt;
  }
  return failing;
}

}  // namespace edge
