/*
 * Oracle for the ADER-DG elastic / viscoelastic tet4 element update.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the CPU checker: a plain-C, FP64
 * restatement of the reference's vanilla path.  Only the test suite, the
 * smoke check and the CPU-baseline leg of the benchmark load it; the product
 * (the HIP library) never calls into it.
 *
 * Each function names the reference lines it restates.  Layouts follow the
 * reference exactly: DOFs [el][q][mode][run] (run innermost), anelastic DOFs
 * [el][rm][6][mode][run], star matrices [el][3][81] (dense 9x9),
 * flux solvers [2][el][4][81], anelastic star [el][3][18] (6x3 dense),
 * anelastic source [el][rm][36] (6x6 dense), anelastic flux solvers
 * [2][el][4][54] (6x9 dense).
 *
 * Pinning: see tests/test_oracle_*.py.  Kernels are pinned by the reference's
 * golden vectors (TimePred/VolInt/SurfInt .test.inc), the setup by the
 * reference's own Elasticity.cpp compiled into oracle/_ref, and the scheme by
 * the plane-wave convergence order.
 */
#include <math.h>
#include <stdint.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif
#include <stdlib.h>
#include <string.h>

#define NQ 9   /* elastic quantities */
#define NQA 6  /* quantities per relaxation mechanism */
#define NO_ID16 0xFFFFu

/* face types, values of the reference's seismic spType enum
 * (src/impl/seismic/const.inc:143-148) */
#define FA_FREE_SURFACE 101
#define FA_OUTFLOW 105
#define FA_PERIODIC 106

static int nmodes(int order) { return order * (order + 1) * (order + 2) / 6; }
static int nmodes_face(int order) { return order * (order + 1) / 2; }

/* -------------------------------------------------------------------------
 * Small fused GEMMs, src/linalg/Matrix.h:296-380 (matMulFusedAC / BC).
 * Row-major, run axis innermost; alpha = 1, beta in {0, 1}.
 * ---------------------------------------------------------------------- */
static void mm_ac(int r, int m, int n, int k, int lda, int ldb, int ldc, int beta,
                  const double *a, const double *b, double *c) {
  if (!beta)
    for (int im = 0; im < m; im++)
      for (int in = 0; in < n; in++)
        for (int ir = 0; ir < r; ir++) c[(im * ldc + in) * r + ir] = 0;
  for (int ik = 0; ik < k; ik++)
    for (int im = 0; im < m; im++)
      for (int in = 0; in < n; in++) {
        double bv = b[ik * ldb + in];
        for (int ir = 0; ir < r; ir++) c[(im * ldc + in) * r + ir] += a[(im * lda + ik) * r + ir] * bv;
      }
}

static void mm_bc(int r, int m, int n, int k, int lda, int ldb, int ldc, int beta,
                  const double *a, const double *b, double *c) {
  if (!beta)
    for (int im = 0; im < m; im++)
      for (int in = 0; in < n; in++)
        for (int ir = 0; ir < r; ir++) c[(im * ldc + in) * r + ir] = 0;
  for (int ik = 0; ik < k; ik++)
    for (int im = 0; im < m; im++) {
      double av = a[im * lda + ik];
      for (int in = 0; in < n; in++)
        for (int ir = 0; ir < r; ir++) c[(im * ldc + in) * r + ir] += av * b[(ik * ldb + in) * r + ir];
    }
}

/* -------------------------------------------------------------------------
 * Reference-element matrices as handed to the kernels.
 * stiffT is stored un-negated; the CK negates it the way
 * src/dg/TimePred.hpp:86 does when it stores the matrices.
 * ---------------------------------------------------------------------- */
typedef struct {
  int order;            /* O = degree + 1 */
  int n_crs;            /* fused runs R */
  int n_rms;            /* relaxation mechanisms */
  const double *stiffT; /* [3][N][N]  K_d^T M^-1 */
  const double *stiffV; /* [3][N][N]  K_d M^-1   */
  const double *fluxL;  /* [4][N][NF] */
  const double *fluxN;  /* [12][N][NF] */
  const double *fluxT;  /* [4][NF][N] */
  const double *rfs;    /* [n_rms] relaxation frequencies */
} orc_dg;

/* -------------------------------------------------------------------------
 * Cauchy-Kowalevski time prediction,
 * src/impl/seismic/kernels/TimePredVanilla.hpp:200-306 (kernels :103-158).
 * derE: [O][9][N][R], tIntE: [9][N][R], derA: [rm][O][6][N][R], tIntA: [rm][6][N][R]
 * ---------------------------------------------------------------------- */
void orc_ck(const orc_dg *dg, double dt, const double *starE, const double *starA,
            const double *srcA, const double *dofsE, const double *dofsA, double *derE,
            double *derA, double *tIntE, double *tIntA) {
  const int O = dg->order, N = nmodes(O), R = dg->n_crs, NR = dg->n_rms;
  const int sz = NQ * N * R, szA = NQA * N * R;
  double *scratch = (double *)calloc((size_t)sz, sizeof(double));
  double *scrA = (double *)calloc((size_t)szA, sizeof(double));
  double *negT = (double *)malloc(sizeof(double) * 3 * N * N);
  for (int i = 0; i < 3 * N * N; i++) negT[i] = -dg->stiffT[i];

  double sca = dt;
  for (int i = 0; i < sz; i++) {
    derE[i] = dofsE[i];
    tIntE[i] = sca * dofsE[i];
  }
  for (int rm = 0; rm < NR; rm++)
    for (int i = 0; i < szA; i++) {
      derA[(rm * O + 0) * szA + i] = dofsA[rm * szA + i];
      tIntA[rm * szA + i] = sca * dofsA[rm * szA + i];
    }

  for (int de = 1; de < O; de++) {
    /* elastic runs truncate to the hierarchical modes still alive; the
     * viscoelastic branch keeps all modes (TimePredVanilla.hpp:232, :296) */
    const int re = (NR == 0) ? de : 1;
    const int nk = (NR == 0) ? nmodes(O - re + 1) : N; /* CE_N_ELEMENT_MODES_CK(l_re-1) */
    const int nn = (NR == 0) ? nmodes(O - re) : nmodes(O - 1);
    double *dcur = derE + (size_t)de * sz;
    const double *dprev = derE + (size_t)(de - 1) * sz;
    for (int i = 0; i < sz; i++) dcur[i] = 0;
    for (int i = 0; i < szA; i++) scrA[i] = 0;
    for (int d = 0; d < 3; d++) {
      mm_ac(R, NQ, nn, nk, N, N, N, 0, dprev, negT + (size_t)d * N * N, scratch);
      mm_bc(R, NQ, nn, NQ, NQ, N, N, 1, starE + d * 81, scratch, dcur);
      if (NR > 0) mm_bc(R, NQA, nn, 3, 3, N, N, 1, starA + d * 18, scratch + NQA * N * R, scrA);
    }
    sca *= dt / (de + 1);
    for (int rm = 0; rm < NR; rm++) {
      mm_bc(R, NQA, N, NQA, NQA, N, N, 1, srcA + rm * 36, derA + ((size_t)rm * O + de - 1) * szA, dcur);
      double *dA = derA + ((size_t)rm * O + de) * szA;
      const double *dAp = derA + ((size_t)rm * O + de - 1) * szA;
      for (int i = 0; i < szA; i++) {
        dA[i] = dg->rfs[rm] * (scrA[i] + dAp[i]);
        tIntA[rm * szA + i] += sca * dA[i];
      }
    }
    const int ncp = (NR == 0) ? nmodes(O - de) : N;
    for (int q = 0; q < NQ; q++)
      for (int md = 0; md < ncp; md++)
        for (int r = 0; r < R; r++) tIntE[(q * N + md) * R + r] += sca * dcur[(q * N + md) * R + r];
  }
  free(scratch);
  free(scrA);
  free(negT);
}

/* -------------------------------------------------------------------------
 * Taylor integration over [0, dt] of stored derivatives,
 * src/impl/seismic/kernels/TimePred.hpp:165-202; and evaluation at a time,
 * TimePred.hpp:130-156.
 * ---------------------------------------------------------------------- */
void orc_integrate(const orc_dg *dg, double dt, const double *derE, double *tIntE) {
  const int O = dg->order, N = nmodes(O), R = dg->n_crs, sz = NQ * N * R;
  double sca = dt;
  for (int i = 0; i < sz; i++) tIntE[i] = 0;
  for (int de = 0; de < O; de++) {
    for (int i = 0; i < sz; i++) tIntE[i] += sca * derE[(size_t)de * sz + i];
    sca *= dt / (de + 2);
  }
}

void orc_eval_time_pred(const orc_dg *dg, double t, const double *derE, double *out) {
  const int O = dg->order, N = nmodes(O), R = dg->n_crs, sz = NQ * N * R;
  double sca = 1.0;
  for (int i = 0; i < sz; i++) out[i] = 0;
  for (int de = 0; de < O; de++) {
    for (int i = 0; i < sz; i++) out[i] += sca * derE[(size_t)de * sz + i];
    sca *= t / (de + 1);
  }
}

/* -------------------------------------------------------------------------
 * Volume integral, src/impl/seismic/kernels/VolIntVanilla.hpp:189-241.
 * ---------------------------------------------------------------------- */
void orc_vol(const orc_dg *dg, const double *starE, const double *starA, const double *srcA,
             const double *tIntE, const double *tIntA, double *dofsE, double *dofsA) {
  const int O = dg->order, N = nmodes(O), R = dg->n_crs, NR = dg->n_rms;
  const int szA = NQA * N * R;
  double *scratch = (double *)calloc((size_t)NQ * N * R, sizeof(double));
  double *scrA = (double *)calloc((size_t)szA, sizeof(double));
  const int nk = nmodes(O - 1);
  for (int d = 0; d < 3; d++) {
    mm_ac(R, NQ, N, nk, N, N, N, 0, tIntE, dg->stiffV + (size_t)d * N * N, scratch);
    mm_bc(R, NQ, N, NQ, NQ, N, N, 1, starE + d * 81, scratch, dofsE);
    if (NR > 0) mm_bc(R, NQA, N, 3, 3, N, N, 1, starA + d * 18, scratch + NQA * N * R, scrA);
  }
  for (int rm = 0; rm < NR; rm++) {
    mm_bc(R, NQA, N, NQA, NQA, N, N, 1, srcA + rm * 36, tIntA + rm * szA, dofsE);
    for (int i = 0; i < szA; i++) dofsA[rm * szA + i] += dg->rfs[rm] * (scrA[i] - tIntA[rm * szA + i]);
  }
  free(scratch);
  free(scrA);
}

/* -------------------------------------------------------------------------
 * Surface integrals, src/impl/seismic/kernels/SurfIntVanilla.hpp:212-357 and
 * SurfInt.hpp:127-156.
 * ---------------------------------------------------------------------- */
void orc_scatter_update_a(const orc_dg *dg, const double *upd, double *dofsA) {
  const int N = nmodes(dg->order), R = dg->n_crs, szA = NQA * N * R;
  for (int rm = 0; rm < dg->n_rms; rm++)
    for (int i = 0; i < szA; i++) dofsA[rm * szA + i] += dg->rfs[rm] * upd[i];
}

void orc_surf_local(const orc_dg *dg, const double *fsE /*[4][81]*/, const double *fsA /*[4][54]*/,
                    const double *tIntE, double *dofsE, double *dofsA) {
  const int O = dg->order, N = nmodes(O), NF = nmodes_face(O), R = dg->n_crs;
  double *s0 = (double *)calloc((size_t)NQ * NF * R, sizeof(double));
  double *s1 = (double *)calloc((size_t)NQ * NF * R, sizeof(double));
  double *upA = (double *)calloc((size_t)NQA * N * R, sizeof(double));
  for (int fa = 0; fa < 4; fa++) {
    mm_ac(R, NQ, NF, N, N, NF, NF, 0, tIntE, dg->fluxL + (size_t)fa * N * NF, s0);
    mm_bc(R, NQ, NF, NQ, NQ, NF, NF, 0, fsE + fa * 81, s0, s1);
    mm_ac(R, NQ, N, NF, NF, N, N, 1, s1, dg->fluxT + (size_t)fa * NF * N, dofsE);
    if (dg->n_rms > 0) {
      mm_bc(R, NQA, NF, NQ, NQ, NF, NF, 0, fsA + fa * 54, s0, s1);
      mm_ac(R, NQA, N, NF, NF, N, N, 1, s1, dg->fluxT + (size_t)fa * NF * N, upA);
    }
  }
  if (dg->n_rms > 0) orc_scatter_update_a(dg, upA, dofsA);
  free(s0);
  free(s1);
  free(upA);
}

void orc_neigh_flux_int(const orc_dg *dg, int fa, unsigned vId, unsigned fId, const double *tIntE,
                        double *out /*[9][NF][R]*/) {
  const int O = dg->order, N = nmodes(O), NF = nmodes_face(O), R = dg->n_crs;
  const double *fm = (vId != NO_ID16) ? dg->fluxN + (size_t)(vId * 4 + fId) * N * NF
                                      : dg->fluxL + (size_t)fa * N * NF;
  mm_ac(R, NQ, NF, N, N, NF, NF, 0, tIntE, fm, out);
}

/* dofsA receives the UNSCALED anelastic update (SurfIntVanilla.hpp:343-355). */
void orc_surf_neigh(const orc_dg *dg, int fa, unsigned vId, unsigned fId, const double *fsE /*[81]*/,
                    const double *fsA /*[54]*/, const double *tIntE, double *dofsE, double *upA) {
  const int O = dg->order, N = nmodes(O), NF = nmodes_face(O), R = dg->n_crs;
  double *s0 = (double *)calloc((size_t)NQ * NF * R, sizeof(double));
  double *s1 = (double *)calloc((size_t)NQ * NF * R, sizeof(double));
  orc_neigh_flux_int(dg, fa, vId, fId, tIntE, s0);
  mm_bc(R, NQ, NF, NQ, NQ, NF, NF, 0, fsE, s0, s1);
  mm_ac(R, NQ, N, NF, NF, N, N, 1, s1, dg->fluxT + (size_t)fa * NF * N, dofsE);
  if (dg->n_rms > 0) {
    mm_bc(R, NQA, NF, NQ, NQ, NF, NF, 0, fsA, s0, s1);
    mm_ac(R, NQA, N, NF, NF, N, N, 1, s1, dg->fluxT + (size_t)fa * NF * N, upA);
  }
  free(s0);
  free(s1);
}

/* -------------------------------------------------------------------------
 * Solver loops, src/impl/seismic/solvers/AderDg.hpp:338-534 (local) and
 * :553-722 (neigh); global time stepping (tDofs[0] only).
 * ---------------------------------------------------------------------- */
void orc_local(const orc_dg *dg, int64_t first, int64_t n_els, double dt, const double *starE,
               const double *starA, const double *srcA, const double *fsE0 /*[el][4][81]*/,
               const double *fsA0 /*[el][4][54]*/, double *dofsE, double *dofsA, double *tIntE) {
  const int O = dg->order, N = nmodes(O), R = dg->n_crs, NR = dg->n_rms;
  const size_t sz = (size_t)NQ * N * R, szA = (size_t)NQA * N * R;
#pragma omp parallel
  {
    double *der = (double *)malloc(sizeof(double) * O * sz);
    double *derA = (double *)malloc(sizeof(double) * (NR > 0 ? NR : 1) * O * szA);
    double *tA = (double *)malloc(sizeof(double) * (NR > 0 ? NR : 1) * szA);
#pragma omp for schedule(static)
    for (int64_t el = first; el < first + n_els; el++) {
      double *dA = (NR > 0) ? dofsA + (size_t)el * NR * szA : NULL;
      orc_ck(dg, dt, starE + el * 243, NR ? starA + el * 54 : NULL, NR ? srcA + el * NR * 36 : NULL,
             dofsE + el * sz, dA, der, derA, tIntE + el * sz, tA);
      orc_vol(dg, starE + el * 243, NR ? starA + el * 54 : NULL, NR ? srcA + el * NR * 36 : NULL,
              tIntE + el * sz, tA, dofsE + el * sz, dA);
      orc_surf_local(dg, fsE0 + el * 324, NR ? fsA0 + el * 216 : NULL, tIntE + el * sz, dofsE + el * sz,
                     dA);
    }
    free(der);
    free(derA);
    free(tA);
  }
}

void orc_neigh(const orc_dg *dg, int64_t first, int64_t n_els, const int32_t *fa_type /*[nFa]*/,
               const int64_t *el_fa /*[el][4]*/, const int64_t *el_fa_el /*[el][4]*/,
               const uint16_t *f_id /*[el][4]*/, const uint16_t *v_id /*[el][4]*/,
               const double *fsE1 /*[el][4][81]*/, const double *fsA1 /*[el][4][54]*/,
               const double *tIntE, double *dofsE, double *dofsA) {
  const int O = dg->order, N = nmodes(O), R = dg->n_crs, NR = dg->n_rms;
  const size_t sz = (size_t)NQ * N * R, szA = (size_t)NQA * N * R;
#pragma omp parallel
  {
    double *upA = (double *)malloc(sizeof(double) * szA);
#pragma omp for schedule(static)
    for (int64_t el = first; el < first + n_els; el++) {
      for (size_t i = 0; i < szA; i++) upA[i] = 0;
      for (int fa = 0; fa < 4; fa++) {
        const int64_t fid = el_fa[el * 4 + fa];
        const int32_t ty = fa_type[fid];
        if ((ty & FA_OUTFLOW) == FA_OUTFLOW) continue;
        const int fs = (ty & FA_FREE_SURFACE) == FA_FREE_SURFACE;
        const int64_t ne = fs ? el : el_fa_el[el * 4 + fa];
        unsigned vId = NO_ID16, fId = NO_ID16;
        if (!fs) {
          vId = v_id[el * 4 + fa];
          fId = f_id[el * 4 + fa];
        }
        orc_surf_neigh(dg, fa, vId, fId, fsE1 + (el * 4 + fa) * 81, NR ? fsA1 + (el * 4 + fa) * 54 : NULL,
                       tIntE + ne * sz, dofsE + el * sz, upA);
      }
      if (NR > 0) orc_scatter_update_a(dg, upA, dofsA + (size_t)el * NR * szA);
    }
    free(upA);
  }
}

/* -------------------------------------------------------------------------
 * Local time stepping (rate 2), elastic: the tDofs[1] / tDofs[2] buffers of
 * AderDg::local (AderDg.hpp:384-417) and the neighbour data AderDg::neigh
 * takes per face (:620-674).
 *   el_lts[el] bit 0: an adjacent element has a larger time step
 *                     (tDofs[1] sums tDofs[0] over the element's steps since
 *                     the first of every pair);
 *              bit 1: an adjacent element has a smaller time step
 *                     (tDofs[2] = the Taylor integral over [0, dt/2]).
 *   fa_lts[el][fa]: 0 equal steps: the neighbour's tDofs[0];
 *                   1 the element's step is the greater: tDofs[1];
 *                   2 the element's step is the smaller: tDofs[2] on the
 *                     first step of a pair, tDofs[0] - tDofs[2] on the second.
 * ---------------------------------------------------------------------- */
void orc_local_lts(const orc_dg *dg, int64_t first, int64_t n_els, double dt, int first_ts, const double *starE,
                   const double *starA, const double *srcA, const double *fsE0, const double *fsA0,
                   const uint8_t *el_lts, double *dofsE, double *dofsA, double *t0, double *t1, double *t2) {
  const int O = dg->order, N = nmodes(O), R = dg->n_crs, NR = dg->n_rms;
  const size_t sz = (size_t)NQ * N * R, szA = (size_t)NQA * N * R;
#pragma omp parallel
  {
    double *der = (double *)malloc(sizeof(double) * O * sz);
    double *derA = (double *)malloc(sizeof(double) * (NR > 0 ? NR : 1) * O * szA);
    double *tA = (double *)malloc(sizeof(double) * (NR > 0 ? NR : 1) * szA);
#pragma omp for schedule(static)
    for (int64_t el = first; el < first + n_els; el++) {
      double *dA = (NR > 0) ? dofsA + (size_t)el * NR * szA : NULL;
      /* AderDg.hpp:371-381: the CK (with the anelastic terms) into tDofs[0] */
      orc_ck(dg, dt, starE + el * 243, NR ? starA + el * 54 : NULL, NR ? srcA + el * NR * 36 : NULL,
             dofsE + el * sz, dA, der, derA, t0 + el * sz, tA);
      /* :384-411: tDofs[1] sums tDofs[0] over a pair of steps */
      if (el_lts[el] & 1) {
        for (size_t i = 0; i < sz; i++) t1[el * sz + i] = (first_ts ? 0.0 : t1[el * sz + i]) + t0[el * sz + i];
      }
      /* :413-417: tDofs[2] integrates the (elastic) derivatives over [0, dt/2] */
      if (el_lts[el] & 2) orc_integrate(dg, 0.5 * dt, der, t2 + el * sz);
      orc_vol(dg, starE + el * 243, NR ? starA + el * 54 : NULL, NR ? srcA + el * NR * 36 : NULL, t0 + el * sz, tA,
              dofsE + el * sz, dA);
      orc_surf_local(dg, fsE0 + el * 324, NR ? fsA0 + el * 216 : NULL, t0 + el * sz, dofsE + el * sz, dA);
    }
    free(der);
    free(derA);
    free(tA);
  }
}

void orc_neigh_lts(const orc_dg *dg, int64_t first, int64_t n_els, int first_ts, const int32_t *fa_type,
                   const int64_t *el_fa, const int64_t *el_fa_el, const uint16_t *f_id, const uint16_t *v_id,
                   const double *fsE1, const double *fsA1, const uint8_t *fa_lts, const double *t0, const double *t1,
                   const double *t2, double *dofsE, double *dofsA) {
  const int O = dg->order, N = nmodes(O), R = dg->n_crs, NR = dg->n_rms;
  const size_t sz = (size_t)NQ * N * R, szA = (size_t)NQA * N * R;
#pragma omp parallel
  {
    double *tn = (double *)malloc(sizeof(double) * sz);
    double *upA = (double *)malloc(sizeof(double) * szA);
#pragma omp for schedule(static)
    for (int64_t el = first; el < first + n_els; el++) {
      for (size_t i = 0; i < szA; i++) upA[i] = 0.0;
      for (int fa = 0; fa < 4; fa++) {
        const int64_t fid = el_fa[el * 4 + fa];
        const int32_t ty = fa_type[fid];
        if ((ty & FA_OUTFLOW) == FA_OUTFLOW) continue;
        const int fs = (ty & FA_FREE_SURFACE) == FA_FREE_SURFACE;
        const int64_t ne = fs ? el : el_fa_el[el * 4 + fa];
        unsigned vId = NO_ID16, fId = NO_ID16;
        if (!fs) {
          vId = v_id[el * 4 + fa];
          fId = f_id[el * 4 + fa];
        }
        /* AderDg.hpp:620-674: the neighbour's tDofs by the faces' time steps */
        const int mode = fa_lts[el * 4 + fa];
        const double *src = t0 + ne * sz;
        if (mode == 1) src = t1 + ne * sz;
        if (mode == 2 && first_ts) src = t2 + ne * sz;
        if (mode == 2 && !first_ts) {
          for (size_t i = 0; i < sz; i++) tn[i] = t0[ne * sz + i] - t2[ne * sz + i];
          src = tn;
        }
        orc_surf_neigh(dg, fa, vId, fId, fsE1 + (el * 4 + fa) * 81, NR ? fsA1 + (el * 4 + fa) * 54 : NULL, src,
                       dofsE + el * sz, upA);
      }
      /* :715-720 */
      if (NR > 0) orc_scatter_update_a(dg, upA, dofsA + (size_t)el * NR * szA);
    }
    free(tn);
    free(upA);
  }
}

/* -------------------------------------------------------------------------
 * Geometry: Jacobian, src/linalg/Mappings.hpp:189-230 with the tet4 shape
 * derivatives (jac[i][j] = dx_j / dxi_i); determinant and inverse.
 * ve: [4][3] vertex coordinates.
 * ---------------------------------------------------------------------- */
void orc_jac(const double *ve, double jac[9], double *det, double jinv[9]) {
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) jac[i * 3 + j] = ve[(i + 1) * 3 + j] - ve[j];
  const double *a = jac;
  double d = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
             a[2] * (a[3] * a[7] - a[4] * a[6]);
  *det = d;
  jinv[0] = (a[4] * a[8] - a[5] * a[7]) / d;
  jinv[1] = (a[2] * a[7] - a[1] * a[8]) / d;
  jinv[2] = (a[1] * a[5] - a[2] * a[4]) / d;
  jinv[3] = (a[5] * a[6] - a[3] * a[8]) / d;
  jinv[4] = (a[0] * a[8] - a[2] * a[6]) / d;
  jinv[5] = (a[2] * a[3] - a[0] * a[5]) / d;
  jinv[6] = (a[3] * a[7] - a[4] * a[6]) / d;
  jinv[7] = (a[1] * a[6] - a[0] * a[7]) / d;
  jinv[8] = (a[0] * a[4] - a[1] * a[3]) / d;
}

/* Elastic star matrices S_d = sum_k A_k jinv[k][d],
 * src/impl/seismic/setups/Elasticity.cpp:167-252 (Jacobians A, B, C of
 * q = (sxx, syy, szz, sxy, syz, sxz, u, v, w), q_t + A q_x + B q_y + C q_z = 0). */
void orc_star(double rho, double lam, double mu, const double jinv[9], double star[243]) {
  memset(star, 0, sizeof(double) * 243);
  const double lp = -lam - 2 * mu, l = -lam, m = -mu, ir = -1.0 / rho;
  /* (row, col, value) of A_k, k = 0..2 */
  const int rc[3][8][2] = {
      {{0, 6}, {1, 6}, {2, 6}, {3, 7}, {5, 8}, {6, 0}, {7, 3}, {8, 5}},
      {{0, 7}, {1, 7}, {2, 7}, {3, 6}, {4, 8}, {6, 3}, {7, 1}, {8, 4}},
      {{0, 8}, {1, 8}, {2, 8}, {4, 7}, {5, 6}, {6, 5}, {7, 4}, {8, 2}}};
  const double val[3][8] = {{lp, l, l, m, m, ir, ir, ir}, {l, lp, l, m, m, ir, ir, ir}, {l, l, lp, m, m, ir, ir, ir}};
  for (int d = 0; d < 3; d++)
    for (int k = 0; k < 3; k++)
      for (int e = 0; e < 8; e++) star[d * 81 + rc[k][e][0] * 9 + rc[k][e][1]] += val[k][e] * jinv[k * 3 + d];
}

/* Rotation of the quantities from face-aligned (n, s, t) to xyz
 * (src/impl/seismic/generated/Trafo3D.inc semantics) and its inverse,
 * computed here from the orthonormal frame. */
void orc_trafo(const double n[3], const double s[3], const double t[3], double T[81], double Ti[81]) {
  const double *f[3] = {n, s, t};
  /* stress index pairs in Voigt order (xx, yy, zz, xy, yz, xz) */
  const int vi[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {1, 2}, {0, 2}};
  memset(T, 0, sizeof(double) * 81);
  memset(Ti, 0, sizeof(double) * 81);
  for (int r = 0; r < 6; r++) {
    const int i = vi[r][0], j = vi[r][1];
    for (int c = 0; c < 6; c++) {
      const int a = vi[c][0], b = vi[c][1];
      /* sigma_ij = sum_ab R_ia R_jb sigma'_ab, R = [n s t] as columns */
      double v = f[a][i] * f[b][j];
      if (a != b) v += f[b][i] * f[a][j];
      T[r * 9 + c] = v;
      /* sigma'_ab = sum_ij R_ia R_jb sigma_ij */
      double w = f[vi[r][0]][vi[c][0]] * f[vi[r][1]][vi[c][1]];
      if (vi[c][0] != vi[c][1]) w += f[vi[r][0]][vi[c][1]] * f[vi[r][1]][vi[c][0]];
      Ti[r * 9 + c] = w;
    }
  }
  for (int i = 0; i < 3; i++)
    for (int a = 0; a < 3; a++) {
      T[(6 + i) * 9 + 6 + a] = f[a][i];
      Ti[(6 + a) * 9 + 6 + i] = f[a][i];
    }
}

/* Left / right contributions to the exact Riemann solution in face-aligned
 * coordinates (the role of generated/FlMid3D.inc and FrMid3D.inc): the flux
 * of the left element A_L q*, with q* the middle state of the welded-interface
 * Riemann problem, differentiated w.r.t. the left and right face states. */
void orc_fs_mid(double rhoL, double rhoR, double lamL, double lamR, double muL, double muR,
                double flMid[81], double frMid[81]) {
  memset(flMid, 0, sizeof(double) * 81);
  memset(frMid, 0, sizeof(double) * 81);
  const double zpL = sqrt((lamL + 2 * muL) * rhoL), zpR = sqrt((lamR + 2 * muR) * rhoR);
  const double zsL = sqrt(muL * rhoL), zsR = sqrt(muR * rhoR);
  /* normal: u* = (sR - sL + zpR uR + zpL uL)/(zpL+zpR),
   *         s* = (zpR sL + zpL sR + zpL zpR (uR - uL))/(zpL+zpR) */
  const double ip = 1.0 / (zpL + zpR);
  const double du_dsL = -ip, du_dsR = ip, du_duL = zpL * ip, du_duR = zpR * ip;
  const double ds_dsL = zpR * ip, ds_dsR = zpL * ip, ds_duL = -zpL * zpR * ip, ds_duR = zpL * zpR * ip;
  /* flux rows: F[snn] = -(lamL+2muL) u*, F[sss] = F[stt] = -lamL u*, F[un] = -(s*) / rhoL */
  const double cu[3] = {-(lamL + 2 * muL), -lamL, -lamL};
  for (int r = 0; r < 3; r++) {
    flMid[r * 9 + 0] = cu[r] * du_dsL;
    flMid[r * 9 + 6] = cu[r] * du_duL;
    frMid[r * 9 + 0] = cu[r] * du_dsR;
    frMid[r * 9 + 6] = cu[r] * du_duR;
  }
  flMid[6 * 9 + 0] = -ds_dsL / rhoL;
  flMid[6 * 9 + 6] = -ds_duL / rhoL;
  frMid[6 * 9 + 0] = -ds_dsR / rhoL;
  frMid[6 * 9 + 6] = -ds_duR / rhoL;
  /* shear pairs (sns, us) = (3, 7) and (snt, ut) = (5, 8) */
  const double is = 1.0 / (zsL + zsR);
  const int pr[2][2] = {{3, 7}, {5, 8}};
  for (int p = 0; p < 2; p++) {
    const int si = pr[p][0], vi = pr[p][1];
    flMid[si * 9 + si] = -muL * (-is);
    flMid[si * 9 + vi] = -muL * (zsL * is);
    frMid[si * 9 + si] = -muL * (is);
    frMid[si * 9 + vi] = -muL * (zsR * is);
    flMid[vi * 9 + si] = -(zsR * is) / rhoL;
    flMid[vi * 9 + vi] = -(-zsL * zsR * is) / rhoL;
    frMid[vi * 9 + si] = -(zsL * is) / rhoL;
    frMid[vi * 9 + vi] = -(zsL * zsR * is) / rhoL;
  }
}

static void mm9(const double *a, const double *b, double *c) {
  for (int i = 0; i < 9; i++)
    for (int j = 0; j < 9; j++) {
      double s = 0;
      for (int k = 0; k < 9; k++) s += a[i * 9 + k] * b[k * 9 + j];
      c[i * 9 + j] = s;
    }
}

/* Elastic flux solvers of one face side, src/impl/seismic/setups/Elasticity.cpp:372-441
 * (fsL = T FlMid T^-1; fsR = T FrMid [mirror] T^-1; the free surface mirrors
 * the face-aligned normal tractions, columns 0, 3, 5). */
void orc_fs(double rhoL, double rhoR, double lamL, double lamR, double muL, double muR, const double n[3],
            const double s[3], const double t[3], int free_surface, double fsL[81], double fsR[81]) {
  double T[81], Ti[81], fl[81], fr[81], tmp[81];
  orc_trafo(n, s, t, T, Ti);
  orc_fs_mid(rhoL, rhoR, lamL, lamR, muL, muR, fl, fr);
  mm9(T, fl, tmp);
  mm9(tmp, Ti, fsL);
  mm9(T, fr, tmp);
  if (free_surface)
    for (int r = 0; r < 9; r++) {
      tmp[r * 9 + 0] *= -1;
      tmp[r * 9 + 3] *= -1;
      tmp[r * 9 + 5] *= -1;
    }
  mm9(tmp, Ti, fsR);
}

/* Per-face set up of the flux solvers of both adjacent elements,
 * src/impl/seismic/solvers/AderDgInit.hpp:412-665 (GTS, single partition):
 * the left element uses the outward normal n, the right one -n with the
 * same tangents; both are scaled by -2 area / |J|. */
void orc_init_fs(int64_t n_fas, const int64_t *fa_el /*[fa][2]*/, const int64_t *el_fa /*[el][4]*/,
                 const int32_t *fa_type, const double *normals /*[fa][3]*/,
                 const double *tangents /*[fa][2][3]*/, const double *areas, const double *jdet /*[el]*/,
                 const double *rho, const double *lam, const double *mu, double *fsE0, double *fsE1) {
  for (int64_t fa = 0; fa < n_fas; fa++) {
    const int64_t eL = fa_el[fa * 2], eR = fa_el[fa * 2 + 1];
    const int exR = eR >= 0;
    int fL = -1, fR = -1;
    for (int f = 0; f < 4; f++) {
      if (el_fa[eL * 4 + f] == fa) fL = f;
      if (exR && el_fa[eR * 4 + f] == fa) fR = f;
    }
    const int fs = (fa_type[fa] & FA_FREE_SURFACE) == FA_FREE_SURFACE;
    const double rL = rho[eL], lL = lam[eL], mL = mu[eL];
    const double rR = exR ? rho[eR] : rL, lR = exR ? lam[eR] : lL, mR = exR ? mu[eR] : mL;
    const double *n = normals + fa * 3, *s = tangents + fa * 6, *t = tangents + fa * 6 + 3;
    double a[81], b[81];
    if (fL >= 0) {
      orc_fs(rL, rR, lL, lR, mL, mR, n, s, t, fs, a, b);
      const double sc = -2.0 * areas[fa] / jdet[eL];
      for (int i = 0; i < 81; i++) {
        fsE0[(eL * 4 + fL) * 81 + i] = a[i] * sc;
        fsE1[(eL * 4 + fL) * 81 + i] = b[i] * sc;
      }
    }
    if (exR && fR >= 0) {
      const double nm[3] = {-n[0], -n[1], -n[2]};
      orc_fs(rR, rL, lR, lL, mR, mL, nm, s, t, 0, a, b);
      const double sc = -2.0 * areas[fa] / jdet[eR];
      for (int i = 0; i < 81; i++) {
        fsE0[(eR * 4 + fR) * 81 + i] = a[i] * sc;
        fsE1[(eR * 4 + fR) * 81 + i] = b[i] * sc;
      }
    }
  }
}

/* Element time step by CFL, src/impl/seismic/common.hpp:84-98, with the
 * reference's scale 1/2 (const.inc:135-137). */
/* -------------------------------------------------------------------------
 * Viscoelastic set-up, src/impl/seismic/setups/ViscoElasticity.{h,cpp}.
 * ---------------------------------------------------------------------- */
/* 2n-1 relaxation frequencies log-spaced around 2 pi f_cen over the ratio
 * f_rat, ViscoElasticity.h:169-182. */
void orc_visco_frequencies(int n_mechs, double f_cen, double f_rat, double *freqs) {
  freqs[0] = 2 * M_PI * f_cen / sqrt(f_rat);
  for (int fr = 1; fr < 2 * n_mechs - 1; fr++)
    freqs[fr] = exp(log(freqs[0]) + (fr * log(f_rat)) / (2 * (n_mechs - 1)));
}

/* Least-squares fit of the anelastic coefficients for a constant quality
 * factor q over the 2n-1 frequencies, ViscoElasticity.cpp:69-128 (the
 * reference solves with an SVD; the full-rank least-squares solution is
 * unique, here by Householder QR). */
void orc_anelastic_coeff_ps(int n_mechs, double q, const double *freqs, double *coeffs) {
  const int m = 2 * n_mechs - 1, n = n_mechs;
  const double qi = 1.0 / q;
  double A[64], b[16];
  for (int fr = 0; fr < m; fr++) {
    for (int me = 0; me < n; me++) {
      const double mf = freqs[me * 2];
      A[fr * n + me] = (mf * freqs[fr] + mf * mf * qi) / (freqs[fr] * freqs[fr] + mf * mf);
    }
    b[fr] = qi;
  }
  for (int k = 0; k < n; k++) { /* Householder reflection of column k */
    double nr = 0;
    for (int i = k; i < m; i++) nr += A[i * n + k] * A[i * n + k];
    nr = sqrt(nr);
    const double alpha = A[k * n + k] > 0 ? -nr : nr;
    double v[16];
    for (int i = 0; i < m; i++) v[i] = i < k ? 0 : A[i * n + k];
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
  for (int k = n - 1; k >= 0; k--) { /* R y = Q^T b */
    double r = b[k];
    for (int j = k + 1; j < n; j++) r -= A[k * n + j] * coeffs[j];
    coeffs[k] = r / A[k * n + k];
  }
}

/* Unrelaxed (elastic) Lame parameters from the phase-velocity ones at the
 * reference frequency, ViscoElasticity.cpp:30-67. */
void orc_elastic_lame(int n_mechs, double f_ref, const double *freqs, const double *c_kappa,
                      const double *c_mu, double lam_phase, double mu_phase, double *lam_e, double *mu_e) {
  double th1[2] = {1, 1}, th2[2] = {0, 0};
  for (int me = 0; me < n_mechs; me++) {
    const double rat = f_ref / freqs[me * 2], sca = 1.0 / (1 + rat * rat);
    th1[0] -= c_kappa[me] * sca;
    th1[1] -= c_mu[me] * sca;
    th2[0] += c_kappa[me] * rat * sca;
    th2[1] += c_mu[me] * rat * sca;
  }
  double r[2];
  for (int i = 0; i < 2; i++) r[i] = sqrt(th1[i] * th1[i] + th2[i] * th2[i]);
  *mu_e = mu_phase * (r[1] + th1[1]) / (2 * r[1] * r[1]);
  const double kappa = (lam_phase + 2 * mu_phase) * (r[0] + th1[0]) / (2 * r[0] * r[0]);
  *lam_e = kappa - 2 * *mu_e;
}

/* Anelastic coefficients in Lame form (lambda part, then mu part) and the
 * elastic Lame parameters, ViscoElasticity.cpp:130-176. */
void orc_anelastic_coeff_lame(int n_mechs, double f_cen, double f_rat, double qp, double qs, double lam_phase,
                              double mu_phase, double *lam_e, double *mu_e, double *coeffs /*[2n]*/) {
  double freqs[16];
  orc_visco_frequencies(n_mechs, f_cen, f_rat, freqs);
  orc_anelastic_coeff_ps(n_mechs, qp, freqs, coeffs);
  orc_anelastic_coeff_ps(n_mechs, qs, freqs, coeffs + n_mechs);
  orc_elastic_lame(n_mechs, 2 * M_PI * f_cen, freqs, coeffs, coeffs + n_mechs, lam_phase, mu_phase, lam_e, mu_e);
  for (int me = 0; me < n_mechs; me++) {
    const double sca = 2.0 * *mu_e / *lam_e;
    coeffs[me] = (1.0 + sca) * coeffs[me] - sca * coeffs[n_mechs + me];
  }
}

/* Anelastic source matrices acting on the stresses, [n][6][6] per element,
 * and the elastic Lame parameters, ViscoElasticity.h:380-458. */
void orc_src_a(int n_mechs, double f_cen, double f_rat, double qp, double qs, double lam_phase, double mu_phase,
               double *lam_e, double *mu_e, double *src /*[n][36]*/) {
  double co[32];
  orc_anelastic_coeff_lame(n_mechs, f_cen, f_rat, qp, qs, lam_phase, mu_phase, lam_e, mu_e, co);
  for (int me = 0; me < n_mechs; me++) {
    double *o = src + me * 36;
    for (int i = 0; i < 36; i++) o[i] = 0;
    const double d = -*lam_e * co[me] - 2 * *mu_e * co[n_mechs + me];
    const double sh = -2 * *mu_e * co[n_mechs + me];
    const double off = -*lam_e * co[me];
    for (int i = 0; i < 3; i++) {
      o[i * 6 + i] = d;
      o[(3 + i) * 6 + 3 + i] = sh;
      for (int j = 0; j < 3; j++)
        if (j != i) o[i * 6 + j] = off;
    }
  }
}

/* Anelastic star matrices [3][6][3] (rows: memory variables, columns:
 * velocities), ViscoElasticity.cpp:233-316. */
void orc_star_a(const double jinv[9], double star_a[54]) {
  for (int i = 0; i < 54; i++) star_a[i] = 0;
  for (int d = 0; d < 3; d++) {
    double *o = star_a + d * 18;
    const double j0 = jinv[0 * 3 + d], j1 = jinv[1 * 3 + d], j2 = jinv[2 * 3 + d];
    o[0 * 3 + 0] += -1.0 * j0;
    o[3 * 3 + 1] += -0.5 * j0;
    o[5 * 3 + 2] += -0.5 * j0;
    o[1 * 3 + 1] += -1.0 * j1;
    o[3 * 3 + 0] += -0.5 * j1;
    o[4 * 3 + 2] += -0.5 * j1;
    o[2 * 3 + 2] += -1.0 * j2;
    o[4 * 3 + 1] += -0.5 * j2;
    o[5 * 3 + 0] += -0.5 * j2;
  }
}

/* Middle-state part of the anelastic flux solvers in face coordinates,
 * [6][9] each, ViscoElasticity.cpp:376-441. */
void orc_fs_mid_a(double rL, double rR, double lL, double lR, double mL, double mR, double midL[54],
                  double midR[54]) {
  const double cpL = sqrt((lL + 2 * mL) / rL), cpR = sqrt((lR + 2 * mR) / rR);
  const double csL = sqrt(mL / rL), csR = sqrt(mR / rR);
  for (int i = 0; i < 54; i++) midL[i] = midR[i] = 0;
  const double dp = lL * cpR + lR * cpL + 2 * mL * cpR + 2 * mR * cpL;
  const double ds = 2 * mL * csR + 2 * mR * csL;
  midL[0 * 9 + 0] = cpL * cpR / dp;
  midL[0 * 9 + 6] = -cpR * (lL + 2 * mL) / dp;
  midL[3 * 9 + 3] = midL[5 * 9 + 5] = csL * csR / ds;
  midL[3 * 9 + 7] = midL[5 * 9 + 8] = -mL * csR / ds;
  midR[0 * 9 + 0] = -cpL * cpR / dp;
  midR[0 * 9 + 6] = -cpL * (lR + 2 * mR) / dp;
  midR[3 * 9 + 3] = midR[5 * 9 + 5] = -csL * csR / ds;
  midR[3 * 9 + 7] = midR[5 * 9 + 8] = -mR * csL / ds;
}

/* Anelastic flux solvers [6][9] of both sides in global coordinates:
 * T_A mid T^-1 with T_A the stress block of T, the right state mirrored
 * on a free surface, ViscoElasticity.cpp:505-570 and
 * solvers/AderDgInit.hpp:288-370. */
void orc_fs_a(double rL, double rR, double lL, double lR, double mL, double mR, const double n[3],
              const double s[3], const double t[3], int free_surface, double fsAl[54], double fsAr[54]) {
  double T[81], Ti[81], mid[2][54], tmp[2][54];
  orc_trafo(n, s, t, T, Ti);
  orc_fs_mid_a(rL, rR, lL, lR, mL, mR, mid[0], mid[1]);
  for (int sd = 0; sd < 2; sd++)
    for (int r = 0; r < 6; r++)
      for (int c = 0; c < 9; c++) {
        double v = 0;
        for (int k = 0; k < 6; k++) v += T[r * 9 + k] * mid[sd][k * 9 + c];
        tmp[sd][r * 9 + c] = v;
      }
  if (free_surface)
    for (int r = 0; r < 6; r++) {
      tmp[1][r * 9 + 0] *= -1;
      tmp[1][r * 9 + 3] *= -1;
      tmp[1][r * 9 + 5] *= -1;
    }
  for (int sd = 0; sd < 2; sd++) {
    double *o = sd == 0 ? fsAl : fsAr;
    for (int r = 0; r < 6; r++)
      for (int c = 0; c < 9; c++) {
        double v = 0;
        for (int k = 0; k < 9; k++) v += tmp[sd][r * 9 + k] * Ti[k * 9 + c];
        o[r * 9 + c] = v;
      }
  }
}

/* Anelastic flux solvers of all element faces, [el][4][54] per side, with
 * the elastic solvers' scaling -2 area / |J| (AderDgInit.hpp:617-663). */
void orc_init_fs_a(int64_t n_fas, const int64_t *fa_el, const int64_t *el_fa, const int32_t *fa_type,
                   const double *normals, const double *tangents, const double *areas, const double *jdet,
                   const double *rho, const double *lam, const double *mu, double *fsA0, double *fsA1) {
  for (int64_t fa = 0; fa < n_fas; fa++) {
    const int64_t eL = fa_el[fa * 2], eR = fa_el[fa * 2 + 1];
    const int exR = eR >= 0;
    int fL = -1, fR = -1;
    for (int f = 0; f < 4; f++) {
      if (el_fa[eL * 4 + f] == fa) fL = f;
      if (exR && el_fa[eR * 4 + f] == fa) fR = f;
    }
    const int fs = (fa_type[fa] & FA_FREE_SURFACE) == FA_FREE_SURFACE;
    const double rL = rho[eL], lL = lam[eL], mL = mu[eL];
    const double rR = exR ? rho[eR] : rL, lR = exR ? lam[eR] : lL, mR = exR ? mu[eR] : mL;
    const double *n = normals + fa * 3, *s = tangents + fa * 6, *t = tangents + fa * 6 + 3;
    double a[54], b[54];
    if (fL >= 0) {
      orc_fs_a(rL, rR, lL, lR, mL, mR, n, s, t, fs, a, b);
      const double sc = -2.0 * areas[fa] / jdet[eL];
      for (int i = 0; i < 54; i++) {
        fsA0[(eL * 4 + fL) * 54 + i] = a[i] * sc;
        fsA1[(eL * 4 + fL) * 54 + i] = b[i] * sc;
      }
    }
    if (exR && fR >= 0) {
      const double nm[3] = {-n[0], -n[1], -n[2]};
      orc_fs_a(rR, rL, lR, lL, mR, mL, nm, s, t, 0, a, b);
      const double sc = -2.0 * areas[fa] / jdet[eR];
      for (int i = 0; i < 54; i++) {
        fsA0[(eR * 4 + fR) * 54 + i] = a[i] * sc;
        fsA1[(eR * 4 + fR) * 54 + i] = b[i] * sc;
      }
    }
  }
}

/* -------------------------------------------------------------------------
 * Dynamic rupture: linear slip-weakening friction
 * (src/impl/seismic/solvers/FrictionLaws.hpp:309-456) and the sub-face net
 * updates of the internal boundary (InternalBoundary.hpp:184-280) with the
 * face solvers of InternalBoundarySolvers<TET4,O,3>::init (:743-923).
 * ---------------------------------------------------------------------- */
/* Middle state of the left side from the jump qR - qL in face coordinates:
 * ms = qL + J (qR - qL), the waves travelling into the left medium
 * (P: sigma_nn, u_n; S: sigma_ns / sigma_nt, u_s / u_t; the tangential
 * normal stresses follow the P wave), as generated/MiddleStateJumpL3D.inc. */
void orc_ms_jump_l(double lamL, double lamR, double muL, double muR, double rhoL, double rhoR, double J[81]) {
  const double zpL = sqrt((lamL + 2 * muL) * rhoL), zpR = sqrt((lamR + 2 * muR) * rhoR);
  const double zsL = sqrt(muL * rhoL), zsR = sqrt(muR * rhoR);
  const double dp = zpL + zpR, ds = zsL + zsR, lr = lamL / (lamL + 2 * muL);
  for (int i = 0; i < 81; i++) J[i] = 0;
  J[0 * 9 + 0] = zpL / dp;
  J[0 * 9 + 6] = zpL * zpR / dp;
  for (int r = 1; r < 3; r++) {
    J[r * 9 + 0] = lr * J[0];
    J[r * 9 + 6] = lr * J[6];
  }
  J[3 * 9 + 3] = J[5 * 9 + 5] = zsL / ds;
  J[3 * 9 + 7] = J[5 * 9 + 8] = zsL * zsR / ds;
  J[6 * 9 + 0] = 1 / dp;
  J[6 * 9 + 6] = zpR / dp;
  J[7 * 9 + 3] = J[8 * 9 + 5] = 1 / ds;
  J[7 * 9 + 7] = J[8 * 9 + 8] = zsR / ds;
}

/* Middle state of the right side: ms = qR + J (qR - qL), the waves
 * travelling into the right medium, as generated/MiddleStateJumpR3D.inc. */
void orc_ms_jump_r(double lamL, double lamR, double muL, double muR, double rhoL, double rhoR, double J[81]) {
  const double zpL = sqrt((lamL + 2 * muL) * rhoL), zpR = sqrt((lamR + 2 * muR) * rhoR);
  const double zsL = sqrt(muL * rhoL), zsR = sqrt(muR * rhoR);
  const double dp = zpL + zpR, ds = zsL + zsR, lr = lamR / (lamR + 2 * muR);
  for (int i = 0; i < 81; i++) J[i] = 0;
  J[0 * 9 + 0] = -zpR / dp;
  J[0 * 9 + 6] = zpL * zpR / dp;
  for (int r = 1; r < 3; r++) {
    J[r * 9 + 0] = lr * J[0];
    J[r * 9 + 6] = lr * J[6];
  }
  J[3 * 9 + 3] = J[5 * 9 + 5] = -zsR / ds;
  J[3 * 9 + 7] = J[5 * 9 + 8] = zsL * zsR / ds;
  J[6 * 9 + 0] = 1 / dp;
  J[6 * 9 + 6] = -zpL / dp;
  J[7 * 9 + 3] = J[8 * 9 + 5] = 1 / ds;
  J[7 * 9 + 7] = J[8 * 9 + 8] = -zsL / ds;
}

/* Flux of a face-coordinate middle state in physical coordinates: T A_n,
 * A_n the normal Jacobian in face coordinates, as
 * generated/MiddleStateFlux3D.inc. */
void orc_ms_flux(const double n[3], const double s[3], const double t[3], double lam, double mu, double rho,
                 double F[81]) {
  double T[81], Ti[81], A[81];
  orc_trafo(n, s, t, T, Ti);
  for (int i = 0; i < 81; i++) A[i] = 0;
  A[0 * 9 + 6] = -(lam + 2 * mu);
  A[1 * 9 + 6] = A[2 * 9 + 6] = -lam;
  A[3 * 9 + 7] = A[5 * 9 + 8] = -mu;
  A[6 * 9 + 0] = A[7 * 9 + 3] = A[8 * 9 + 5] = -1 / rho;
  mm9(T, A, F);
}

/* Rotation taking the direction of v0 to that of v1 (Rodrigues),
 * src/linalg/Geom.hpp:939-989. */
void orc_rot_mat(const double v0[3], const double v1[3], double R[9]) {
  double a[3], b[3], na = 0, nb = 0;
  for (int i = 0; i < 3; i++) {
    na += v0[i] * v0[i];
    nb += v1[i] * v1[i];
  }
  na = sqrt(na);
  nb = sqrt(nb);
  for (int i = 0; i < 3; i++) {
    a[i] = v0[i] / na;
    b[i] = v1[i] / nb;
  }
  const double c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
  const double sm[9] = {0, -c[2], c[1], c[2], 0, -c[0], -c[1], c[0], 0};
  const double sca = 1.0 / (1.0 + a[0] * b[0] + a[1] * b[1] + a[2] * b[2]);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double v = 0;
      for (int k = 0; k < 3; k++) v += sm[i * 3 + k] * sm[k * 3 + j];
      R[i * 3 + j] = v * sca + sm[i * 3 + j] + (i == j ? 1.0 : 0.0);
    }
}

/* Solvers of one rupture face [4][81]: T^-1 of the face frame, the middle
 * state jump solver, the flux solvers of the left and right elements scaled
 * by (area / n_sf) / (volume / n_sc); with a boundary frame bnd [3][3] the
 * face frame follows the boundary's normal orientation and tangents
 * (InternalBoundary.hpp:743-923).  Returns 1 if the face normal points in
 * the boundary's preferred direction. */
int orc_ib_solvers(const double normal[3], const double tangent0[3], const double tangent1[3],
                   const double *bnd /*[3][3] or NULL*/, double rhoL, double lamL, double muL, double rhoR,
                   double lamR, double muR, double area, double volL, double volR, int n_sf, int n_sc,
                   double out[4 * 81]) {
  double f[3][3];
  for (int d = 0; d < 3; d++) {
    f[0][d] = normal[d];
    f[1][d] = tangent0[d];
    f[2][d] = tangent1[d];
  }
  int pref = 1;
  if (bnd) {
    pref = (bnd[0] * f[0][0] + bnd[1] * f[0][1] + bnd[2] * f[0][2]) > 0;
    if (!pref)
      for (int d = 0; d < 3; d++) f[0][d] = -f[0][d];
    double R[9];
    orc_rot_mat(bnd, f[0], R);
    for (int k = 1; k < 3; k++)
      for (int i = 0; i < 3; i++) {
        double v = 0;
        for (int j = 0; j < 3; j++) v += R[i * 3 + j] * bnd[k * 3 + j];
        f[k][i] = v;
      }
  }
  double T[81];
  orc_trafo(f[0], f[1], f[2], T, out);
  if (pref) {
    orc_ms_jump_l(lamL, lamR, muL, muR, rhoL, rhoR, out + 81);
  } else {
    orc_ms_jump_r(lamR, lamL, muR, muL, rhoR, rhoL, out + 81);
    for (int i = 0; i < 81; i++) out[81 + i] = -out[81 + i];
  }
  orc_ms_flux(f[0], f[1], f[2], lamL, muL, rhoL, out + 162);
  orc_ms_flux(f[0], f[1], f[2], lamR, muR, rhoR, out + 243);
  for (int sd = 0; sd < 2; sd++) {
    double sca = (area / n_sf) / ((sd == 0 ? volL : volR) / n_sc);
    if (!pref) sca = -sca;
    for (int i = 0; i < 81; i++) out[(2 + sd) * 81 + i] *= sca;
  }
  return pref;
}

/* Linear slip weakening for one sub-face and run: strength, failure,
 * perturbed minus / plus middle states, slip rate, slip and friction
 * coefficient (FrictionLaws.hpp:309-392).  state: sn0, co0, ss0[2], ss0A,
 * muf, st, tr[2], sr[2], dd[2] (14 values). */
void orc_lsw(double dt, double csDmuM, double csDmuP, double mus, double mud, double dcInv, double *st_,
             const double ms[9], double msM[9], double msP[9], int *per) {
  double *sn0 = st_ + 0, *co0 = st_ + 1, *ss0 = st_ + 2, *muf = st_ + 5, *st = st_ + 6, *tr = st_ + 7,
         *sr = st_ + 9, *dd = st_ + 11;
  for (int q = 0; q < 9; q++) msM[q] = msP[q] = ms[q];
  double strength = *muf * (*sn0 + ms[0]) + *co0;
  strength = strength < 0 ? -strength : 0; /* only compression holds the fault */
  *st = strength;
  const double s1 = ss0[0] + ms[3], s2 = ss0[1] + ms[5];
  const double shear = sqrt(s1 * s1 + s2 * s2);
  *per = shear > strength;
  const double scale = strength / shear;
  tr[0] = *per ? s1 * scale - ss0[0] : ms[3];
  tr[1] = *per ? s2 * scale - ss0[1] : ms[5];
  msM[3] = msP[3] = tr[0];
  msM[5] = msP[5] = tr[1];
  const double d1 = tr[0] - ms[3], d2 = tr[1] - ms[5];
  msM[7] -= csDmuM * d1;
  msM[8] -= csDmuM * d2;
  msP[7] += csDmuP * d1;
  msP[8] += csDmuP * d2;
  sr[0] = msP[7] - msM[7];
  sr[1] = msP[8] - msM[8];
  dd[0] += fabs(sr[0]) * dt;
  dd[1] += fabs(sr[1]) * dt;
  const double dda = sqrt(dd[0] * dd[0] + dd[1] * dd[1]);
  const double arg = mus - (mus - mud) * dcInv * dda;
  *muf = arg > mud ? arg : mud;
  tr[0] += ss0[0];
  tr[1] += ss0[1];
}

/* Net updates of one rupture face for n_sf sub-faces and R runs
 * (InternalBoundary.hpp:184-280 with FrictionLaws<3,R>::perturb):
 * solv [4][81] from orc_ib_solvers, dofs [9][n_sf][R] of the two sides'
 * face-adjacent sub-cells, l_eq_m: the left element is the minus side;
 * glob [3][R]: mus, mud, 1/Dc; state [n_sf][R][14] (see orc_lsw). */
void orc_net_updates(int n_sf, int R, const double *solv, const double *dofsL, const double *dofsR, double dt,
                     int l_eq_m, double csDmuM, double csDmuP, const double *glob, double *state, double *netL,
                     double *netR, int *per) {
  const double *tm1 = solv, *J = solv + 81, *FL = solv + 162, *FR = solv + 243;
  const size_t sz = (size_t)9 * n_sf * R;
  double *rot = (double *)malloc(sizeof(double) * 2 * sz), *ms2 = (double *)malloc(sizeof(double) * 2 * sz);
  for (int sd = 0; sd < 2; sd++) {
    const double *d = sd == 0 ? dofsL : dofsR;
    for (int q = 0; q < 9; q++)
      for (int i = 0; i < n_sf * R; i++) {
        double v = 0;
        for (int p = 0; p < 9; p++) v += tm1[q * 9 + p] * d[p * n_sf * R + i];
        rot[sd * sz + q * n_sf * R + i] = v;
      }
  }
  for (int sf = 0; sf < n_sf; sf++)
    for (int r = 0; r < R; r++) {
      const int i = sf * R + r;
      double qL[9], jmp[9], ms[9], msA[9], msB[9];
      for (int q = 0; q < 9; q++) {
        qL[q] = rot[q * n_sf * R + i];
        jmp[q] = rot[sz + q * n_sf * R + i] - qL[q];
      }
      for (int q = 0; q < 9; q++) {
        double v = qL[q];
        for (int p = 0; p < 9; p++) v += J[q * 9 + p] * jmp[p];
        ms[q] = v;
      }
      int pf;
      orc_lsw(dt, csDmuM, csDmuP, glob[r], glob[R + r], glob[2 * R + r], state + (size_t)i * 14, ms, msA, msB,
              &pf);
      per[i] = pf;
      /* minus / plus to left / right */
      const double *mL = l_eq_m ? msA : msB, *mR = l_eq_m ? msB : msA;
      for (int q = 0; q < 9; q++) {
        ms2[q * n_sf * R + i] = -mL[q] * dt;
        ms2[sz + q * n_sf * R + i] = mR[q] * dt;
      }
    }
  for (int sd = 0; sd < 2; sd++) {
    const double *F = sd == 0 ? FL : FR;
    double *o = sd == 0 ? netL : netR;
    for (int q = 0; q < 9; q++)
      for (int i = 0; i < n_sf * R; i++) {
        double v = 0;
        for (int p = 0; p < 9; p++) v += F[q * 9 + p] * ms2[sd * sz + p * n_sf * R + i];
        o[q * n_sf * R + i] = v;
      }
  }
  free(rot);
  free(ms2);
}

/* Integral over [x1, x2] of a series sampled at x0 + i dx (i < n), linear
 * between samples; where [x1, x2] leaves the samples the series takes the
 * value na.  Series<1>::integrate, src/linalg/Series.hpp:59-133 (FP64
 * tolerance 1e-9, src/constants.hpp:289-300). */
double orc_series_integrate(double dx, double x0, long n, const double *v, double x1, double x2, double na) {
  const double tol = 1e-9;
  double a = (x1 - x0 + dx) / dx, b = (x2 - x0) / dx;
  size_t first = (size_t)(a > 0 ? a : 0), last = (size_t)(b > 0 ? b : 0);
  if (last > (size_t)(n - 1)) last = (size_t)(n - 1);
  if (first > last) first = last;
  double s = 0;
  for (size_t i = first + 1; i <= last; i++) s += 0.5 * dx * (v[i] - v[i - 1]) + v[i - 1] * dx;
  const double t_end = x0 + (n - 1) * dx;
  const int in1 = x1 + tol > x0 && x1 - tol < t_end, in2 = x2 + tol > x0 && x2 - tol < t_end;
  double d = x0 + first * dx - x1; /* the part before the first covered sample */
  if (first > 0 && in1)
    s += 0.5 * d * (d / dx) * (v[first - 1] - v[first]) + v[first] * d;
  else
    s += na * d;
  d = x2 - (x0 + last * dx); /* the part after the last covered sample */
  if (last < (size_t)(n - 1) && in2)
    s += 0.5 * d * (d / dx) * (v[last + 1] - v[last]) + v[last] * d;
  else
    s += na * d;
  return s;
}

/* Point sources of one step (PointSources::apply, src/impl/seismic/setups/
 * PointSources.hpp:887-930): for every source (element el, run, basis
 * evaluations scaled by the inverse mass and 1/|J| bev[N], quantity
 * scalings sca[9], series ts[ptr[i]..ptr[i+1]) sampled from t0 with dts):
 * dofs[el][q][md][run] += int_{t1}^{t2} s(t) dt * sca[q] * bev[md]. */
void orc_src_apply(int n_modes, int n_runs, long n_src, const long *el, const int *run, const double *t0,
                   const double *dts, const double *bev, const double *sca, const long *ptr, const double *ts,
                   double t1, double t2, double *dofs) {
  for (long i = 0; i < n_src; i++) {
    const double in = orc_series_integrate(dts[i], t0[i], ptr[i + 1] - ptr[i], ts + ptr[i], t1, t2, 0.0);
    for (int q = 0; q < NQ; q++)
      for (int md = 0; md < n_modes; md++)
        dofs[((size_t)el[i] * NQ + q) * n_modes * n_runs + (size_t)md * n_runs + run[i]] +=
            in * sca[i * NQ + q] * bev[i * n_modes + md];
  }
}

double orc_dt_cfl(double rho, double lam, double mu, double in_dia, int order) {
  return in_dia * 0.5 / sqrt((lam + 2 * mu) / rho) / (2.0 * order - 1);
}
