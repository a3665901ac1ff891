/*
 * edge_amd: MI355X-native ADER-DG element update for elastic tet4 meshes.
 *
 * C ABI of libedge_amd.so.  Plain pointers and sizes only.  Device pointers
 * are HIP device allocations owned by the caller; host pointers are ordinary
 * host memory.  Every entry point returns EDGE_OK (0) or an error code;
 * edge_last_error() returns a message for the calling thread's last failure.
 *
 * Layouts at the boundary are the reference's:
 *   elastic DOFs / time-integrated DOFs  [el][9][mode][run]   (run innermost)
 *   star matrices                        [el][3][9*9]
 *   flux solvers                         [2][el][4][9*9]      (0: local, 1: neighbour)
 * with mode counts N = O(O+1)(O+2)/6 and face modes NF = O(O+1)/2 for order O.
 */
#ifndef EDGE_AMD_H
#define EDGE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define EDGE_OK 0
#define EDGE_ERR_ARG 1
#define EDGE_ERR_HIP 2
#define EDGE_ERR_MESH 3
#define EDGE_ERR_UNSUPPORTED 4
#define EDGE_ERR_NOMEM 5

/* ---- sentinels and face types (values of the reference's seismic spType,
 *      src/impl/seismic/const.inc:143-148; "no id" as in
 *      src/impl/seismic/solvers/AderDg.hpp:692-698) ---- */
#define EDGE_NO_ID16 0xFFFFu
#define EDGE_FA_INTERIOR 0
#define EDGE_FA_FREE_SURFACE 101
#define EDGE_FA_OUTFLOW 105
#define EDGE_FA_PERIODIC 106
#define EDGE_FA_RUPTURE 201

/* Message of the calling thread's last error ("" if none). */
const char *edge_last_error(void);

/* Library version string. */
const char *edge_version(void);

/* ====================================================================== */
/* Reference-element matrices (tet4), replaces edge::dg::Basis /           */
/* dg::{TimePred,VolInt,SurfInt}::store*Dense fed from edge::pre::dg::g_*  */
/* (src/dg/Basis.h:34-62, :344-401; src/dg/TimePred.hpp:69-127,            */
/* src/dg/VolInt.hpp:63-118, src/dg/SurfInt.hpp:76-170).                   */
/* ====================================================================== */

/* Mode counts for order O (= polynomial degree + 1). */
int edge_dg_sizes(int order, int *n_modes, int *n_face_modes);

/* Copies the matrices of order O into host arrays (any pointer may be NULL):
 *   mass[N] (diagonal), stiffT[3][N][N] = K_d^T M^-1 (not negated),
 *   stiffV[3][N][N] = K_d M^-1, fluxL[4][N][NF], fluxN[12][N][NF],
 *   fluxT[4][NF][N]. */
int edge_dg_matrices(int order, double *mass, double *stiff_t, double *stiff_v, double *flux_l,
                     double *flux_n, double *flux_t);

/* Evaluates the N modal basis functions of order O at reference points
 * pts[n][3] (xi in the unit tet): val[n][N].  (src/dg/Basis.cpp:136-161) */
int edge_dg_eval_basis(int order, int64_t n_pts, const double *pts, double *val);

/* Collapsed Gauss-Legendre rule with n1d points per direction on the unit
 * tet (src/dg/QuadraturePoints.h:424-530): pts[n1d^3][3], w[n1d^3]. */
int edge_dg_quad_tet(int n1d, double *pts, double *w);

/* The fault phase's tables of order O (host arrays, either may be NULL):
 * *n_sub_faces = n_sf = (2 O - 1)^2; scat[16][N][n_sf], the means over each
 * sub-face of the traces of the N element basis functions (t < 4: on face t
 * in its own coordinates; t = 4 + 4 v + f: on a neighbour's face f seen from
 * the element's face with vertex orientation v, SurfInt.hpp:127-130);
 * lift[16][n_sf][N], the DOF update of a net update of one sub-face (the
 * least-squares face polynomial of the sub-face values, lifted through the
 * inverse mass matrix, times n_sf / (3 n_sc) with n_sc = (2 O - 1)^3). */
int edge_dg_fault_tables(int order, int32_t *n_sub_faces, double *scat, double *lift);

/* ====================================================================== */
/* Structured box meshes (the conventions of tools/edge_v, see mesh.cpp)   */
/* ====================================================================== */
typedef struct {
  int64_t n[3];      /* hexes per axis */
  double origin[3];  /* lower corner */
  double extent[3];  /* edge lengths */
  int32_t bc[3];     /* per axis: EDGE_FA_PERIODIC, _FREE_SURFACE or _OUTFLOW */
  double jitter;     /* interior vertices moved by U[-jitter h, jitter h] per axis */
  uint64_t seed;     /* jitter seed */
  int64_t block;     /* element order: hex blocks of block^3 (1 = lexicographic) */
} edge_box_params;

typedef struct { /* caller-allocated outputs, sizes from edge_mesh_box_sizes */
  double *ve_crds;   /* [nVe][3] */
  int64_t *el_ve;    /* [nEl][4] */
  int64_t *el_fa;    /* [nEl][4] */
  int64_t *fa_el;    /* [nFa][2], -1 = none */
  int64_t *el_fa_el; /* [nEl][4], -1 = none */
  uint16_t *f_id;    /* [nEl][4] fIdElFaEl */
  uint16_t *v_id;    /* [nEl][4] vIdElFaEl */
  int32_t *fa_type;  /* [nFa] */
  double *normals;   /* [nFa][3] outward w.r.t. fa_el[.][0] */
  double *tangents;  /* [nFa][2][3] */
  double *areas;     /* [nFa] */
  double *in_dia;    /* [nEl] */
  double *volume;    /* [nEl] */
} edge_mesh_out;

int edge_mesh_box_sizes(const edge_box_params *p, int64_t *n_ve, int64_t *n_el, int64_t *n_fa);
int edge_mesh_box(const edge_box_params *p, edge_mesh_out *m);

/* Any tet4 mesh (the entity layout EDGE-V writes for the solver,
 * src/mesh/EdgeV.cpp:182-466, built here from vertices and elements): the
 * elements' vertex order is made positive, faces are matched through their
 * vertices, and each face gets the type of a tagged triangle bf_ve with the
 * same vertices (boundary conditions and internal boundaries, EDGE-V's
 * physical groups), else 0 inside and default_bc on the boundary. */
typedef struct {
  int64_t n_ve, n_el;
  const double *ve_crds;  /* [nVe][3] */
  const int64_t *el_ve;   /* [nEl][4] */
  int64_t n_bf;           /* tagged triangles */
  const int64_t *bf_ve;   /* [nBf][3] */
  const int32_t *bf_type; /* [nBf] */
  int32_t default_bc;     /* EDGE_FA_FREE_SURFACE or EDGE_FA_OUTFLOW */
} edge_tets_in;

int edge_mesh_tets_sizes(const edge_tets_in *in, int64_t *n_fa);
int edge_mesh_tets(const edge_tets_in *in, edge_mesh_out *m);

/* ====================================================================== */
/* Element-update solver: replaces                                          */
/* edge::seismic::solvers::AderDg<real, 0, TET4, O, O, R>                   */
/* (src/impl/seismic/solvers/AderDg.hpp:225-312 constructor, :338-534       */
/* local, :553-722 neigh) with its kernels::Kernels members                 */
/* (src/impl/seismic/kernels/Kernels.hpp:71-142).                           */
/* ====================================================================== */
typedef struct edge_el_solver edge_el_solver;

typedef struct {
  int32_t order;     /* O: 2..5 */
  int32_t n_runs;    /* fused runs R (innermost DOF axis): 1, 8 or 16 */
  int32_t precision; /* 64 or 32 (arithmetic and storage type of DOFs) */
  int32_t kernel;    /* 0 = pick automatically, 1 = generic, 2 = MFMA-tiled (FP32, 16 runs),
                        3 = MFMA-tiled with the neighbour step reading tInt tiles,
                        4 = FP64 single run on FP64 MFMA */
} edge_el_config;

typedef struct { /* host arrays in the reference's entity layout */
  int64_t n_el, n_fa;
  const double *ve_crds;    /* [nVe][3] */
  const int64_t *el_ve;     /* [nEl][4] */
  const int64_t *el_fa;     /* [nEl][4] */
  const int64_t *fa_el;     /* [nFa][2] */
  const int64_t *el_fa_el;  /* [nEl][4] */
  const uint16_t *f_id;     /* [nEl][4] */
  const uint16_t *v_id;     /* [nEl][4] */
  const int32_t *fa_type;   /* [nFa] */
  const double *normals;    /* [nFa][3] */
  const double *tangents;   /* [nFa][2][3] */
  const double *areas;      /* [nFa] */
} edge_mesh_in;

/* Builds the solver: copies mesh connectivity to the device and builds the
 * star matrices (AderDgInit.hpp:129-174) and the per-face Godunov flux
 * solvers (AderDgInit.hpp:394-665) in FP64 on the GPU from per-element
 * density / Lame parameters (host arrays [nEl]). */
int edge_el_create(const edge_el_config *cfg, const edge_mesh_in *mesh, const double *rho,
                   const double *lam, const double *mu, edge_el_solver **out);

/* Same, with caller-provided star matrices [nEl][3][81] and flux solvers
 * [2][nEl][4][81] (host, FP64); the mesh only supplies connectivity and face
 * types (normals, tangents, areas and vertices may be NULL). */
int edge_el_create_raw(const edge_el_config *cfg, const edge_mesh_in *mesh, const double *star,
                       const double *fs, edge_el_solver **out);

int edge_el_destroy(edge_el_solver *s);

/* Copies the solver's star matrices [nEl][3][81] / flux solvers
 * [2][nEl][4][81] to host FP64 arrays (either may be NULL). */
int edge_el_get_setup(const edge_el_solver *s, double *star, double *fs);

/* Local step for elements [first, first+n): CK time prediction into tInt,
 * volume integral and local surface integral into dofs.  dofs and tInt are
 * device pointers for the whole mesh in the reference layout and precision.
 * stream: hipStream_t (NULL = default stream).  Mirrors AderDg::local with
 * global time stepping (AderDg.hpp:338-534). */
int edge_el_local(edge_el_solver *s, int64_t first, int64_t n_els, double dt, void *dofs, void *t_int,
                  void *stream);

/* Neighbour step for elements [first, first+n): surface integrals of the
 * face-adjacent elements' tInt (free surface: own tInt and local flux
 * matrix; outflow: skipped).  Mirrors AderDg::neigh (AderDg.hpp:553-722).
 * With the MFMA-tiled kernels (selector 0 or 2, FP32, 16 runs) the local step
 * also writes, per element face, the projection of the element's tInt that
 * its neighbour needs (the reference's neighFluxInt, AderDg.hpp:419-473), and
 * this step reads those instead of t_int: t_int must then be the array the
 * local steps of the neighbouring elements wrote, unchanged since.  The
 * solver records the tInt array of the local steps since the last neighbour
 * step and returns EDGE_ERR_ARG when t_int is another array (a substituted
 * tInt, such as the receive buffers of AderDg.hpp:675-684, needs kernel 3),
 * when those local steps wrote different arrays, or when no local step has
 * written the projections since set-up or the last edge_el_autotune /
 * edge_el_set_variant. */
int edge_el_neigh(edge_el_solver *s, int64_t first, int64_t n_els, const void *t_int, void *dofs,
                  void *stream);

/* n_steps full time steps (local then neigh over all elements) with time
 * step dt.  use_graph != 0 captures one step into a HIP graph on first use
 * and replays it.  FP32 MFMA elastic solvers that exchange face projections
 * (edge_el_kernel_info flag bit 4) run n_steps >= 2 as one launch per time
 * step: the local step of step 0, then n_steps - 1 launches that each carry
 * the neighbour step of step k - 1 and the local step of step k (the
 * projections alternate between two solver-held buffers), then the
 * neighbour step of the last step; with use_graph != 0 the whole sequence
 * is one graph (recaptured when n_steps changes).  An attached fault's phase
 * runs between the launches.  The DOFs are those of n_steps full steps
 * either way; after the call t_int holds the last step's tInt and no
 * separate edge_el_neigh may follow. */
int edge_el_step(edge_el_solver *s, int64_t n_steps, double dt, void *dofs, void *t_int, int use_graph,
                 void *stream);

/* AderDg::neigh of step k followed by AderDg::local of step k + 1 (time
 * step dt) over all elements, as one launch: equals edge_el_neigh(t_int,
 * dofs) then edge_el_local(dt, dofs, t_int), so a caller's time loop
 * (steps.inc:24-77) becomes local, then neigh_local per further step (with
 * point sources or a fault phase between the calls, where they run between
 * local and neigh), then a last neigh.  For FP32 MFMA elastic solvers that
 * exchange face projections (edge_el_kernel_info flag bit 4;
 * EDGE_ERR_UNSUPPORTED otherwise).  t_int must be the array the pending
 * local step wrote, as edge_el_neigh checks; EDGE_ERR_ARG when no local step
 * is pending.  No counterpart in the reference (its neigh and local are
 * separate element loops, AderDg.hpp:338-534, 553-722). */
int edge_el_neigh_local(edge_el_solver *s, double dt, void *dofs, void *t_int, void *stream);

/* The DOFs each fused launch's local step starts from (after the
 * neighbour step) exist only inside that launch; edge_el_neigh_local and the
 * fused edge_el_step also write them to taps (a device array in the DOF
 * layout and size, the caller's) for the n elements el[n], and leave its
 * other elements untouched: receivers sampled at the step start
 * (AderDg.hpp:475-496) read them there (edge_recv_eval on taps).  n = 0
 * switches the taps off.  Elastic solvers only (EDGE_ERR_UNSUPPORTED for
 * viscoelastic ones, whose fused launches write no taps).  Drops a captured
 * step graph.  No counterpart in the reference. */
int edge_el_set_taps(edge_el_solver *s, int64_t n, const int64_t *el, void *taps);

/* Number of elements / modes / runs / bytes of one DOF array of a solver. */
int edge_el_info(const edge_el_solver *s, int64_t *n_el, int32_t *n_modes, int32_t *n_runs,
                 int64_t *dof_bytes);

/* ====================================================================== */
/* Viscoelastic (attenuating) solvers: AderDg with TL_N_RMS > 0            */
/* (src/impl/seismic/solvers/AderDg.hpp:225-312, :338-534, :553-722) with  */
/* the set-up of AderDgInit::initSrcA / initStar / initFs                  */
/* (AderDgInit.hpp:91-117, :129-174, :394-665) and the constant-Q fit of   */
/* src/impl/seismic/setups/ViscoElasticity.{h,cpp}.                        */
/* ====================================================================== */
typedef struct {
  int32_t n_rms;   /* relaxation mechanisms, 1..5 */
  double freq_cen; /* central frequency of the constant-Q band */
  double freq_rat; /* ratio of the band's highest to its lowest frequency */
} edge_visco_config;

/* Builds a viscoelastic solver from phase-velocity Lame parameters and
 * quality factors per element (host arrays [nEl]).  The fit replaces lam /
 * mu by the unrelaxed ones for the elastic star matrices and flux solvers.
 * Runs 1 or 8 fused runs.  cfg->kernel: 1 the generic kernels; 2 the MFMA
 * kernels (FP32, 8 runs, orders 3-5, 1-3 mechanisms, rank-3 factored flux
 * solvers, EDGE_ERR_UNSUPPORTED otherwise), whose local step also writes
 * the face projections the neighbour step reads (as with edge_el_neigh,
 * t_int must then be the array the neighbours' local steps wrote); 3 the
 * MFMA kernels with the neighbour step reading t_int; 4 the FP64 MFMA
 * kernels (FP64, 1 run, 1-3 mechanisms); 0 the MFMA kernels where they
 * apply, else the generic ones.  edge_el_kernel_info reports the choice. */
int edge_el_create_visco(const edge_el_config *cfg, const edge_visco_config *vc, const edge_mesh_in *mesh,
                         const double *rho, const double *lam, const double *mu, const double *qp, const double *qs,
                         edge_el_solver **out);

/* Local / neighbour steps with the anelastic DOFs dofs_a [nEl][n_rms][6][mode][run]
 * (device, the solver's precision) beside the elastic ones; as
 * edge_el_local / edge_el_neigh otherwise.  dofs_a may be NULL for an
 * elastic solver. */
int edge_el_local_a(edge_el_solver *s, int64_t first, int64_t n_els, double dt, void *dofs, void *dofs_a,
                    void *t_int, void *stream);
int edge_el_neigh_a(edge_el_solver *s, int64_t first, int64_t n_els, const void *t_int, void *dofs, void *dofs_a,
                    void *stream);

/* edge_el_step with the anelastic DOFs (time loop of the reference's
 * TimeGroupStatic::computeStep over AderDg::local / neigh with TL_N_RMS > 0);
 * use_graph != 0 captures one step into a HIP graph on first use and
 * replays it.  FP32 MFMA viscoelastic solvers of 8 runs that exchange face
 * projections with flag bit 4 run n_steps >= 2 as fused launches, as
 * edge_el_step describes (after the call no separate edge_el_neigh_a may
 * follow).  As edge_el_step for an elastic solver (dofs_a ignored). */
int edge_el_step_a(edge_el_solver *s, int64_t n_steps, double dt, void *dofs, void *dofs_a, void *t_int,
                   int use_graph, void *stream);

/* Copies a viscoelastic solver's set-up to host FP64 arrays (any may be
 * NULL): anelastic star matrices [nEl][3][6][3], source matrices
 * [nEl][n_rms][6][6], anelastic flux solvers [2][nEl][4][6][9], relaxation
 * frequencies [n_rms], unrelaxed Lame parameters [nEl]. */
int edge_el_get_setup_a(const edge_el_solver *s, double *star_a, double *src_a, double *fs_a, double *rfs,
                        double *lam_e, double *mu_e);

/* ====================================================================== */
/* Receivers and point sources.                                           */
/* ====================================================================== */
typedef struct edge_points edge_points;

/* The N basis functions at reference points xi[n][3] of elements el[n]
 * (src/io/Receivers.cpp:137-145); with src_scaled != 0 divided by the
 * diagonal mass and the element's Jacobian determinant, as the point
 * sources' bEvals (src/impl/seismic/setups/PointSources.hpp:389-470).
 * out[n][N]. */
int edge_el_point_basis(const edge_el_solver *s, int64_t n, const int64_t *el, const double *xi, int32_t src_scaled,
                        double *out);

/* Receivers in elements el[n] with basis values basis[n][N] (edge_el_point_basis,
 * src_scaled 0); replaces io::Receivers' per-element evaluation
 * (src/io/Receivers.cpp:250-290) fed by AderDg::local's evalTimePrediction
 * (src/impl/seismic/solvers/AderDg.hpp:475-496). */
int edge_recv_create(const edge_el_solver *s, int64_t n, const int64_t *el, const double *basis, edge_points **out);

/* Evaluates every receiver at n_tau (<= 8) relative times tau[n][n_tau] of
 * the coming step from the DOFs at its start (device, the solver's layout and
 * precision): out[n][n_tau][9][R] (device, the solver's precision), the time
 * prediction of the element at the receiver's point.  Call before the step's
 * local update; the values for tau < 0 are not meaningful. */
int edge_recv_eval(edge_points *p, const edge_el_solver *s, const void *dofs, int32_t n_tau, const double *tau,
                   void *out, void *stream);

/* edge_recv_eval for a viscoelastic solver: the time prediction comes from
 * the CK with the anelastic terms, which needs the anelastic DOFs dofs_a
 * (device, [nEl][n_rms][6][mode][run]) beside the elastic ones
 * (TimePredVanilla.hpp:200-306 with TL_N_RMS > 0, as the receiver branch of
 * AderDg::local uses it, AderDg.hpp:475-496).  1 or 8 runs.  As
 * edge_recv_eval for an elastic solver (dofs_a ignored). */
int edge_recv_eval_a(edge_points *p, const edge_el_solver *s, const void *dofs, const void *dofs_a, int32_t n_tau,
                     const double *tau, void *out, void *stream);

/* Point sources (PointSources::t_ps, PointSources.hpp:82-107, one run per
 * source): element el[n], fused run run[n], series start times t0[n] and
 * sampling steps dts[n], basis evaluations bev[n][N] (edge_el_point_basis,
 * src_scaled 1), quantity scalings sca[n][9], series samples ts[] with
 * offsets ts_ptr[n+1] (ts_ptr[0] = 0). */
int edge_src_create(const edge_el_solver *s, int64_t n, const int64_t *el, const int32_t *run, const double *t0,
                    const double *dts, const double *bev, const double *sca, const int64_t *ts_ptr,
                    const double *ts, edge_points **out);

/* Adds the sources' contributions of [t1, t2] to the DOFs (device):
 * PointSources::apply (PointSources.hpp:887-930) with Series<1>::integrate
 * (src/linalg/Series.hpp:59-133), FP64. */
int edge_src_apply(edge_points *p, const edge_el_solver *s, double t1, double t2, void *dofs, void *stream);

int edge_points_destroy(edge_points *p);

/* ====================================================================== */
/* Dynamic rupture: the internal-boundary net updates of fault faces with  */
/* linear slip-weakening friction.  Replaces                               */
/* InternalBoundarySolvers<TET4,O,3>::init (src/impl/seismic/solvers/      */
/* InternalBoundary.hpp:743-923), the face data of RuptureInit             */
/* (src/impl/seismic/setups/RuptureInit.hpp:170-206) and, per time step,   */
/* InternalBoundary::netUpdates (InternalBoundary.hpp:184-280) with        */
/* FrictionLaws<3,R>::linSlipWeak (FrictionLaws.hpp:309-456).  FP64.       */
/* ====================================================================== */
typedef struct edge_rupture edge_rupture;

typedef struct {
  int32_t n_sub_faces; /* sub-faces per fault face (n_sf) */
  int32_t n_sub_cells; /* sub-cells per element (scales the flux solvers) */
  int32_t n_runs;      /* fused runs R */
} edge_rupture_config;

/* Per fault face (host arrays): outward normal [nF][3] and tangents
 * [nF][2][3] w.r.t. the left element, area [nF], the fault frame [3][3]
 * (normal, strike, dip; unit vectors) that orients every face, density /
 * lambda / mu of the left and right elements [nF][3], their volumes [nF];
 * friction parameters [3][R] (static and dynamic coefficient, 1/Dc) and the
 * initial sub-face state [nF][n_sf][R][14]: sn0, co0, ss0[2], ss0A, muf, st,
 * tr[2], sr[2], dd[2] (t_LinSlipWeakSubFace, FrictionLaws.type:58-148). */
int edge_rupture_create(const edge_rupture_config *cfg, int64_t n_faces, const double *normals,
                        const double *tangents, const double *areas, const double *fault_frame, const double *mat_l,
                        const double *mat_r, const double *vol_l, const double *vol_r, const double *glob,
                        const double *state0, edge_rupture **out);
int edge_rupture_destroy(edge_rupture *r);

/* Net updates of faces [first, first+n) for the sub-cell DOFs of both sides
 * dofs_l / dofs_r [nF][9][n_sf][R] (device): net_l / net_r [nF][9][n_sf][R]
 * (device; the left-going updates negative, as the reference adds both).
 * Advances the friction state (slip, slip rate, traction, friction
 * coefficient) with the time step dt. */
int edge_rupture_net_updates(edge_rupture *r, int64_t first, int64_t n_faces, double dt, const double *dofs_l,
                             const double *dofs_r, double *net_l, double *net_r, void *stream);

/* Copies the face solvers [nF][4][9][9] (T^-1, jump, left and right flux),
 * the sub-face state [nF][n_sf][R][14] and the failure flags [nF][n_sf][R]
 * to the host (any may be NULL). */
int edge_rupture_get(const edge_rupture *r, double *solvers, double *state, uint8_t *per);

/* ====================================================================== */
/* A fault coupled to the element update (dynamic rupture in the time     */
/* step).  The reference has the pieces (InternalBoundarySolvers<TET4,O,3> */
/* ::init, InternalBoundary.hpp:743-923; RuptureInit, RuptureInit.hpp:     */
/* 170-206; netUpdates with linSlipWeak, InternalBoundary.hpp:184-280) but */
/* no call site in its time loop (src/impl/seismic/steps.inc:24-77); this  */
/* defines one.  Faces of type EDGE_FA_RUPTURE carry no Godunov flux in    */
/* the element update (edge_el_create zeroes their solvers and the         */
/* neighbour step skips them); the fault phase, run between the local and */
/* the neighbour step, adds their flux:                                     */
/*  1. per fault face and sub-face (n_sf = (2 O - 1)^2 sub-triangles of the */
/*     face, in the coordinates of the face's first element L, rows along  */
/*     its first face edge, upward then downward triangles), the means of  */
/*     both elements' time-averaged traces tInt / dt;                       */
/*  2. the net updates (edge_rupture_net_updates with n_sub_faces = n_sf,  */
/*     n_sub_cells = (2 O - 1)^3), advancing the friction state;          */
/*  3. the face polynomial of degree O - 1 whose sub-face means are closest */
/*     to the net updates (least squares), lifted into both elements as a  */
/*     surface flux.  A fault that does not slip gives the welded Godunov  */
/*     flux exactly.                                                       */
/* ====================================================================== */
typedef struct edge_fault edge_fault;

/* Builds the fault of an elastic solver's mesh (the same mesh and
 * materials as edge_el_create; faces of type EDGE_FA_RUPTURE, each between
 * two elements): fault_frame [3][3] (normal, strike, dip; unit vectors),
 * friction parameters glob [3][R] (static and dynamic coefficient, 1/Dc)
 * and the initial sub-face state state0 [nF][n_sf][R][14] as
 * edge_rupture_create, fault faces in mesh-face order.  The precision of
 * the DOFs is the solver's; the friction and net updates run in FP64. */
int edge_fault_create(const edge_el_solver *s, const edge_mesh_in *mesh, const double *rho, const double *lam,
                      const double *mu, const double *fault_frame, const double *glob, const double *state0,
                      edge_fault **out);
int edge_fault_destroy(edge_fault *f);

/* The fault phase for the solver's time step dt: reads tInt (after the
 * local step) and adds the fault faces' fluxes to dofs (device arrays of
 * the solver's layout and precision). */
int edge_fault_apply(edge_fault *f, const edge_el_solver *s, double dt, const void *t_int, void *dofs,
                     void *stream);

/* Number of fault faces and sub-faces; per fault face (any may be NULL) the
 * mesh face id faces[nF], its elements L, R elements[nF][2] and L's local
 * face local_faces[nF] (the sub-faces lie in that face's coordinates). */
int edge_fault_info(const edge_fault *f, int64_t *n_faces, int32_t *n_sub_faces, int64_t *faces, int64_t *elements,
                    int32_t *local_faces);

/* The friction state [nF][n_sf][R][14] and slipping flags [nF][n_sf][R] of
 * the last step (either may be NULL). */
int edge_fault_get(const edge_fault *f, double *state, uint8_t *slipping);

/* Attaches a fault built for this solver (NULL detaches it): edge_el_step
 * then runs local, the fault phase, neighbour.  Elastic solvers with global
 * time stepping.  Drops a captured step graph. */
int edge_el_set_fault(edge_el_solver *s, edge_fault *f);

/* ---------------------------------------------------------------------- */
/* Rate-2 local time stepping (AderDg.hpp:384-417 and :620-674): elements   */
/* are sorted by time group and advanced group-range by group-range.        */
/* ---------------------------------------------------------------------- */
/* The C_LTS_EL / C_LTS_AD bits of the elements' spType: el_lts[nEl] bit 0
 * "an adjacent element has a larger time step" (the element keeps tDofs[1]),
 * bit 1 "an adjacent element has a smaller time step" (tDofs[2]);
 * fa_lts[nEl][4]: 0 equal steps, 1 the element's step is the greater, 2 the
 * smaller (interior faces only).  NULL, NULL switches local time stepping
 * off.  Elastic and viscoelastic solvers (the latter step with
 * edge_el_local_lts_a / edge_el_neigh_lts_a). */
int edge_el_set_lts(edge_el_solver *s, const uint8_t *el_lts, const uint8_t *fa_lts);

/* AderDg::local with local time stepping over elements [first, first+n):
 * t_int0 = tDofs[0] (as edge_el_local), t_int1 = tDofs[1] (reset on the
 * first step of a pair, first_ts != 0, then summed), t_int2 = tDofs[2]
 * (integral over [0, dt/2]); device arrays [nEl][9][N][R]. */
int edge_el_local_lts(edge_el_solver *s, int64_t first, int64_t n_els, double dt, int32_t first_ts, void *dofs,
                      void *t_int0, void *t_int1, void *t_int2, void *stream);

/* AderDg::neigh with local time stepping: faces between equal steps take the
 * neighbour's tDofs[0], faces to a smaller step its tDofs[1], faces to a
 * larger step its tDofs[2] (first_ts != 0) or tDofs[0] - tDofs[2]. */
int edge_el_neigh_lts(edge_el_solver *s, int64_t first, int64_t n_els, int32_t first_ts, const void *t_int0,
                      const void *t_int1, const void *t_int2, void *dofs, void *stream);

/* The same two steps with the anelastic DOFs dofs_a [nEl][n_rms][6][N][R]
 * of a viscoelastic solver (AderDg::local / neigh with TL_N_RMS > 0 and
 * local time stepping, AderDg.hpp:384-417, :620-720): tDofs[0..2] hold the
 * elastic time-integrated DOFs, tDofs[2] the integral over [0, dt/2] of the
 * CK with the anelastic terms; an interface face adds the anelastic flux
 * update to every mechanism.  1 or 8 runs.  As the elastic calls for an
 * elastic solver (dofs_a ignored). */
int edge_el_local_lts_a(edge_el_solver *s, int64_t first, int64_t n_els, double dt, int32_t first_ts, void *dofs,
                        void *dofs_a, void *t_int0, void *t_int1, void *t_int2, void *stream);
int edge_el_neigh_lts_a(edge_el_solver *s, int64_t first, int64_t n_els, int32_t first_ts, const void *t_int0,
                        const void *t_int1, const void *t_int2, void *dofs, void *dofs_a, void *stream);

/* Which kernels a solver runs: *kernel = 1 (generic), 2 (MFMA-tiled, runs
 * on the MFMA N dimension), 4 (FP64 single run on FP64 MFMA, quantities on
 * the M dimension) or 5 (single run, 16 elements on the N dimension);
 * *flags bit 0: sparse (elastic-pattern) star matrices, bit 1: rank-3
 * factored flux solvers, bit 2: the neighbour step reads face projections,
 * bit 3: viscoelastic (anelastic DOFs), bit 4: multi-step calls fuse the
 * neighbour step of step k - 1 with the local step of step k (edge_el_step),
 * bit 5: the higher-occupancy variant of the kernels (FP32 viscoelastic MFMA
 * local kernel: tIntA in LDS, five waves per SIMD).
 * Diagnostic with no counterpart in the reference (for tests and tools). */
int edge_el_kernel_info(const edge_el_solver *s, int32_t *kernel, int32_t *flags);

/* Autotuner: times every kernel variant and schedule the solver can switch
 * to (tile shapes: runs on the MFMA N dimension, quantities on the M
 * dimension or 16 elements on the N dimension; the FP32 MFMA kernels with
 * and without the face-projection exchange, for elastic solvers with it
 * also the fused multi-step form, for viscoelastic ones each in two
 * occupancy variants; each with the XCD-aware element schedule or the
 * linear one; the generic kernels; all parity-tested against the oracle) over
 * n_reps full steps of time step dt on scratch DOFs on its own mesh (a fused
 * candidate: one fused sequence of n_reps steps), and
 * keeps the fastest for the solver's later calls.  In: *n_cands = capacity
 * of the arrays (may be NULL); out: *n_cands = number of candidates,
 * kernels[i] (as edge_el_kernel_info), flags[i] (bit 0: projections
 * exchanged, bit 1: XCD-aware schedule, bit 2: fused multi-step form, bit 3:
 * the higher-occupancy variant) and
 * ms[i] (per step).  Drops a
 * captured step graph.  Its steps overwrite the face projections: a
 * neighbour step pending from before the call returns EDGE_ERR_ARG.  No counterpart in the reference (its back-end is a
 * compile-time choice, Kernels.hpp:30-44). */
int edge_el_autotune(edge_el_solver *s, double dt, int32_t n_reps, void *stream, int32_t *n_cands,
                     int32_t *kernels, int32_t *flags, double *ms);

/* Switches the solver to one of the candidates edge_el_autotune times
 * (kernel and flags as it reports them; flags are ignored for the generic
 * kernel), without timing: EDGE_ERR_ARG if the solver has no such variant.
 * Drops a captured step graph.  For parity tests of every candidate; no
 * counterpart in the reference. */
int edge_el_set_variant(edge_el_solver *s, int32_t kernel, int32_t flags);

/* FP64 L2 norms per quantity and run, exact through the diagonal mass
 * matrix: for device DOF arrays a (element precision prec_a: 32 or 64) and
 * b (prec_b; NULL for none), sets diff_sq[9][R] = ||a - b||^2 and
 * ref_sq[9][R] = ||b||^2 (or ||a||^2 without b), each sum_el |J| sum_n M_n x_n^2.
 * The relative L2 error of a against b is sqrt(diff_sq / ref_sq).  Needs a
 * solver built by edge_el_create (element Jacobians).  The error-norm
 * reduction of src/setups/InitialDofs.hpp:277-370 in modal form. */
int edge_el_l2_norms(edge_el_solver *s, const void *a, int prec_a, const void *b, int prec_b, double *diff_sq,
                     double *ref_sq, void *stream);

/* edge_el_l2_norms without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops).  No counterpart in the reference.  Same arguments and results,
 * bit for bit: the per-block partial sums live in a static device array of
 * the library (1024 blocks x 2 x 9 x 16 doubles, 2.25 MiB, held on every
 * GPU the library runs on until the process ends) and are summed on the
 * device in the same order.  Solvers of at most 16 runs (EDGE_ERR_ARG
 * beyond).  Calls on one GPU share the array: calls that overlap on one GPU
 * (other streams or threads) race.  Synchronises the stream before it
 * returns (the results are on the host). */
int edge_el_l2_norms_noalloc(edge_el_solver *s, const void *a, int prec_a, const void *b, int prec_b,
                             double *diff_sq, double *ref_sq, void *stream);


/* Elastic energy per run of a solver's device DOFs (its precision), FP64 on
 * the GPU: kinetic 1/2 rho |v|^2 and strain 1/2 sigma : C^-1 : sigma with the
 * element's rho, lambda, mu (the unrelaxed ones for viscoelastic solvers),
 * integrated exactly through |J| and the diagonal mass matrix; kinetic[R],
 * strain[R] on the host (either may be NULL).  Solvers built from materials
 * only (EDGE_ERR_UNSUPPORTED for edge_el_create_raw).  No counterpart in the
 * reference's kernels (its energy and error output lives in the setups,
 * InitialDofs.hpp:277-370). */
int edge_el_energy(edge_el_solver *s, const void *dofs, double *kinetic, double *strain, void *stream);

/* L1, squared L2 and Linf errors per quantity and run (each [9][R], host),
 * FP64 on the GPU, of a solver's device DOFs against the sum of n_waves
 * plane waves q(x, t) = sum_w r_w sin(k_w . x - omega_w t + phase_w) per run
 * (waves [R][n_waves][14]: r[9], k[3], omega, phase; host), evaluated at
 * time t on the collapsed Gauss rule with n1d points per direction (0:
 * order + 1) of every element.  Replaces InitialDofs::err
 * (src/setups/InitialDofs.hpp:277-370), whose exact solutions are exprtk
 * expressions, for plane-wave solutions.  Solvers built from geometry only. */
int edge_el_error_norms(edge_el_solver *s, const void *dofs, int32_t n1d, int32_t n_waves, const double *waves,
                        double t, double *l1, double *l2sq, double *linf, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* EDGE_AMD_H */
