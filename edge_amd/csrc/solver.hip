// The element-update solver behind the C ABI: device copies of the
// reference-element matrices and of the neighbour relations, the set-up, and
// the local / neighbour steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <utility>
#include <vector>

#include "kernels.hpp"
#include "ec64_pack.hpp"
#include "kernels_mfma.hpp"
#include "mfma_pack.hpp"
#include "setup.hpp"
#include "visco.hpp"

struct edge_el_solver {
  edge_el_config cfg;
  int64_t n_el = 0;
  int n_modes = 0, n_face_modes = 0;
  int kernel = 1;  // 1 generic, 2 MFMA-tiled
  int tune_kernel = 0;  // the kernel set-up chose (0: not recorded yet): the autotuner's candidates derive from it
  size_t real_bytes = 8;
  // device allocations
  void* mats = nullptr;  // stiffTn | stiffV | fluxLN | fluxT  (precision T); FP64: then edge::PadLayout
  float* pk = nullptr;   // MFMA-tiled path: A operands packed in lane order
  void* star = nullptr;  // [nEl][3][81]
  void* fs = nullptr;    // [2][nEl][4][81]
  int32_t* ne = nullptr;
  uint8_t* fmat = nullptr;
  double* jdet = nullptr;  // [nEl] Jacobian determinants (geometry-built solvers only)
  bool sparse_star = false;  // all star matrices have the elastic 24-entry pattern
  double* mass = nullptr;  // [N] diagonal reference mass matrix
  double* mat = nullptr;   // [3][nEl] rho, lambda, mu (FP64; solvers built from materials)
  double* geo = nullptr;   // [nEl][12] v0, v1 - v0, v2 - v0, v3 - v0 (FP64; solvers built from geometry)
  // MFMA-tiled path: rank-3 flux-solver factors and face projections
  void* fsf = nullptr;       // [2][nEl][4][54]
  bool fs_rank3 = false;
  bool want_proj = false;    // kernel selector 2: exchange face projections when possible
  bool use_proj = true;      // the projections are exchanged when they exist (the autotuner may turn this off)
  bool xcd_map = true;       // MFMA kernels: contiguous element ranges per XCD (the autotuner may turn this off)
  bool hi_occ = false;       // kernels with an occupancy variant: the higher-occupancy one (an autotuner choice)
  bool proj_tables = false;  // build the projection slot tables whatever the kernel (MFMA viscoelastic)
  void* proj = nullptr;      // [nEl][4][768]
  void* proj_alt = nullptr;  // the second projection buffer of the fused step (allocated on first use)
  bool use_fused = true;     // multi-step calls fuse neighbour step k - 1 with local step k (the autotuner may turn this off)
  uint8_t* tap_on = nullptr; // [nEl] elements whose step-start DOFs the fused launches also write to tap
  void* tap = nullptr;       // caller's array in the DOF layout (edge_el_set_taps), null: none
  int32_t* snd_v = nullptr;  // [nEl][4]
  uint8_t* snd_fm = nullptr; // [nEl][4]
  int32_t* rcv = nullptr;    // [nEl][4]
  bool proj_ok = false;      // every projection slot feeds at most one element face
  // The face projections stand in for the tInt array of the local step that
  // wrote them: the neighbour step checks that it is handed the same array.
  const void* proj_tint = nullptr;  // tInt of the local steps since the last neighbour step (null: none valid)
  bool proj_mixed = false;          // those local steps wrote different tInt arrays
  bool proj_fresh = false;          // a local step ran since the last neighbour step
  // element-batched single-run kernels (kernel 5): packed operands,
  // face-orientation maps and their face projections [nEl][4][3][16]
  // (solver precision)
  bool want_ec = false;  // set-up builds their data when the solver qualifies
  bool ec_ok = false;
  void* ec_pk = nullptr;
  void* ec_rot = nullptr;
  void* ec_proj = nullptr;
  // viscoelastic solvers
  int n_rms = 0;
  void* rfs = nullptr;       // [n_rms]
  void* star_a = nullptr;    // [nEl][54]
  void* src_a = nullptr;     // [nEl][n_rms][36]
  void* fs_a = nullptr;      // [2][nEl][4][54]
  void* fsfa = nullptr;      // [2][nEl][4][18] anelastic factors (MFMA viscoelastic kernels)
  std::vector<double> lam_e, mu_e, h_rfs;
  // element faces on a fault (face type EDGE_FA_RUPTURE), el * 4 + fa
  std::vector<int64_t> h_fault_ef;
  edge_fault* fault = nullptr;  // attached fault: the step runs its phase between local and neighbour
  // rate-2 local time stepping
  uint8_t* fa_lts = nullptr;   // [nEl][4]: 0 equal steps, 1 this element's step greater, 2 smaller
  int32_t* ne_lts = nullptr;   // ne with the interface faces skipped
  int32_t* rcv_lts = nullptr;  // rcv with the interface faces skipped
  int64_t* lts_lt = nullptr;   // elements with a larger-step neighbour (tDofs[1]), ascending
  int64_t* lts_gt = nullptr;   // elements with a smaller-step neighbour (tDofs[2]), ascending
  int64_t* lts_if = nullptr;   // elements with interface faces, ascending
  std::vector<int64_t> h_lts_lt, h_lts_gt, h_lts_if;
  // graph path
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  void* graph_dofs = nullptr;
  void* graph_dofs_a = nullptr;
  void* graph_tint = nullptr;
  double graph_dt = 0;
  hipStream_t graph_stream = nullptr;
  int64_t graph_steps = 0;  // > 0: the graph holds a whole fused sequence of this many steps
};

namespace {

template <typename T>
edge::SolverDev<T> dev_view(const edge_el_solver* s) {
  edge::SolverDev<T> d;
  d.order = s->cfg.order;
  d.n_runs = s->cfg.n_runs;
  d.n_el = s->n_el;
  const size_t N = s->n_modes, NF = s->n_face_modes;
  const T* m = static_cast<const T*>(s->mats);
  d.dg.stiffTn = m;
  d.dg.stiffV = m + 3 * N * N;
  d.dg.fluxLN = m + 6 * N * N;
  d.dg.fluxT = m + 6 * N * N + 16 * N * NF;
  if constexpr (sizeof(T) == 4) d.dg.pk = s->pk;
  if constexpr (sizeof(T) == 8) d.dg.pad = m + 6 * N * N + 16 * N * NF + 4 * NF * N;
  d.nb.ne = s->ne;
  d.nb.fmat = s->fmat;
  d.star = static_cast<const T*>(s->star);
  d.fs = static_cast<const T*>(s->fs);
  d.sparse_star = s->sparse_star;
  d.xcd_map = s->xcd_map;
  d.hi_occ = s->hi_occ;
  d.fsf = static_cast<const T*>(s->fsf);
  d.fs_rank3 = s->fs_rank3 && s->fsf;
  d.proj = d.fs_rank3 && s->use_proj ? static_cast<T*>(s->proj) : nullptr;
  d.snd_v = s->snd_v;
  d.snd_fm = s->snd_fm;
  d.rcv = s->rcv;
  d.tap_on = s->tap ? s->tap_on : nullptr;
  d.tap = static_cast<T*>(s->tap);
  d.n_rms = s->n_rms;
  d.rfs = static_cast<const T*>(s->rfs);
  d.star_a = static_cast<const T*>(s->star_a);
  d.src_a = static_cast<const T*>(s->src_a);
  d.fs_a = static_cast<const T*>(s->fs_a);
  d.fsfa = static_cast<const T*>(s->fsfa);
  return d;
}

template <typename T>
int upload_mats(edge_el_solver* s, const edge::DgTables& t) {
  const size_t N = t.n, NF = t.nf;
  const size_t base = 6 * N * N + 16 * N * NF + 4 * NF * N;
  // FP64: the zero-padded copies of the FP64 single-run kernels follow
  const size_t total = base + (sizeof(T) == 8 ? edge::PadLayout::total((int)N, (int)NF) : 0);
  std::vector<T> h(total, T(0));
  size_t o = 0;
  for (size_t i = 0; i < 3 * N * N; i++) h[o++] = (T)(-t.stiffT[i]);
  for (size_t i = 0; i < 3 * N * N; i++) h[o++] = (T)t.stiffV[i];
  for (size_t i = 0; i < 4 * N * NF; i++) h[o++] = (T)t.fluxL[i];
  for (size_t i = 0; i < 12 * N * NF; i++) h[o++] = (T)t.fluxN[i];
  for (size_t i = 0; i < 4 * NF * N; i++) h[o++] = (T)t.fluxT[i];
  if (sizeof(T) == 8) {
    using P = edge::PadLayout;
    const int n = (int)N, nf = (int)NF;
    // mats [nm][rows][cols] (dense, starting at src) into [nm][P4(rows)][P16(cols)] at dst
    auto pad = [&](size_t src, size_t dst, int nm, int rows, int cols) {
      for (int m = 0; m < nm; m++)
        for (int r = 0; r < rows; r++)
          for (int c = 0; c < cols; c++)
            h[base + dst + ((size_t)m * P::p4(rows) + r) * P::p16(cols) + c] =
                h[src + ((size_t)m * rows + r) * cols + c];
    };
    pad(0, 0, 3, n, n);
    pad(3 * N * N, P::stiff_v(n), 3, n, n);
    pad(6 * N * N, P::flux_ln(n), 16, n, nf);
    pad(6 * N * N + 16 * N * NF, P::flux_t(n, nf), 4, nf, n);
  }
  EDGE_HIP_TRY(hipMalloc(&s->mats, sizeof(T) * total));
  EDGE_HIP_TRY(hipMemcpy(s->mats, h.data(), sizeof(T) * total, hipMemcpyHostToDevice));
  if (s->kernel == 2) {
    const std::vector<float> p = edge::mfma_pack_operands(t);
    EDGE_HIP_TRY(hipMalloc(&s->pk, sizeof(float) * p.size()));
    EDGE_HIP_TRY(hipMemcpy(s->pk, p.data(), sizeof(float) * p.size(), hipMemcpyHostToDevice));
  }
  return EDGE_OK;
}

int check_cfg(const edge_el_config* c) {
  if (!c) return edge::fail(EDGE_ERR_ARG, "config missing");
  if (c->order < edge::kMinOrder || c->order > edge::kMaxOrder)
    return edge::fail(EDGE_ERR_UNSUPPORTED, "order must be in 2..5");
  if (c->n_runs != 1 && c->n_runs != 8 && c->n_runs != 16)
    return edge::fail(EDGE_ERR_UNSUPPORTED, "n_runs must be 1, 8 or 16");
  if (c->precision != 64 && c->precision != 32) return edge::fail(EDGE_ERR_UNSUPPORTED, "precision must be 32 or 64");
  if (c->kernel < 0 || c->kernel > 5) return edge::fail(EDGE_ERR_ARG, "kernel selector");
  return EDGE_OK;
}

int check_mesh(const edge_mesh_in* m) {
  if (!m || m->n_el <= 0 || m->n_fa <= 0 || !m->el_fa || !m->fa_el || !m->el_fa_el || !m->f_id || !m->v_id ||
      !m->fa_type)
    return edge::fail(EDGE_ERR_ARG, "mesh connectivity missing");
  if (m->n_el >= (int64_t)1 << 31) return edge::fail(EDGE_ERR_UNSUPPORTED, "more than 2^31 elements");
  return EDGE_OK;
}

// neighbour element and flux-matrix id per element face
// (src/impl/seismic/solvers/AderDg.hpp:586-600, :686-698; kernels/SurfInt.hpp:127-130)
int build_neigh(edge_el_solver* s, const edge_mesh_in* m) {
  const int64_t ne = m->n_el;
  std::vector<int32_t> nb(4 * ne);
  std::vector<uint8_t> fm(4 * ne);
  for (int64_t el = 0; el < ne; el++)
    for (int fa = 0; fa < 4; fa++) {
      const int64_t f = m->el_fa[el * 4 + fa];
      if (f < 0 || f >= m->n_fa) return edge::fail(EDGE_ERR_MESH, "element face id out of range");
      const int32_t ty = m->fa_type[f];
      if ((ty & EDGE_FA_RUPTURE) == EDGE_FA_RUPTURE) {
        // a fault face: its flux comes from the fault phase (sub-face net
        // updates), neither from the Godunov solvers nor the neighbour step
        nb[el * 4 + fa] = -1;
        fm[el * 4 + fa] = 0;
        s->h_fault_ef.push_back(el * 4 + fa);
      } else if ((ty & EDGE_FA_OUTFLOW) == EDGE_FA_OUTFLOW) {
        nb[el * 4 + fa] = -1;
        fm[el * 4 + fa] = 0;
      } else if ((ty & EDGE_FA_FREE_SURFACE) == EDGE_FA_FREE_SURFACE) {
        nb[el * 4 + fa] = (int32_t)el;
        fm[el * 4 + fa] = (uint8_t)fa;
      } else {
        const int64_t n = m->el_fa_el[el * 4 + fa];
        const unsigned v = m->v_id[el * 4 + fa], fi = m->f_id[el * 4 + fa];
        if (n < 0 || n >= ne || v > 2 || fi > 3)
          return edge::fail(EDGE_ERR_MESH, "interior face without a valid neighbour");
        nb[el * 4 + fa] = (int32_t)n;
        fm[el * 4 + fa] = (uint8_t)(4 + v * 4 + fi);
      }
    }
  EDGE_HIP_TRY(hipMalloc(&s->ne, sizeof(int32_t) * 4 * ne));
  EDGE_HIP_TRY(hipMalloc(&s->fmat, 4 * ne));
  EDGE_HIP_TRY(hipMemcpy(s->ne, nb.data(), sizeof(int32_t) * 4 * ne, hipMemcpyHostToDevice));
  EDGE_HIP_TRY(hipMemcpy(s->fmat, fm.data(), 4 * ne, hipMemcpyHostToDevice));
  if (s->kernel != 2 && !s->proj_tables && !s->want_ec) return EDGE_OK;
  // Face projections: element el's face fa receives the projection that its
  // neighbour n computes through the shared face (n's face fId, or fa itself
  // on a free surface) with el's flux matrix and neighbour-side solver.
  std::vector<int32_t> snd_v(4 * ne, -1), rcv(4 * ne, -1);
  std::vector<uint8_t> snd_fm(4 * ne, 0);
  s->proj_ok = true;
  for (int64_t el = 0; el < ne && s->proj_ok; el++)
    for (int fa = 0; fa < 4; fa++) {
      const int32_t n = nb[el * 4 + fa];
      if (n < 0) continue;
      const int f = fm[el * 4 + fa] < 4 ? fa : (int)m->f_id[el * 4 + fa];
      const int64_t slot = (int64_t)n * 4 + f;
      if (snd_v[slot] >= 0) {
        s->proj_ok = false;
        break;
      }
      snd_v[slot] = (int32_t)(el * 4 + fa);
      snd_fm[slot] = fm[el * 4 + fa];
      rcv[el * 4 + fa] = (int32_t)slot;
    }
  if (!s->proj_ok) return EDGE_OK;
  EDGE_HIP_TRY(hipMalloc(&s->snd_v, sizeof(int32_t) * 4 * ne));
  EDGE_HIP_TRY(hipMalloc(&s->snd_fm, 4 * ne));
  EDGE_HIP_TRY(hipMalloc(&s->rcv, sizeof(int32_t) * 4 * ne));
  EDGE_HIP_TRY(hipMemcpy(s->snd_v, snd_v.data(), sizeof(int32_t) * 4 * ne, hipMemcpyHostToDevice));
  EDGE_HIP_TRY(hipMemcpy(s->snd_fm, snd_fm.data(), 4 * ne, hipMemcpyHostToDevice));
  EDGE_HIP_TRY(hipMemcpy(s->rcv, rcv.data(), sizeof(int32_t) * 4 * ne, hipMemcpyHostToDevice));
  return EDGE_OK;
}

// after the flux solvers (and their factors) exist: projection buffer
// Zeroes both sides' Godunov solvers (and their factors) of the element
// faces listed in ef ([n], el * 4 + fa): arr [2][nEl][4][per].
template <typename T>
__global__ void k_zero_faces(const int64_t* ef, int64_t n, int64_t n_el, int per, T* arr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * 2 * per) return;
  const int64_t f = i / (2 * per), side = (i / per) % 2, k = i % per;
  arr[(side * n_el * 4 + ef[f]) * per + k] = T(0);
}

template <typename T>
int zero_fault_faces_t(edge_el_solver* s) {
  const int64_t n = (int64_t)s->h_fault_ef.size();
  int64_t* d = nullptr;
  EDGE_HIP_TRY(hipMalloc(&d, sizeof(int64_t) * n));
  hipError_t e = hipMemcpy(d, s->h_fault_ef.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice);
  const std::pair<void*, int> arrs[] = {{s->fs, 81}, {s->fsf, 54}, {s->fs_a, 54}, {s->fsfa, 18}};
  for (const auto& a : arrs) {
    if (!a.first || e != hipSuccess) continue;
    const int64_t tot = n * 2 * a.second;
    hipLaunchKernelGGL(k_zero_faces<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, 0, d, n, s->n_el,
                       a.second, (T*)a.first);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  (void)hipFree(d);
  if (e != hipSuccess) return edge::fail(EDGE_ERR_HIP, std::string("fault faces: ") + hipGetErrorString(e));
  return EDGE_OK;
}

// Fault faces carry no Godunov flux: their local and neighbour solvers are
// zero (the fault phase adds their flux instead).
int zero_fault_faces(edge_el_solver* s) {
  if (s->h_fault_ef.empty()) return EDGE_OK;
  return s->cfg.precision == 64 ? zero_fault_faces_t<double>(s) : zero_fault_faces_t<float>(s);
}

int finish_mfma(edge_el_solver* s) {
  if (s->kernel == 2 && s->want_proj && s->fs_rank3 && s->proj_ok)
    EDGE_HIP_TRY(hipMalloc(&s->proj, sizeof(float) * 4 * 768 * s->n_el));
  return EDGE_OK;
}

// Uploads FP64 host values as the solver's precision T.
template <typename T>
int upload_as(const std::vector<double>& h, void** dst) {
  const std::vector<T> v(h.begin(), h.end());
  EDGE_HIP_TRY(hipMalloc(dst, sizeof(T) * v.size()));
  EDGE_HIP_TRY(hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return EDGE_OK;
}

// The element-batched kernels' data, once the flux solvers and their factors
// exist: they need the elastic star pattern, rank-3 solvers, one projection
// slot per element face and face-orientation maps.  Solvers that may run
// them and were not given another kernel take them by default.
int finish_ec(edge_el_solver* s) {
  if (!s->want_ec) return EDGE_OK;
  edge::DgTables t;
  std::vector<double> rot;
  const bool ok = s->sparse_star && s->fs_rank3 && s->fsf && s->proj_ok && edge::dg_tables(s->cfg.order, &t) &&
                  edge::ec_rotations(t, &rot);
  if (!ok) {
    if (s->kernel == 5)
      return edge::fail(EDGE_ERR_UNSUPPORTED, "element-batched kernels: they need the elastic star pattern and "
                                              "rank-3 flux solvers");
    return EDGE_OK;
  }
  const bool f32 = s->cfg.precision == 32;
  const std::vector<double> pk = edge::ec_pack_operands(t, f32);
  int rc = f32 ? upload_as<float>(pk, &s->ec_pk) : upload_as<double>(pk, &s->ec_pk);
  if (rc == EDGE_OK) rc = f32 ? upload_as<float>(rot, &s->ec_rot) : upload_as<double>(rot, &s->ec_rot);
  if (rc) return rc;
  EDGE_HIP_TRY(hipMalloc(&s->ec_proj, s->real_bytes * 4 * 48 * s->n_el));
  s->ec_ok = true;
  if (s->cfg.kernel == 0) s->kernel = 5;
  return EDGE_OK;
}

template <typename T>
edge::EcDev<T> ec_view(const edge_el_solver* s) {
  edge::EcDev<T> x;
  x.pk = static_cast<const T*>(s->ec_pk);
  x.rot = static_cast<const T*>(s->ec_rot);
  x.proj = static_cast<T*>(s->ec_proj);
  return x;
}

int pick_kernel(edge_el_solver* s) {
  const bool mfma_ok = edge::mfma_supported(s->cfg.order, s->cfg.n_runs, s->cfg.precision);
  const bool gen_ok = edge::generic_supported(s->cfg.order, s->cfg.n_runs, (int)s->real_bytes);
  const bool p64_ok = edge::plain64_supported(s->cfg.order, s->cfg.n_runs, s->cfg.precision / 8);
  const bool ec_ok = edge::ec_supported(s->cfg.order, s->cfg.n_runs);
  // the element-batched kernels' data is built whenever they may run
  // (requested, the default, or an autotuner candidate)
  s->want_ec = ec_ok && (s->cfg.kernel == 0 || s->cfg.kernel == 4 || s->cfg.kernel == 5);
  if (s->cfg.kernel == 5) {
    if (!ec_ok) return edge::fail(EDGE_ERR_UNSUPPORTED, "element-batched kernels: one run, orders 2-5");
    s->kernel = 5;
  } else if (s->cfg.kernel == 2 || s->cfg.kernel == 3) {
    if (!mfma_ok) return edge::fail(EDGE_ERR_UNSUPPORTED, "no MFMA-tiled kernel for this configuration");
    s->kernel = 2;
    s->want_proj = s->cfg.kernel == 2;
  } else if (s->cfg.kernel == 1) {
    if (!gen_ok) return edge::fail(EDGE_ERR_UNSUPPORTED, "generic kernel does not fit this configuration");
    s->kernel = 1;
  } else if (s->cfg.kernel == 4) {
    if (!p64_ok) return edge::fail(EDGE_ERR_UNSUPPORTED, "the FP64 MFMA kernel needs FP64 and one run");
    s->kernel = 4;
  } else {
    s->kernel = mfma_ok ? 2 : (p64_ok ? 4 : 1);
    s->want_proj = mfma_ok;
    if (!mfma_ok && !p64_ok && !gen_ok) return edge::fail(EDGE_ERR_UNSUPPORTED, "no kernel for this configuration");
  }
  return EDGE_OK;
}

int create_common(const edge_el_config* cfg, const edge_mesh_in* mesh, edge_el_solver** out,
                  edge_el_solver** made, bool proj_tables = false) {
  int rc = check_cfg(cfg);
  if (rc) return rc;
  if ((rc = check_mesh(mesh))) return rc;
  if (!out) return edge::fail(EDGE_ERR_ARG, "output handle missing");
  edge::DgTables t;
  if (!edge::dg_tables(cfg->order, &t)) return edge::fail(EDGE_ERR_UNSUPPORTED, "order not generated");
  edge_el_solver* s = new (std::nothrow) edge_el_solver();
  if (!s) return edge::fail(EDGE_ERR_NOMEM, "solver");
  *made = s;
  s->cfg = *cfg;
  s->n_el = mesh->n_el;
  s->n_modes = t.n;
  s->n_face_modes = t.nf;
  s->real_bytes = cfg->precision == 64 ? 8 : 4;
  s->proj_tables = proj_tables;
  if ((rc = pick_kernel(s))) return rc;
  rc = cfg->precision == 64 ? upload_mats<double>(s, t) : upload_mats<float>(s, t);
  if (rc) return rc;
  if ((rc = build_neigh(s, mesh))) return rc;
  EDGE_HIP_TRY(hipMalloc(&s->star, s->real_bytes * 243 * s->n_el));
  EDGE_HIP_TRY(hipMalloc(&s->fs, s->real_bytes * 2 * 324 * s->n_el));
  if (s->kernel == 2 || s->want_ec) EDGE_HIP_TRY(hipMalloc(&s->fsf, s->real_bytes * 2 * 4 * 54 * s->n_el));
  EDGE_HIP_TRY(hipMalloc(&s->mass, sizeof(double) * t.n));
  EDGE_HIP_TRY(hipMemcpy(s->mass, t.mass, sizeof(double) * t.n, hipMemcpyHostToDevice));
  return EDGE_OK;
}

// The MFMA kernels exchange face projections between the local and the
// neighbour step: the neighbour step then reads them instead of tInt.
bool proj_active(const edge_el_solver* s) {
  return (s->kernel == 2 && s->fs_rank3 && s->fsf && s->use_proj && s->proj) || (s->kernel == 5 && s->ec_ok);
}

// A local step that writes projections records the tInt array it wrote.
void note_local(edge_el_solver* s, const void* t_int) {
  if (!proj_active(s)) return;
  if (!s->proj_fresh) {
    s->proj_tint = t_int;
    s->proj_mixed = false;
    s->proj_fresh = true;
  } else if (s->proj_tint != t_int) {
    s->proj_mixed = true;
  }
}

// The neighbour step reads the projections in place of t_int: they must come
// from local steps that wrote this very array (a substituted tInt, such as a
// receive buffer, would be silently ignored otherwise).
int check_neigh(edge_el_solver* s, const void* t_int) {
  if (!proj_active(s)) return EDGE_OK;
  if (!s->proj_tint)
    return edge::fail(EDGE_ERR_ARG, "neighbour step: no local step has written the face projections since set-up "
                                    "or the last autotune (run the local step first)");
  if (s->proj_mixed || s->proj_tint != t_int)
    return edge::fail(EDGE_ERR_ARG, "neighbour step: t_int is not the array the local step wrote; this solver "
                                    "exchanges face projections and ignores t_int (create it with kernel 3 to read "
                                    "a substituted tInt)");
  s->proj_fresh = false;
  return EDGE_OK;
}

template <typename T>
int local_t(edge_el_solver* s, int64_t first, int64_t n, double dt, void* dofs, void* t_int, hipStream_t st) {
  if (s->n_rms > 0) return edge::fail(EDGE_ERR_ARG, "viscoelastic solver: the step needs dofs_a (edge_el_local_a)");
  const auto d = dev_view<T>(s);
  note_local(s, t_int);
  if (s->kernel == 2) return edge::launch_local_mfma<T>(d, first, n, (T)dt, (T*)dofs, (T*)t_int, st);
  if (s->kernel == 5) return edge::launch_local_ec<T>(d, ec_view<T>(s), first, n, (T)dt, (T*)dofs, (T*)t_int, st);
  if constexpr (sizeof(T) == 8)
    if (s->kernel == 4) return edge::launch_local_plain64(d, first, n, dt, (double*)dofs, (double*)t_int, st);
  return edge::launch_local_generic<T>(d, first, n, (T)dt, (T*)dofs, (T*)t_int, st);
}

template <typename T>
int neigh_t(edge_el_solver* s, int64_t first, int64_t n, const void* t_int, void* dofs, hipStream_t st) {
  if (s->n_rms > 0) return edge::fail(EDGE_ERR_ARG, "viscoelastic solver: the step needs dofs_a (edge_el_neigh_a)");
  const auto d = dev_view<T>(s);
  if (int rc = check_neigh(s, t_int)) return rc;
  if (s->kernel == 2) return edge::launch_neigh_mfma<T>(d, first, n, (const T*)t_int, (T*)dofs, st);
  if (s->kernel == 5) return edge::launch_neigh_ec<T>(d, ec_view<T>(s), first, n, (T*)dofs, st);
  if constexpr (sizeof(T) == 8)
    if (s->kernel == 4) return edge::launch_neigh_plain64(d, first, n, (const double*)t_int, (double*)dofs, st);
  return edge::launch_neigh_generic<T>(d, first, n, (const T*)t_int, (T*)dofs, st);
}

void drop_lts(edge_el_solver* s) {
  for (void* p : {(void*)s->fa_lts, (void*)s->ne_lts, (void*)s->rcv_lts, (void*)s->lts_lt, (void*)s->lts_gt,
                  (void*)s->lts_if})
    if (p) (void)hipFree(p);
  s->fa_lts = nullptr;
  s->ne_lts = s->rcv_lts = nullptr;
  s->lts_lt = s->lts_gt = s->lts_if = nullptr;
  s->h_lts_lt.clear();
  s->h_lts_gt.clear();
  s->h_lts_if.clear();
}

// [lo, hi) of the entries of an ascending element list inside [first, first + n)
std::pair<int64_t, int64_t> list_range(const std::vector<int64_t>& l, int64_t first, int64_t n) {
  const auto a = std::lower_bound(l.begin(), l.end(), first), b = std::lower_bound(l.begin(), l.end(), first + n);
  return {a - l.begin(), b - l.begin()};
}

int check_range(const edge_el_solver* s, int64_t first, int64_t n) {
  if (!s) return edge::fail(EDGE_ERR_ARG, "solver handle missing");
  if (first < 0 || n < 0 || first + n > s->n_el) return edge::fail(EDGE_ERR_ARG, "element range out of bounds");
  return EDGE_OK;
}

template <typename T>
int local_lts_t(edge_el_solver* s, int64_t first, int64_t n, double dt, int first_ts, void* dofs, void* t0, void* t1,
                void* t2, hipStream_t st) {
  const auto g = list_range(s->h_lts_gt, first, n);
  int rc = edge::launch_lts_half<T>(dev_view<T>(s), s->lts_gt + g.first, g.second - g.first, (T)(0.5 * dt),
                                    (const T*)dofs, (T*)t2, st);
  if (rc) return rc;
  rc = local_t<T>(s, first, n, dt, dofs, t0, st);
  if (rc) return rc;
  const auto l = list_range(s->h_lts_lt, first, n);
  return edge::launch_lts_sum<T>(s->lts_lt + l.first, l.second - l.first, 9 * s->n_modes * s->cfg.n_runs,
                                 first_ts, (const T*)t0, (T*)t1, st);
}

template <typename T>
int neigh_lts_t(edge_el_solver* s, int64_t first, int64_t n, int first_ts, const void* t0, const void* t1,
                const void* t2, void* dofs, hipStream_t st) {
  auto d = dev_view<T>(s);
  d.nb.ne = s->ne_lts;
  if (d.rcv) d.rcv = s->rcv_lts;
  int rc = check_neigh(s, t0);
  if (rc) return rc;
  if (s->kernel == 2)
    rc = edge::launch_neigh_mfma<T>(d, first, n, (const T*)t0, (T*)dofs, st);
  else if (s->kernel == 5)
    rc = edge::launch_neigh_ec<T>(d, ec_view<T>(s), first, n, (T*)dofs, st);
  else if constexpr (sizeof(T) == 8)
    rc = s->kernel == 4 ? edge::launch_neigh_plain64(d, first, n, (const double*)t0, (double*)dofs, st)
                        : edge::launch_neigh_generic<T>(d, first, n, (const T*)t0, (T*)dofs, st);
  else
    rc = edge::launch_neigh_generic<T>(d, first, n, (const T*)t0, (T*)dofs, st);
  if (rc) return rc;
  const auto f = list_range(s->h_lts_if, first, n);
  return edge::launch_lts_faces<T>(dev_view<T>(s), s->fa_lts, s->lts_if + f.first, f.second - f.first, first_ts,
                                   (const T*)t0, (const T*)t1, (const T*)t2, (T*)dofs, st);
}

// AderDg::local / neigh with local time stepping and anelastic DOFs
template <typename T>
int local_lts_a_t(edge_el_solver* s, int64_t first, int64_t n, double dt, int first_ts, void* dofs, void* dofs_a,
                  void* t0, void* t1, void* t2, hipStream_t st) {
  const auto g = list_range(s->h_lts_gt, first, n);
  int rc = edge::launch_lts_half_visco<T>(dev_view<T>(s), s->lts_gt + g.first, g.second - g.first, (T)(0.5 * dt),
                                          (const T*)dofs, (const T*)dofs_a, (T*)t2, st);
  if (rc) return rc;
  rc = edge_el_local_a(s, first, n, dt, dofs, dofs_a, t0, st);
  if (rc) return rc;
  const auto l = list_range(s->h_lts_lt, first, n);
  return edge::launch_lts_sum<T>(s->lts_lt + l.first, l.second - l.first, 9 * s->n_modes * s->cfg.n_runs,
                                 first_ts, (const T*)t0, (T*)t1, st);
}

template <typename T>
int neigh_lts_a_t(edge_el_solver* s, int64_t first, int64_t n, int first_ts, const void* t0, const void* t1,
                  const void* t2, void* dofs, void* dofs_a, hipStream_t st) {
  auto d = dev_view<T>(s);
  d.nb.ne = s->ne_lts;
  if (d.rcv) d.rcv = s->rcv_lts;
  int rc = check_neigh(s, t0);
  if (rc) return rc;
  if (s->kernel == 2) {
    if constexpr (sizeof(T) == 4)
      rc = edge::launch_neigh_visco_mfma(d, first, n, (const float*)t0, (float*)dofs, (float*)dofs_a, st);
  } else if (s->kernel == 4) {
    if constexpr (sizeof(T) == 8)
      rc = edge::launch_neigh_plain64_visco(d, first, n, (const double*)t0, (double*)dofs, (double*)dofs_a, st);
  } else {
    rc = edge::launch_neigh_visco<T>(d, first, n, (const T*)t0, (T*)dofs, (T*)dofs_a, st);
  }
  if (rc) return rc;
  const auto f = list_range(s->h_lts_if, first, n);
  return edge::launch_lts_faces_visco<T>(dev_view<T>(s), s->fa_lts, s->lts_if + f.first, f.second - f.first,
                                         first_ts, (const T*)t0, (const T*)t1, (const T*)t2, (T*)dofs, (T*)dofs_a,
                                         st);
}

int one_step(edge_el_solver* s, double dt, void* dofs, void* dofs_a, void* t_int, hipStream_t st) {
  // a whole step writes the projections its neighbour step reads: a local
  // step pending from before is superseded
  s->proj_fresh = false;
  s->proj_mixed = false;
  if (s->n_rms > 0) {
    const int rc = edge_el_local_a(s, 0, s->n_el, dt, dofs, dofs_a, t_int, st);
    return rc ? rc : edge_el_neigh_a(s, 0, s->n_el, t_int, dofs, dofs_a, st);
  }
  int rc = s->cfg.precision == 64 ? local_t<double>(s, 0, s->n_el, dt, dofs, t_int, st)
                                  : local_t<float>(s, 0, s->n_el, dt, dofs, t_int, st);
  if (rc) return rc;
  if (s->fault && (rc = edge::fault_apply(s->fault, dt, t_int, dofs, st))) return rc;
  return s->cfg.precision == 64 ? neigh_t<double>(s, 0, s->n_el, t_int, dofs, st)
                                : neigh_t<float>(s, 0, s->n_el, t_int, dofs, st);
}

// Multi-step calls of FP32 MFMA elastic solvers that exchange face
// projections fuse the neighbour step of step k - 1 with the local step of
// step k (one launch per time step; the projections alternate between two
// buffers, since a launch reads the previous step's while it writes its
// own).  Any fault phase runs between the launches, where it runs between
// local and neighbour step otherwise.
bool fused_active(const edge_el_solver* s) {
  return s->kernel == 2 && s->use_fused && s->cfg.precision == 32 && proj_active(s) &&
         s->cfg.n_runs == (s->n_rms == 0 ? 16 : 8);
}

// The viscoelastic form of fused_steps (FP32, 8 runs): the same launch
// sequence with the anelastic DOFs.
int fused_steps_a(edge_el_solver* s, int64_t n_steps, double dt, void* dofs, void* dofs_a, void* t_int,
                  hipStream_t st) {
  float* P[2] = {static_cast<float*>(s->proj), static_cast<float*>(s->proj_alt)};
  auto d = dev_view<float>(s);
  const float fdt = (float)dt;
  float *q = (float*)dofs, *qa = (float*)dofs_a, *ti = (float*)t_int;
  d.proj = P[0];
  int rc = edge::launch_local_visco_mfma(d, 0, s->n_el, fdt, q, qa, ti, st);
  for (int64_t k = 1; k < n_steps && rc == EDGE_OK; k++) {
    d.proj = P[k % 2];
    rc = edge::launch_fused_visco_mfma(d, fdt, P[(k - 1) % 2], q, qa, ti, st);
  }
  d.proj = P[(n_steps - 1) % 2];
  if (rc == EDGE_OK) rc = edge::launch_neigh_visco_mfma(d, 0, s->n_el, ti, q, qa, st);
  return rc;
}

int fused_steps(edge_el_solver* s, int64_t n_steps, double dt, void* dofs, void* dofs_a, void* t_int,
                hipStream_t st) {
  // the fused kernels are the FP32 MFMA kernels of 16 elastic or 8
  // viscoelastic runs: nothing else may get here
  if (!fused_active(s) || (s->n_rms > 0) != (dofs_a != nullptr))
    return edge::fail(EDGE_ERR_UNSUPPORTED, "fused steps: FP32 MFMA solvers, 16 elastic or 8 viscoelastic runs");
  if (!s->proj_alt) EDGE_HIP_TRY(hipMalloc(&s->proj_alt, sizeof(float) * 4 * 768 * s->n_el));
  // the projection buffers no longer stand for a tInt array a separate
  // neighbour step could be handed
  s->proj_tint = nullptr;
  s->proj_fresh = false;
  if (s->n_rms > 0) return fused_steps_a(s, n_steps, dt, dofs, dofs_a, t_int, st);
  float* P[2] = {static_cast<float*>(s->proj), static_cast<float*>(s->proj_alt)};
  auto d = dev_view<float>(s);
  const float fdt = (float)dt;
  d.proj = P[0];
  int rc = edge::launch_local_mfma<float>(d, 0, s->n_el, fdt, (float*)dofs, (float*)t_int, st);
  if (rc == EDGE_OK && s->fault) rc = edge::fault_apply(s->fault, dt, t_int, dofs, st);
  for (int64_t k = 1; k < n_steps && rc == EDGE_OK; k++) {
    d.proj = P[k % 2];
    rc = edge::launch_fused_mfma(d, fdt, P[(k - 1) % 2], (float*)dofs, (float*)t_int, st);
    if (rc == EDGE_OK && s->fault) rc = edge::fault_apply(s->fault, dt, t_int, dofs, st);
  }
  d.proj = P[(n_steps - 1) % 2];
  if (rc == EDGE_OK) rc = edge::launch_neigh_mfma<float>(d, 0, s->n_el, (const float*)t_int, (float*)dofs, st);
  return rc;
}

void drop_graph(edge_el_solver* s) {
  if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  s->graph_exec = nullptr;
  s->graph = nullptr;
  s->graph_steps = 0;
}

}  // namespace

namespace {
// rho, lambda, mu per element (energy reduction) and the element maps
// (error norms) to the device
int upload_element_data(edge_el_solver* s, const edge_mesh_in* m, const double* rho, const double* lam,
                        const double* mu) {
  const size_t ne = s->n_el;
  EDGE_HIP_TRY(hipMalloc(&s->mat, sizeof(double) * 3 * ne));
  EDGE_HIP_TRY(hipMemcpy(s->mat, rho, sizeof(double) * ne, hipMemcpyHostToDevice));
  EDGE_HIP_TRY(hipMemcpy(s->mat + ne, lam, sizeof(double) * ne, hipMemcpyHostToDevice));
  EDGE_HIP_TRY(hipMemcpy(s->mat + 2 * ne, mu, sizeof(double) * ne, hipMemcpyHostToDevice));
  std::vector<double> geo(12 * ne);
  for (size_t el = 0; el < ne; el++) {
    const double* v0 = m->ve_crds + 3 * m->el_ve[el * 4];
    for (int d = 0; d < 3; d++) geo[el * 12 + d] = v0[d];
    for (int i = 1; i < 4; i++) {
      const double* vi = m->ve_crds + 3 * m->el_ve[el * 4 + i];
      for (int d = 0; d < 3; d++) geo[el * 12 + 3 * i + d] = vi[d] - v0[d];
    }
  }
  EDGE_HIP_TRY(hipMalloc(&s->geo, sizeof(double) * 12 * ne));
  EDGE_HIP_TRY(hipMemcpy(s->geo, geo.data(), sizeof(double) * 12 * ne, hipMemcpyHostToDevice));
  return EDGE_OK;
}
}  // namespace

extern "C" int edge_el_create(const edge_el_config* cfg, const edge_mesh_in* mesh, const double* rho,
                              const double* lam, const double* mu, edge_el_solver** out) {
  if (out) *out = nullptr;
  if (!mesh || !mesh->ve_crds || !mesh->el_ve || !mesh->normals || !mesh->tangents || !mesh->areas || !rho ||
      !lam || !mu)
    return edge::fail(EDGE_ERR_ARG, "mesh geometry or material missing");
  for (int64_t el = 0; mesh && el < mesh->n_el; el++)
    if (!(rho[el] > 0) || !(mu[el] > 0) || !(lam[el] + 2 * mu[el] > 0))
      return edge::fail(EDGE_ERR_ARG, "non-physical material parameters");
  edge_el_solver* s = nullptr;
  int rc = create_common(cfg, mesh, out, &s);
  if (rc == EDGE_OK) {
    int64_t nve = 0;
    for (int64_t i = 0; i < 4 * mesh->n_el; i++) nve = std::max(nve, mesh->el_ve[i] + 1);
    edge::SetupInput in{mesh->n_el, mesh->n_fa, nve,         mesh->ve_crds, mesh->el_ve, mesh->el_fa,
                        mesh->fa_el, mesh->fa_type, mesh->normals, mesh->tangents, mesh->areas, rho,
                        lam,         mu};
    s->sparse_star = true;  // built from the Jacobians' pattern only
    if (hipMalloc(&s->jdet, sizeof(double) * s->n_el) != hipSuccess)
      rc = edge::fail(EDGE_ERR_HIP, "allocation of Jacobian determinants");
    else
      rc = cfg->precision == 64
               ? edge::build_setup<double>(in, (double*)s->star, (double*)s->fs, s->jdet, (double*)s->fsf,
                                           &s->fs_rank3, nullptr, nullptr, nullptr, nullptr)
               : edge::build_setup<float>(in, (float*)s->star, (float*)s->fs, s->jdet, (float*)s->fsf,
                                          &s->fs_rank3, nullptr, nullptr, nullptr, nullptr);
    if (rc == EDGE_OK) rc = zero_fault_faces(s);
    if (rc == EDGE_OK) rc = finish_mfma(s);
    if (rc == EDGE_OK) rc = finish_ec(s);
    if (rc == EDGE_OK) rc = upload_element_data(s, mesh, rho, lam, mu);
  }
  if (rc != EDGE_OK) {
    if (s) edge_el_destroy(s);
    return rc;
  }
  *out = s;
  return EDGE_OK;
}

extern "C" int edge_el_create_raw(const edge_el_config* cfg, const edge_mesh_in* mesh, const double* star,
                                  const double* fs, edge_el_solver** out) {
  if (out) *out = nullptr;
  if (!star || !fs) return edge::fail(EDGE_ERR_ARG, "star matrices or flux solvers missing");
  edge_el_solver* s = nullptr;
  int rc = create_common(cfg, mesh, out, &s);
  if (rc == EDGE_OK) {
    const size_t ns = 243 * (size_t)s->n_el, nf = 648 * (size_t)s->n_el;
    // the sparse-star kernels apply only when every entry off the pattern is zero
    bool pat[81] = {};
    for (int q = 0; q < 9; q++)
      for (int j = 0; j < edge::kStarNnzRow[q]; j++) pat[q * 9 + edge::kStarCol[q][j]] = true;
    bool sparse = true;
    for (size_t i = 0; i < ns && sparse; i++) sparse = pat[i % 81] || star[i] == 0.0;
    s->sparse_star = sparse;
    if (cfg->precision == 64) {
      if (hipMemcpy(s->star, star, 8 * ns, hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(s->fs, fs, 8 * nf, hipMemcpyHostToDevice) != hipSuccess)
        rc = edge::fail(EDGE_ERR_HIP, "upload of star matrices / flux solvers");
    } else {
      std::vector<float> a(star, star + ns), b(fs, fs + nf);
      if (hipMemcpy(s->star, a.data(), 4 * ns, hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(s->fs, b.data(), 4 * nf, hipMemcpyHostToDevice) != hipSuccess)
        rc = edge::fail(EDGE_ERR_HIP, "upload of star matrices / flux solvers");
    }
    if (rc == EDGE_OK && s->fsf)
      rc = cfg->precision == 64 ? edge::factor_flux_solvers<double>(fs, 8 * s->n_el, (double*)s->fsf, &s->fs_rank3)
                                : edge::factor_flux_solvers<float>(fs, 8 * s->n_el, (float*)s->fsf, &s->fs_rank3);
    if (rc == EDGE_OK) rc = zero_fault_faces(s);
    if (rc == EDGE_OK) rc = finish_mfma(s);
    if (rc == EDGE_OK) rc = finish_ec(s);
  }
  if (rc != EDGE_OK) {
    if (s) edge_el_destroy(s);
    return rc;
  }
  *out = s;
  return EDGE_OK;
}

extern "C" int edge_el_destroy(edge_el_solver* s) {
  if (!s) return EDGE_OK;
  drop_graph(s);
  (void)hipFree(s->mats);
  (void)hipFree(s->pk);
  (void)hipFree(s->star);
  (void)hipFree(s->fs);
  (void)hipFree(s->ne);
  (void)hipFree(s->fmat);
  (void)hipFree(s->jdet);
  (void)hipFree(s->mass);
  (void)hipFree(s->mat);
  (void)hipFree(s->geo);
  (void)hipFree(s->fsf);
  (void)hipFree(s->proj);
  (void)hipFree(s->proj_alt);
  (void)hipFree(s->tap_on);
  (void)hipFree(s->snd_v);
  (void)hipFree(s->snd_fm);
  (void)hipFree(s->rcv);
  (void)hipFree(s->rfs);
  (void)hipFree(s->star_a);
  (void)hipFree(s->src_a);
  (void)hipFree(s->fs_a);
  (void)hipFree(s->fsfa);
  (void)hipFree(s->ec_pk);
  (void)hipFree(s->ec_rot);
  (void)hipFree(s->ec_proj);
  drop_lts(s);
  delete s;
  return EDGE_OK;
}

extern "C" int edge_el_get_setup(const edge_el_solver* s, double* star, double* fs) {
  if (!s) return edge::fail(EDGE_ERR_ARG, "solver handle missing");
  const size_t ns = 243 * (size_t)s->n_el, nf = 648 * (size_t)s->n_el;
  if (s->cfg.precision == 64) {
    if (star) EDGE_HIP_TRY(hipMemcpy(star, s->star, 8 * ns, hipMemcpyDeviceToHost));
    if (fs) EDGE_HIP_TRY(hipMemcpy(fs, s->fs, 8 * nf, hipMemcpyDeviceToHost));
  } else {
    std::vector<float> a;
    if (star) {
      a.resize(ns);
      EDGE_HIP_TRY(hipMemcpy(a.data(), s->star, 4 * ns, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < ns; i++) star[i] = a[i];
    }
    if (fs) {
      a.resize(nf);
      EDGE_HIP_TRY(hipMemcpy(a.data(), s->fs, 4 * nf, hipMemcpyDeviceToHost));
      for (size_t i = 0; i < nf; i++) fs[i] = a[i];
    }
  }
  return EDGE_OK;
}

namespace {

template <typename T>
int upload_as(void** dst, const std::vector<double>& h) {
  std::vector<T> v(h.begin(), h.end());
  EDGE_HIP_TRY(hipMalloc(dst, sizeof(T) * v.size()));
  EDGE_HIP_TRY(hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return EDGE_OK;
}

template <typename T>
int download_as(double* dst, const void* src, size_t n) {
  std::vector<T> v(n);
  EDGE_HIP_TRY(hipMemcpy(v.data(), src, sizeof(T) * n, hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n; i++) dst[i] = v[i];
  return EDGE_OK;
}

}  // namespace

extern "C" int edge_el_create_visco(const edge_el_config* cfg, const edge_visco_config* vc, const edge_mesh_in* mesh,
                                    const double* rho, const double* lam, const double* mu, const double* qp,
                                    const double* qs, edge_el_solver** out) {
  if (out) *out = nullptr;
  if (!cfg || !vc) return edge::fail(EDGE_ERR_ARG, "configuration missing");
  if (vc->n_rms < 1 || vc->n_rms > edge::kMaxMechs)
    return edge::fail(EDGE_ERR_UNSUPPORTED, "relaxation mechanisms must be 1..5");
  if (!(vc->freq_cen > 0) || !(vc->freq_rat > 1))
    return edge::fail(EDGE_ERR_ARG, "central frequency must be positive and the frequency ratio above 1");
  if (!mesh || !mesh->ve_crds || !mesh->el_ve || !mesh->normals || !mesh->tangents || !mesh->areas || !rho ||
      !lam || !mu || !qp || !qs)
    return edge::fail(EDGE_ERR_ARG, "mesh geometry, material or quality factors missing");
  if (cfg->kernel < 0 || cfg->kernel > 4) return edge::fail(EDGE_ERR_ARG, "kernel selector");
  const bool p64_ok = edge::plain64_visco_supported(cfg->order, cfg->n_runs, cfg->precision / 8, vc->n_rms);
  if (cfg->kernel == 4 && !p64_ok)
    return edge::fail(EDGE_ERR_UNSUPPORTED, "FP64 MFMA viscoelastic kernels: FP64, 1 run, 1-3 mechanisms");
  if (!edge::visco_supported(cfg->order, cfg->n_runs, cfg->precision / 8, vc->n_rms))
    return edge::fail(EDGE_ERR_UNSUPPORTED, "viscoelastic kernels: 1 or 8 runs whose tiles fit LDS");
  const bool vm_ok = edge::visco_mfma_supported(cfg->order, cfg->n_runs, cfg->precision, vc->n_rms);
  if ((cfg->kernel == 2 || cfg->kernel == 3) && !vm_ok)
    return edge::fail(EDGE_ERR_UNSUPPORTED, "MFMA viscoelastic kernels: FP32, 8 runs, orders 3-5, 1-3 mechanisms");
  const int64_t ne = mesh->n_el;
  for (int64_t el = 0; el < ne; el++)
    if (!(rho[el] > 0) || !(mu[el] > 0) || !(lam[el] + 2 * mu[el] > 0) || !(qp[el] > 0) || !(qs[el] > 0))
      return edge::fail(EDGE_ERR_ARG, "non-physical material parameters or quality factors");
  const int nr = vc->n_rms;
  // constant-Q fit per element (FP64, host)
  double freqs[2 * edge::kMaxMechs];
  edge::visco_frequencies(nr, vc->freq_cen, vc->freq_rat, freqs);
  std::vector<double> lam_e(ne), mu_e(ne), src((size_t)ne * nr * 36), rfs(nr);
  for (int rm = 0; rm < nr; rm++) rfs[rm] = freqs[2 * rm];
  for (int64_t el = 0; el < ne; el++) {
    edge::visco_element(nr, freqs, vc->freq_cen, qp[el], qs[el], lam[el], mu[el], &lam_e[el], &mu_e[el],
                        src.data() + (size_t)el * nr * 36);
    if (!(mu_e[el] > 0) || !(lam_e[el] + 2 * mu_e[el] > 0))
      return edge::fail(EDGE_ERR_ARG, "constant-Q fit gives non-physical unrelaxed Lame parameters");
  }
  // the MFMA kernels read the source matrices by their pattern: the
  // normal-stress block and the diagonal of the shear rows
  bool src_pattern = true;
  for (size_t i = 0; i < src.size(); i++) {
    const int row = (int)(i % 36) / 6, col = (int)(i % 6);
    const bool nz = row < 3 ? col < 3 : col == row;
    if (!nz && src[i] != 0) src_pattern = false;
  }
  const bool want_vm = vm_ok && cfg->kernel != 1 && cfg->kernel != 4 && src_pattern;
  if ((cfg->kernel == 2 || cfg->kernel == 3) && !want_vm)
    return edge::fail(EDGE_ERR_UNSUPPORTED, "MFMA viscoelastic kernels: source matrices off the pattern");
  if (cfg->kernel == 4 && !src_pattern)
    return edge::fail(EDGE_ERR_UNSUPPORTED, "FP64 MFMA viscoelastic kernels: source matrices off the pattern");
  edge_el_config c = *cfg;
  c.kernel = 1;
  edge_el_solver* s = nullptr;
  int rc = create_common(&c, mesh, out, &s, want_vm);
  if (rc == EDGE_OK) {
    s->n_rms = nr;
    s->lam_e = lam_e;
    s->mu_e = mu_e;
    s->h_rfs = rfs;
    int64_t nve = 0;
    for (int64_t i = 0; i < 4 * ne; i++) nve = std::max(nve, mesh->el_ve[i] + 1);
    edge::SetupInput in{ne, mesh->n_fa, nve, mesh->ve_crds, mesh->el_ve, mesh->el_fa, mesh->fa_el, mesh->fa_type,
                        mesh->normals, mesh->tangents, mesh->areas, rho, s->lam_e.data(), s->mu_e.data()};
    s->sparse_star = true;
    const size_t rb = s->real_bytes;
    if (hipMalloc(&s->jdet, sizeof(double) * ne) != hipSuccess || hipMalloc(&s->star_a, rb * 54 * ne) != hipSuccess ||
        hipMalloc(&s->fs_a, rb * 2 * 4 * 54 * ne) != hipSuccess)
      rc = edge::fail(EDGE_ERR_HIP, "allocation of the viscoelastic set-up");
    if (rc == EDGE_OK && want_vm &&
        (hipMalloc(&s->fsf, rb * 2 * 4 * 54 * ne) != hipSuccess || hipMalloc(&s->fsfa, rb * 2 * 4 * 18 * ne) != hipSuccess))
      rc = edge::fail(EDGE_ERR_HIP, "allocation of the factored flux solvers");
    if (rc == EDGE_OK)
      rc = cfg->precision == 64 ? upload_as<double>(&s->src_a, src) : upload_as<float>(&s->src_a, src);
    if (rc == EDGE_OK)
      rc = cfg->precision == 64 ? upload_as<double>(&s->rfs, rfs) : upload_as<float>(&s->rfs, rfs);
    if (rc == EDGE_OK)
      rc = cfg->precision == 64
               ? edge::build_setup<double>(in, (double*)s->star, (double*)s->fs, s->jdet, (double*)s->fsf,
                                           &s->fs_rank3, (double*)s->star_a, (double*)s->fs_a, (double*)s->fsfa,
                                           nullptr)
               : edge::build_setup<float>(in, (float*)s->star, (float*)s->fs, s->jdet, (float*)s->fsf,
                                          &s->fs_rank3, (float*)s->star_a, (float*)s->fs_a, (float*)s->fsfa,
                                          nullptr);
    if (rc == EDGE_OK) rc = zero_fault_faces(s);
    if (rc == EDGE_OK) rc = upload_element_data(s, mesh, rho, s->lam_e.data(), s->mu_e.data());
    if (rc == EDGE_OK && p64_ok && src_pattern && (cfg->kernel == 0 || cfg->kernel == 4)) s->kernel = 4;
    if (rc == EDGE_OK && want_vm) {
      if (s->fs_rank3) {
        s->kernel = 2;
        // face projections exchanged between the local and neighbour steps
        // ([nEl][4][512]: the paired projection tiles, lane-major)
        if (cfg->kernel != 3 && s->proj_ok && hipMalloc(&s->proj, sizeof(float) * 4 * 512 * ne) != hipSuccess)
          rc = edge::fail(EDGE_ERR_HIP, "allocation of the face projections");
      } else if (cfg->kernel == 2 || cfg->kernel == 3) {
        rc = edge::fail(EDGE_ERR_UNSUPPORTED, "MFMA viscoelastic kernels: flux solvers not of rank 3");
      } else {  // the generic kernels
        (void)hipFree(s->fsf);
        (void)hipFree(s->fsfa);
        s->fsf = s->fsfa = nullptr;
        s->fs_rank3 = false;
      }
    }
  }
  if (rc != EDGE_OK) {
    if (s) edge_el_destroy(s);
    return rc;
  }
  *out = s;
  return EDGE_OK;
}

extern "C" int edge_el_local_a(edge_el_solver* s, int64_t first, int64_t n_els, double dt, void* dofs, void* dofs_a,
                               void* t_int, void* stream) {
  int rc = check_range(s, first, n_els);
  if (rc) return rc;
  if (s->n_rms == 0) return edge_el_local(s, first, n_els, dt, dofs, t_int, stream);
  if (!dofs || !dofs_a || !t_int) return edge::fail(EDGE_ERR_ARG, "device arrays missing");
  hipStream_t st = (hipStream_t)stream;
  if (s->kernel == 4)
    return edge::launch_local_plain64_visco(dev_view<double>(s), first, n_els, dt, (double*)dofs, (double*)dofs_a,
                                            (double*)t_int, st);
  if (s->kernel == 2) {
    note_local(s, t_int);
    return edge::launch_local_visco_mfma(dev_view<float>(s), first, n_els, (float)dt, (float*)dofs, (float*)dofs_a,
                                         (float*)t_int, st);
  }
  if (s->cfg.precision == 64)
    return edge::launch_local_visco<double>(dev_view<double>(s), first, n_els, dt, (double*)dofs, (double*)dofs_a,
                                            (double*)t_int, st);
  return edge::launch_local_visco<float>(dev_view<float>(s), first, n_els, (float)dt, (float*)dofs, (float*)dofs_a,
                                         (float*)t_int, st);
}

extern "C" int edge_el_neigh_a(edge_el_solver* s, int64_t first, int64_t n_els, const void* t_int, void* dofs,
                               void* dofs_a, void* stream) {
  int rc = check_range(s, first, n_els);
  if (rc) return rc;
  if (s->n_rms == 0) return edge_el_neigh(s, first, n_els, t_int, dofs, stream);
  if (!dofs || !dofs_a || !t_int) return edge::fail(EDGE_ERR_ARG, "device arrays missing");
  hipStream_t st = (hipStream_t)stream;
  if (s->kernel == 4)
    return edge::launch_neigh_plain64_visco(dev_view<double>(s), first, n_els, (const double*)t_int, (double*)dofs,
                                            (double*)dofs_a, st);
  if (s->kernel == 2) {
    if ((rc = check_neigh(s, t_int))) return rc;
    return edge::launch_neigh_visco_mfma(dev_view<float>(s), first, n_els, (const float*)t_int, (float*)dofs,
                                         (float*)dofs_a, st);
  }
  if (s->cfg.precision == 64)
    return edge::launch_neigh_visco<double>(dev_view<double>(s), first, n_els, (const double*)t_int, (double*)dofs,
                                            (double*)dofs_a, st);
  return edge::launch_neigh_visco<float>(dev_view<float>(s), first, n_els, (const float*)t_int, (float*)dofs,
                                         (float*)dofs_a, st);
}

extern "C" int edge_el_get_setup_a(const edge_el_solver* s, double* star_a, double* src_a, double* fs_a, double* rfs,
                                   double* lam_e, double* mu_e) {
  if (!s) return edge::fail(EDGE_ERR_ARG, "solver handle missing");
  if (s->n_rms == 0) return edge::fail(EDGE_ERR_ARG, "not a viscoelastic solver");
  const size_t ne = s->n_el, nr = s->n_rms;
  const bool d = s->cfg.precision == 64;
  int rc = EDGE_OK;
  if (star_a && rc == EDGE_OK) rc = d ? download_as<double>(star_a, s->star_a, 54 * ne) : download_as<float>(star_a, s->star_a, 54 * ne);
  if (src_a && rc == EDGE_OK) rc = d ? download_as<double>(src_a, s->src_a, 36 * nr * ne) : download_as<float>(src_a, s->src_a, 36 * nr * ne);
  if (fs_a && rc == EDGE_OK) rc = d ? download_as<double>(fs_a, s->fs_a, 432 * ne) : download_as<float>(fs_a, s->fs_a, 432 * ne);
  if (rc) return rc;
  if (rfs)
    for (size_t i = 0; i < nr; i++) rfs[i] = s->h_rfs[i];
  if (lam_e)
    for (size_t i = 0; i < ne; i++) lam_e[i] = s->lam_e[i];
  if (mu_e)
    for (size_t i = 0; i < ne; i++) mu_e[i] = s->mu_e[i];
  return EDGE_OK;
}

namespace edge {
int solver_ref(const edge_el_solver* s, SolverRef* out) {
  if (!s) return fail(EDGE_ERR_ARG, "solver handle missing");
  out->precision = s->cfg.precision;
  out->order = s->cfg.order;
  out->n_runs = s->cfg.n_runs;
  out->n_modes = s->n_modes;
  out->n_rms = s->n_rms;
  out->n_el = s->n_el;
  out->jdet = s->jdet;
  out->mass = s->mass;
  if (s->cfg.precision == 64)
    out->d = dev_view<double>(s);
  else
    out->f = dev_view<float>(s);
  return EDGE_OK;
}
}  // namespace edge

extern "C" int edge_el_kernel_info(const edge_el_solver* s, int32_t* kernel, int32_t* flags) {
  if (!s) return edge::fail(EDGE_ERR_ARG, "solver handle missing");
  if (kernel) *kernel = s->kernel;
  if (flags) {
    const bool r3 = (s->kernel == 2 || s->kernel == 5) && s->fs_rank3 && s->fsf;
    *flags = (s->sparse_star ? 1 : 0) | (r3 ? 2 : 0) | (proj_active(s) ? 4 : 0) |
             (s->n_rms > 0 ? 8 : 0) | (fused_active(s) ? 16 : 0) | (s->hi_occ ? 32 : 0);
  }
  return EDGE_OK;
}

namespace {
struct TuneCand {
  int kernel;
  bool proj;
  bool xcd = true;
  bool fused = false;  // multi-step calls fuse neighbour step k - 1 with local step k
  bool occ = false;    // the higher-occupancy variant of the kernels (hi_occ)
};

// The kernel variants a solver can switch between with the data it holds.
std::vector<TuneCand> tune_candidates(edge_el_solver* s) {
  std::vector<TuneCand> c;
  if (!s->tune_kernel) s->tune_kernel = s->kernel;  // before the autotuner or a variant switch first changes it
  const int base = s->tune_kernel;
  const int prec_bytes = (int)s->real_bytes;
  const bool gen = s->n_rms > 0 ? edge::visco_supported(s->cfg.order, s->cfg.n_runs, prec_bytes, s->n_rms)
                                : edge::generic_supported(s->cfg.order, s->cfg.n_runs, prec_bytes);
  if (base == 2) {
    const bool fusable = s->proj && s->cfg.precision == 32;
    // the viscoelastic local kernel also comes with tIntA in LDS, five waves per SIMD
    const int n_occ = s->n_rms > 0 ? 2 : 1;
    for (int x = 1; x >= 0; x--)  // XCD-aware element schedule, then the linear one
      for (int o = 0; o < n_occ; o++) {
        if (fusable && o == 0) c.push_back({2, true, x == 1, true});
        if (s->proj) c.push_back({2, true, x == 1, false, o == 1});
        c.push_back({2, false, x == 1, false, o == 1});
      }
  } else if (base == 4 || base == 5) {
    const bool p64 = edge::plain64_supported(s->cfg.order, s->cfg.n_runs, prec_bytes);
    for (int x = 1; x >= 0; x--) {  // XCD-aware schedule, then the linear one
      if (p64) c.push_back({4, false, x == 1});
      if (s->ec_ok) c.push_back({5, true, x == 1});  // element-batched: projections always exchanged
    }
  }
  if (gen || c.empty()) c.push_back({1, false});
  return c;
}
}  // namespace

extern "C" int edge_el_autotune(edge_el_solver* s, double dt, int32_t n_reps, void* stream, int32_t* n_cands,
                                int32_t* kernels, int32_t* flags, double* ms) {
  if (!s) return edge::fail(EDGE_ERR_ARG, "solver handle missing");
  if (n_reps < 1 || !(dt > 0)) return edge::fail(EDGE_ERR_ARG, "autotune: n_reps >= 1 and dt > 0");
  const std::vector<TuneCand> cand = tune_candidates(s);
  hipStream_t st = (hipStream_t)stream;
  const size_t dof_bytes = s->real_bytes * 9 * s->n_modes * s->cfg.n_runs * s->n_el;
  const size_t a_bytes = dof_bytes / 9 * 6 * s->n_rms;
  void *dofs = nullptr, *t_int = nullptr, *dofs_a = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  const int old_kernel = s->kernel;
  const bool old_proj = s->use_proj, old_xcd = s->xcd_map, old_fused = s->use_fused, old_occ = s->hi_occ;
  std::vector<double> t(cand.size(), 0.0);
  int rc = EDGE_OK;
  auto hip = [&](hipError_t e, const char* what) {
    if (e != hipSuccess && rc == EDGE_OK) rc = edge::fail(EDGE_ERR_HIP, std::string("autotune: ") + what + ": " +
                                                                        hipGetErrorString(e));
    return rc == EDGE_OK;
  };
  // scratch DOFs (zeros: the same instruction stream as any data)
  if (hip(hipMalloc(&dofs, dof_bytes), "allocation") && hip(hipMalloc(&t_int, dof_bytes), "allocation") &&
      (s->n_rms == 0 || hip(hipMalloc(&dofs_a, a_bytes), "allocation")) &&
      hip(hipMemsetAsync(dofs, 0, dof_bytes, st), "memset") && hip(hipMemsetAsync(t_int, 0, dof_bytes, st), "memset") &&
      (s->n_rms == 0 || hip(hipMemsetAsync(dofs_a, 0, a_bytes, st), "memset")) &&
      hip(hipEventCreate(&e0), "event") && hip(hipEventCreate(&e1), "event")) {
    drop_graph(s);
    for (size_t i = 0; i < cand.size() && rc == EDGE_OK; i++) {
      s->kernel = cand[i].kernel;
      s->use_proj = cand[i].proj;
      s->xcd_map = cand[i].xcd;
      s->use_fused = cand[i].fused;
      s->hi_occ = cand[i].occ;
      rc = one_step(s, dt, dofs, dofs_a, t_int, st);  // warm-up
      if (rc == EDGE_OK) hip(hipEventRecord(e0, st), "event");
      if (cand[i].fused)  // one fused sequence of n_reps steps (n_reps + 1 launches)
        rc = n_reps >= 2 ? fused_steps(s, n_reps, dt, dofs, dofs_a, t_int, st)
                         : one_step(s, dt, dofs, dofs_a, t_int, st);
      else
        for (int r = 0; r < n_reps && rc == EDGE_OK; r++) rc = one_step(s, dt, dofs, dofs_a, t_int, st);
      if (rc == EDGE_OK && hip(hipEventRecord(e1, st), "event") && hip(hipEventSynchronize(e1), "sync")) {
        float f = 0.f;
        if (hip(hipEventElapsedTime(&f, e0, e1), "event")) t[i] = f / n_reps;
      }
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(dofs);
  (void)hipFree(t_int);
  (void)hipFree(dofs_a);
  // the candidates' steps overwrote the face projections: a neighbour step
  // pending from before must not read them
  s->proj_tint = nullptr;
  s->proj_fresh = false;
  if (rc != EDGE_OK) {
    s->kernel = old_kernel;
    s->use_proj = old_proj;
    s->xcd_map = old_xcd;
    s->use_fused = old_fused;
    s->hi_occ = old_occ;
    return rc;
  }
  size_t best = 0;
  for (size_t i = 1; i < cand.size(); i++)
    if (t[i] < t[best]) best = i;
  s->kernel = cand[best].kernel;
  s->use_proj = cand[best].proj;
  s->xcd_map = cand[best].xcd;
  s->use_fused = cand[best].fused;
  s->hi_occ = cand[best].occ;
  if (n_cands) {
    const int32_t cap = *n_cands;
    *n_cands = (int32_t)cand.size();
    for (int32_t i = 0; i < (int32_t)cand.size() && i < cap; i++) {
      if (kernels) kernels[i] = cand[i].kernel;
      if (flags)
        flags[i] = (cand[i].proj ? 1 : 0) | (cand[i].xcd ? 2 : 0) | (cand[i].fused ? 4 : 0) | (cand[i].occ ? 8 : 0);
      if (ms) ms[i] = t[i];
    }
  }
  return EDGE_OK;
}

extern "C" int edge_el_set_fault(edge_el_solver* s, edge_fault* f) {
  if (!s) return edge::fail(EDGE_ERR_ARG, "solver handle missing");
  if (f && edge::fault_solver(f) != s) return edge::fail(EDGE_ERR_ARG, "fault: built for another solver");
  if (f && s->n_rms > 0) return edge::fail(EDGE_ERR_UNSUPPORTED, "fault: elastic solvers only");
  drop_graph(s);
  s->fault = f;
  return EDGE_OK;
}

extern "C" int edge_el_set_variant(edge_el_solver* s, int32_t kernel, int32_t flags) {
  if (!s) return edge::fail(EDGE_ERR_ARG, "solver handle missing");
  for (const TuneCand& c : tune_candidates(s)) {
    if (c.kernel != kernel) continue;
    if (c.kernel != 1 &&
        (c.proj != ((flags & 1) != 0) || c.xcd != ((flags & 2) != 0) || c.fused != ((flags & 4) != 0) ||
         c.occ != ((flags & 8) != 0)))
      continue;
    drop_graph(s);
    s->kernel = c.kernel;
    s->use_proj = c.proj;
    s->xcd_map = c.xcd;
    s->use_fused = c.fused;
    s->hi_occ = c.occ;
    s->proj_tint = nullptr;
    s->proj_fresh = false;
    return EDGE_OK;
  }
  return edge::fail(EDGE_ERR_ARG, "not a kernel variant of this solver (see edge_el_autotune's candidates)");
}

extern "C" int edge_el_local(edge_el_solver* s, int64_t first, int64_t n_els, double dt, void* dofs, void* t_int,
                             void* stream) {
  int rc = check_range(s, first, n_els);
  if (rc) return rc;
  if (!dofs || !t_int) return edge::fail(EDGE_ERR_ARG, "device arrays missing");
  hipStream_t st = (hipStream_t)stream;
  return s->cfg.precision == 64 ? local_t<double>(s, first, n_els, dt, dofs, t_int, st)
                                : local_t<float>(s, first, n_els, dt, dofs, t_int, st);
}

extern "C" int edge_el_neigh(edge_el_solver* s, int64_t first, int64_t n_els, const void* t_int, void* dofs,
                             void* stream) {
  int rc = check_range(s, first, n_els);
  if (rc) return rc;
  if (!dofs || !t_int) return edge::fail(EDGE_ERR_ARG, "device arrays missing");
  hipStream_t st = (hipStream_t)stream;
  return s->cfg.precision == 64 ? neigh_t<double>(s, first, n_els, t_int, dofs, st)
                                : neigh_t<float>(s, first, n_els, t_int, dofs, st);
}

extern "C" int edge_el_set_lts(edge_el_solver* s, const uint8_t* el_lts, const uint8_t* fa_lts) {
  using edge::fail;
  if (!s) return fail(EDGE_ERR_ARG, "solver handle missing");
  drop_lts(s);
  if (!el_lts && !fa_lts) return EDGE_OK;
  if (!el_lts || !fa_lts) return fail(EDGE_ERR_ARG, "local time stepping: element and face characteristics");
  if (s->n_rms > 0 ? !edge::lts_visco_supported(s->cfg.order, s->cfg.n_runs, s->cfg.precision / 8, s->n_rms)
                   : !edge::lts_supported(s->cfg.order, s->cfg.n_runs, s->cfg.precision / 8))
    return fail(EDGE_ERR_UNSUPPORTED, "local time stepping: order, runs and precision do not fit one workgroup");
  const int64_t ne4 = s->n_el * 4;
  std::vector<int32_t> ne(ne4), rcv;
  EDGE_HIP_TRY(hipMemcpy(ne.data(), s->ne, sizeof(int32_t) * ne4, hipMemcpyDeviceToHost));
  if (s->rcv) {
    rcv.resize(ne4);
    EDGE_HIP_TRY(hipMemcpy(rcv.data(), s->rcv, sizeof(int32_t) * ne4, hipMemcpyDeviceToHost));
  }
  std::vector<int64_t> lt, gt, itf;
  for (int64_t el = 0; el < s->n_el; el++) {
    bool any1 = false, any2 = false;
    for (int fa = 0; fa < 4; fa++) {
      const int m = fa_lts[el * 4 + fa];
      if (m > 2) return fail(EDGE_ERR_ARG, "local time stepping: face characteristic not 0, 1 or 2");
      if (m != 0 && (ne[el * 4 + fa] < 0 || ne[el * 4 + fa] == el))
        return fail(EDGE_ERR_ARG, "local time stepping: an interface face without a neighbouring element");
      any1 |= m == 1;
      any2 |= m == 2;
      if (m != 0) {
        ne[el * 4 + fa] = -1;
        if (!rcv.empty()) rcv[el * 4 + fa] = -1;
      }
    }
    if (el_lts[el] > 3 || ((el_lts[el] & 1) != 0) != any2 || ((el_lts[el] & 2) != 0) != any1)
      return fail(EDGE_ERR_ARG, "local time stepping: element characteristics do not match its faces");
    if (any2) lt.push_back(el);
    if (any1) gt.push_back(el);
    if (any1 || any2) itf.push_back(el);
  }
  hipError_t e = hipMalloc((void**)&s->fa_lts, ne4);
  if (e == hipSuccess) e = hipMemcpy(s->fa_lts, fa_lts, ne4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc((void**)&s->ne_lts, sizeof(int32_t) * ne4);
  if (e == hipSuccess) e = hipMemcpy(s->ne_lts, ne.data(), sizeof(int32_t) * ne4, hipMemcpyHostToDevice);
  if (e == hipSuccess && !rcv.empty()) e = hipMalloc((void**)&s->rcv_lts, sizeof(int32_t) * ne4);
  if (e == hipSuccess && !rcv.empty())
    e = hipMemcpy(s->rcv_lts, rcv.data(), sizeof(int32_t) * ne4, hipMemcpyHostToDevice);
  for (auto pr : {std::make_pair(&s->lts_lt, &lt), std::make_pair(&s->lts_gt, &gt), std::make_pair(&s->lts_if, &itf)}) {
    if (e == hipSuccess) e = hipMalloc((void**)pr.first, sizeof(int64_t) * (pr.second->size() + 1));
    if (e == hipSuccess && !pr.second->empty())
      e = hipMemcpy(*pr.first, pr.second->data(), sizeof(int64_t) * pr.second->size(), hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    drop_lts(s);
    return fail(EDGE_ERR_HIP, std::string("local time stepping: ") + hipGetErrorString(e));
  }
  s->h_lts_lt = lt;
  s->h_lts_gt = gt;
  s->h_lts_if = itf;
  return EDGE_OK;
}

extern "C" int edge_el_local_lts_a(edge_el_solver* s, int64_t first, int64_t n_els, double dt, int32_t first_ts,
                                   void* dofs, void* dofs_a, void* t_int0, void* t_int1, void* t_int2, void* stream) {
  int rc = check_range(s, first, n_els);
  if (rc) return rc;
  if (s->n_rms == 0) return edge_el_local_lts(s, first, n_els, dt, first_ts, dofs, t_int0, t_int1, t_int2, stream);
  if (!s->fa_lts) return edge::fail(EDGE_ERR_ARG, "local time stepping is not set up (edge_el_set_lts)");
  if (!dofs || !dofs_a || !t_int0 || !t_int1 || !t_int2) return edge::fail(EDGE_ERR_ARG, "device arrays missing");
  hipStream_t st = (hipStream_t)stream;
  return s->cfg.precision == 64
             ? local_lts_a_t<double>(s, first, n_els, dt, first_ts, dofs, dofs_a, t_int0, t_int1, t_int2, st)
             : local_lts_a_t<float>(s, first, n_els, dt, first_ts, dofs, dofs_a, t_int0, t_int1, t_int2, st);
}

extern "C" int edge_el_neigh_lts_a(edge_el_solver* s, int64_t first, int64_t n_els, int32_t first_ts,
                                   const void* t_int0, const void* t_int1, const void* t_int2, void* dofs,
                                   void* dofs_a, void* stream) {
  int rc = check_range(s, first, n_els);
  if (rc) return rc;
  if (s->n_rms == 0) return edge_el_neigh_lts(s, first, n_els, first_ts, t_int0, t_int1, t_int2, dofs, stream);
  if (!s->fa_lts) return edge::fail(EDGE_ERR_ARG, "local time stepping is not set up (edge_el_set_lts)");
  if (!dofs || !dofs_a || !t_int0 || !t_int1 || !t_int2) return edge::fail(EDGE_ERR_ARG, "device arrays missing");
  hipStream_t st = (hipStream_t)stream;
  return s->cfg.precision == 64
             ? neigh_lts_a_t<double>(s, first, n_els, first_ts, t_int0, t_int1, t_int2, dofs, dofs_a, st)
             : neigh_lts_a_t<float>(s, first, n_els, first_ts, t_int0, t_int1, t_int2, dofs, dofs_a, st);
}

extern "C" int edge_el_local_lts(edge_el_solver* s, int64_t first, int64_t n_els, double dt, int32_t first_ts,
                                 void* dofs, void* t_int0, void* t_int1, void* t_int2, void* stream) {
  int rc = check_range(s, first, n_els);
  if (rc) return rc;
  if (s->n_rms > 0) return edge::fail(EDGE_ERR_ARG, "viscoelastic solver: the step needs dofs_a (edge_el_local_lts_a)");
  if (!s->fa_lts) return edge::fail(EDGE_ERR_ARG, "local time stepping is not set up (edge_el_set_lts)");
  if (!dofs || !t_int0 || !t_int1 || !t_int2) return edge::fail(EDGE_ERR_ARG, "device arrays missing");
  hipStream_t st = (hipStream_t)stream;
  return s->cfg.precision == 64 ? local_lts_t<double>(s, first, n_els, dt, first_ts, dofs, t_int0, t_int1, t_int2, st)
                                : local_lts_t<float>(s, first, n_els, dt, first_ts, dofs, t_int0, t_int1, t_int2, st);
}

extern "C" int edge_el_neigh_lts(edge_el_solver* s, int64_t first, int64_t n_els, int32_t first_ts,
                                 const void* t_int0, const void* t_int1, const void* t_int2, void* dofs,
                                 void* stream) {
  int rc = check_range(s, first, n_els);
  if (rc) return rc;
  if (s->n_rms > 0) return edge::fail(EDGE_ERR_ARG, "viscoelastic solver: the step needs dofs_a (edge_el_neigh_lts_a)");
  if (!s->fa_lts) return edge::fail(EDGE_ERR_ARG, "local time stepping is not set up (edge_el_set_lts)");
  if (!dofs || !t_int0 || !t_int1 || !t_int2) return edge::fail(EDGE_ERR_ARG, "device arrays missing");
  hipStream_t st = (hipStream_t)stream;
  return s->cfg.precision == 64 ? neigh_lts_t<double>(s, first, n_els, first_ts, t_int0, t_int1, t_int2, dofs, st)
                                : neigh_lts_t<float>(s, first, n_els, first_ts, t_int0, t_int1, t_int2, dofs, st);
}

namespace {
int step_impl(edge_el_solver* s, int64_t n_steps, double dt, void* dofs, void* dofs_a, void* t_int, int use_graph,
              hipStream_t st) {
  if (n_steps == 0) return EDGE_OK;
  const bool fuse = n_steps >= 2 && fused_active(s) && (s->n_rms > 0) == (dofs_a != nullptr);
  if (fuse && !s->proj_alt) EDGE_HIP_TRY(hipMalloc(&s->proj_alt, sizeof(float) * 4 * 768 * s->n_el));
  if (!use_graph && fuse) return fused_steps(s, n_steps, dt, dofs, dofs_a, t_int, st);
  if (!use_graph) {
    for (int64_t i = 0; i < n_steps; i++) {
      int rc = one_step(s, dt, dofs, dofs_a, t_int, st);
      if (rc) return rc;
    }
    return EDGE_OK;
  }
  // a fused sequence is captured whole (one replay per call), a single step
  // otherwise (one replay per step)
  const int64_t gsteps = fuse ? n_steps : 0;
  if (s->graph_exec && (s->graph_dofs != dofs || s->graph_dofs_a != dofs_a || s->graph_tint != t_int ||
                        s->graph_dt != dt || s->graph_stream != st || s->graph_steps != gsteps))
    drop_graph(s);
  if (!s->graph_exec) {
    // capture needs a non-default stream
    hipStream_t cap = st;
    bool own = false;
    if (!cap) {
      EDGE_HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
      own = true;
    }
    const hipError_t eb = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (eb != hipSuccess) {
      if (own) (void)hipStreamDestroy(cap);
      return edge::fail(EDGE_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(eb));
    }
    int rc = fuse ? fused_steps(s, n_steps, dt, dofs, dofs_a, t_int, cap) : one_step(s, dt, dofs, dofs_a, t_int, cap);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(cap, &g);
    if (rc == EDGE_OK && e != hipSuccess) rc = edge::fail(EDGE_ERR_HIP, hipGetErrorString(e));
    if (rc == EDGE_OK) {
      e = hipGraphInstantiate(&s->graph_exec, g, nullptr, nullptr, 0);
      if (e != hipSuccess) rc = edge::fail(EDGE_ERR_HIP, hipGetErrorString(e));
    }
    s->graph = g;
    if (own) (void)hipStreamDestroy(cap);
    if (rc) {
      drop_graph(s);
      return rc;
    }
    s->graph_dofs = dofs;
    s->graph_dofs_a = dofs_a;
    s->graph_tint = t_int;
    s->graph_dt = dt;
    s->graph_stream = st;
    s->graph_steps = gsteps;
  }
  if (gsteps) {
    s->proj_tint = nullptr;  // as after an eager fused sequence
    s->proj_fresh = false;
    EDGE_HIP_TRY(hipGraphLaunch(s->graph_exec, st));
    return EDGE_OK;
  }
  for (int64_t i = 0; i < n_steps; i++) EDGE_HIP_TRY(hipGraphLaunch(s->graph_exec, st));
  return EDGE_OK;
}
}  // namespace

extern "C" int edge_el_step(edge_el_solver* s, int64_t n_steps, double dt, void* dofs, void* t_int, int use_graph,
                            void* stream) {
  if (!s) return edge::fail(EDGE_ERR_ARG, "solver handle missing");
  if (!dofs || !t_int || n_steps < 0) return edge::fail(EDGE_ERR_ARG, "step arguments");
  if (s->n_rms > 0) return edge::fail(EDGE_ERR_ARG, "viscoelastic solver: the step needs dofs_a (edge_el_step_a)");
  return step_impl(s, n_steps, dt, dofs, nullptr, t_int, use_graph, (hipStream_t)stream);
}

extern "C" int edge_el_step_a(edge_el_solver* s, int64_t n_steps, double dt, void* dofs, void* dofs_a, void* t_int,
                              int use_graph, void* stream) {
  if (!s) return edge::fail(EDGE_ERR_ARG, "solver handle missing");
  if (s->n_rms == 0) return edge_el_step(s, n_steps, dt, dofs, t_int, use_graph, stream);
  if (!dofs || !dofs_a || !t_int || n_steps < 0) return edge::fail(EDGE_ERR_ARG, "step arguments");
  return step_impl(s, n_steps, dt, dofs, dofs_a, t_int, use_graph, (hipStream_t)stream);
}

extern "C" int edge_el_neigh_local(edge_el_solver* s, double dt, void* dofs, void* t_int, void* stream) {
  if (!s) return edge::fail(EDGE_ERR_ARG, "solver handle missing");
  if (!dofs || !t_int) return edge::fail(EDGE_ERR_ARG, "device arrays missing");
  if (!fused_active(s) || s->n_rms > 0)
    return edge::fail(EDGE_ERR_UNSUPPORTED, "fused neighbour and local step: FP32 MFMA elastic solvers that exchange "
                                            "face projections (edge_el_kernel_info flag bit 4)");
  if (!s->proj_fresh)
    return edge::fail(EDGE_ERR_ARG, "fused neighbour and local step: no local step is pending (run edge_el_local "
                                    "first, or its neighbour step has run)");
  int rc = check_neigh(s, t_int);
  if (rc) return rc;
  if (!s->proj_alt) EDGE_HIP_TRY(hipMalloc(&s->proj_alt, sizeof(float) * 4 * 768 * s->n_el));
  auto d = dev_view<float>(s);
  d.proj = static_cast<float*>(s->proj_alt);
  rc = edge::launch_fused_mfma(d, (float)dt, static_cast<const float*>(s->proj), (float*)dofs, (float*)t_int,
                               (hipStream_t)stream);
  if (rc) return rc;
  std::swap(s->proj, s->proj_alt);
  note_local(s, t_int);
  return EDGE_OK;
}

extern "C" int edge_el_set_taps(edge_el_solver* s, int64_t n, const int64_t* el, void* taps) {
  if (!s) return edge::fail(EDGE_ERR_ARG, "solver handle missing");
  if (n < 0 || (n > 0 && (!el || !taps))) return edge::fail(EDGE_ERR_ARG, "taps: elements or array missing");
  for (int64_t i = 0; i < n; i++)
    if (el[i] < 0 || el[i] >= s->n_el) return edge::fail(EDGE_ERR_ARG, "taps: element out of range");
  if (n > 0 && s->n_rms > 0)
    return edge::fail(EDGE_ERR_UNSUPPORTED, "taps: elastic solvers (the fused viscoelastic launch writes none)");
  drop_graph(s);
  s->tap = nullptr;
  if (n == 0) return EDGE_OK;
  std::vector<uint8_t> on(s->n_el, 0);
  for (int64_t i = 0; i < n; i++) on[el[i]] = 1;
  if (!s->tap_on) EDGE_HIP_TRY(hipMalloc(&s->tap_on, s->n_el));
  EDGE_HIP_TRY(hipMemcpy(s->tap_on, on.data(), s->n_el, hipMemcpyHostToDevice));
  s->tap = taps;
  return EDGE_OK;
}

extern "C" int edge_el_info(const edge_el_solver* s, int64_t* n_el, int32_t* n_modes, int32_t* n_runs,
                            int64_t* dof_bytes) {
  if (!s) return edge::fail(EDGE_ERR_ARG, "solver handle missing");
  if (n_el) *n_el = s->n_el;
  if (n_modes) *n_modes = s->n_modes;
  if (n_runs) *n_runs = s->cfg.n_runs;
  if (dof_bytes) *dof_bytes = (int64_t)s->real_bytes * 9 * s->n_modes * s->cfg.n_runs * s->n_el;
  return EDGE_OK;
}

namespace {

// Per-block partial sums of |J| M_n (a - b)^2 and |J| M_n b^2 per (quantity, run).
template <typename TA, typename TB>
__global__ __launch_bounds__(256) void k_l2(int64_t n_el, int n_md, int n_crs, const double* __restrict__ jdet,
                                            const double* __restrict__ mass, const TA* __restrict__ a,
                                            const TB* __restrict__ b, double* __restrict__ part) {
  // one block per chunk of elements; thread t accumulates entry t of [9][R]
  const int nqr = 9 * n_crs;
  const int64_t per = (n_el + gridDim.x - 1) / gridDim.x;
  const int64_t e0 = blockIdx.x * per, e1 = min(n_el, e0 + per);
  for (int qr = threadIdx.x; qr < nqr; qr += blockDim.x) {
    const int q = qr / n_crs, r = qr % n_crs;
    double d2 = 0, r2 = 0;
    for (int64_t el = e0; el < e1; el++) {
      const double j = fabs(jdet[el]);
      const size_t base = ((size_t)el * 9 + q) * n_md * n_crs + r;
      for (int md = 0; md < n_md; md++) {
        const double av = (double)a[base + (size_t)md * n_crs];
        const double bv = b ? (double)b[base + (size_t)md * n_crs] : av;
        const double w = j * mass[md];
        d2 += w * (av - bv) * (av - bv);
        r2 += w * bv * bv;
      }
    }
    part[(size_t)blockIdx.x * 2 * nqr + qr] = d2;
    part[(size_t)blockIdx.x * 2 * nqr + nqr + qr] = r2;
  }
}

template <typename TA, typename TB>
int l2_launch(edge_el_solver* s, const void* a, const void* b, double* d_part, unsigned nblk, hipStream_t st) {
  hipLaunchKernelGGL((k_l2<TA, TB>), dim3(nblk), dim3(256), 0, st, s->n_el, s->n_modes, s->cfg.n_runs, s->jdet,
                     s->mass, (const TA*)a, (const TB*)b, d_part);
  EDGE_HIP_TRY(hipGetLastError());
  return EDGE_OK;
}

}  // namespace

namespace edge {
// Per-block partial sums of the L2 norms (k_l2) into part[nblk][2][9 R].
int l2_partials(edge_el_solver* s, const void* a, int prec_a, const void* b, int prec_b, double* part, unsigned nblk,
                hipStream_t st) {
  const int pb = b ? prec_b : prec_a;
  if (prec_a == 64 && pb == 64) return l2_launch<double, double>(s, a, b, part, nblk, st);
  if (prec_a == 64) return l2_launch<double, float>(s, a, b, part, nblk, st);
  if (pb == 64) return l2_launch<float, double>(s, a, b, part, nblk, st);
  return l2_launch<float, float>(s, a, b, part, nblk, st);
}
}  // namespace edge

extern "C" int edge_el_l2_norms(edge_el_solver* s, const void* a, int prec_a, const void* b, int prec_b,
                                double* diff_sq, double* ref_sq, void* stream) {
  if (!s) return edge::fail(EDGE_ERR_ARG, "solver handle missing");
  if (!s->jdet) return edge::fail(EDGE_ERR_UNSUPPORTED, "norms need a solver built from geometry");
  if (!a || (prec_a != 32 && prec_a != 64) || (b && prec_b != 32 && prec_b != 64))
    return edge::fail(EDGE_ERR_ARG, "norm arguments");
  hipStream_t st = (hipStream_t)stream;
  const int nqr = 9 * s->cfg.n_runs;
  const unsigned nblk = edge::kL2Blocks;
  double* d_part = nullptr;
  EDGE_HIP_TRY(hipMalloc(&d_part, sizeof(double) * 2 * nqr * nblk));
  int rc = edge::l2_partials(s, a, prec_a, b, prec_b, d_part, nblk, st);
  std::vector<double> h(2 * nqr * nblk);
  if (rc == EDGE_OK &&
      hipMemcpyAsync(h.data(), d_part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st) == hipSuccess &&
      hipStreamSynchronize(st) == hipSuccess) {
    for (int i = 0; i < nqr; i++) {
      double d2 = 0, r2 = 0;
      for (unsigned k = 0; k < nblk; k++) {
        d2 += h[(size_t)k * 2 * nqr + i];
        r2 += h[(size_t)k * 2 * nqr + nqr + i];
      }
      if (diff_sq) diff_sq[i] = d2;
      if (ref_sq) ref_sq[i] = r2;
    }
  } else if (rc == EDGE_OK) {
    rc = edge::fail(EDGE_ERR_HIP, "norm reduction");
  }
  (void)hipFree(d_part);
  return rc;
}

namespace {
// Per-block partial sums per run of the kinetic energy 1/2 rho |v|^2 and the
// strain energy 1/2 sigma : C^-1 : sigma (isotropic compliance
// C^-1 sigma = sigma / (2 mu) - lambda / (2 mu (3 lambda + 2 mu)) tr(sigma) I),
// integrated exactly through |J| and the diagonal reference mass matrix.
template <typename T>
__global__ __launch_bounds__(256) void k_energy(int64_t n_el, int n_md, int n_crs, const double* __restrict__ jdet,
                                                const double* __restrict__ mass, const double* __restrict__ mat,
                                                const T* __restrict__ dofs, double* __restrict__ part) {
  __shared__ double sk[256], ss[256];
  const int64_t per = (n_el + gridDim.x - 1) / gridDim.x;
  const int64_t e0 = blockIdx.x * per, e1 = min(n_el, e0 + per);
  const int t = threadIdx.x, r = t % n_crs, m0 = t / n_crs, mstep = blockDim.x / n_crs;
  double ek = 0, es = 0;
  if (m0 < mstep) {
    for (int64_t el = e0; el < e1; el++) {
      const double j = fabs(jdet[el]), rho = mat[el], lam = mat[n_el + el], mu = mat[2 * n_el + el];
      const double c1 = 0.5 / mu, c2 = lam / (2 * mu * (3 * lam + 2 * mu));
      for (int md = m0; md < n_md; md += mstep) {
        double x[9];
        for (int q = 0; q < 9; q++) x[q] = (double)dofs[(((size_t)el * 9 + q) * n_md + md) * n_crs + r];
        const double w = j * mass[md];
        const double tr = x[0] + x[1] + x[2];
        const double ss2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + 2 * (x[3] * x[3] + x[4] * x[4] + x[5] * x[5]);
        ek += w * 0.5 * rho * (x[6] * x[6] + x[7] * x[7] + x[8] * x[8]);
        es += w * 0.5 * (c1 * ss2 - c2 * tr * tr);
      }
    }
  }
  sk[t] = ek;
  ss[t] = es;
  __syncthreads();
  if (t < n_crs) {
    double k = 0, s2 = 0;
    for (int i = 0; i < mstep; i++) {
      k += sk[i * n_crs + t];
      s2 += ss[i * n_crs + t];
    }
    part[(size_t)blockIdx.x * 2 * n_crs + t] = k;
    part[(size_t)blockIdx.x * 2 * n_crs + n_crs + t] = s2;
  }
}
}  // namespace

extern "C" int edge_el_energy(edge_el_solver* s, const void* dofs, double* kinetic, double* strain, void* stream) {
  if (!s) return edge::fail(EDGE_ERR_ARG, "solver handle missing");
  if (!s->jdet || !s->mat) return edge::fail(EDGE_ERR_UNSUPPORTED, "the energy needs a solver built from materials");
  if (!dofs) return edge::fail(EDGE_ERR_ARG, "device DOFs missing");
  const int R = s->cfg.n_runs;
  if (256 % R) return edge::fail(EDGE_ERR_UNSUPPORTED, "energy: runs must divide 256");
  hipStream_t st = (hipStream_t)stream;
  const unsigned nblk = 1024;
  double* d_part = nullptr;
  EDGE_HIP_TRY(hipMalloc(&d_part, sizeof(double) * 2 * R * nblk));
  if (s->cfg.precision == 64)
    hipLaunchKernelGGL(k_energy<double>, dim3(nblk), dim3(256), 0, st, s->n_el, s->n_modes, R, s->jdet, s->mass,
                       s->mat, (const double*)dofs, d_part);
  else
    hipLaunchKernelGGL(k_energy<float>, dim3(nblk), dim3(256), 0, st, s->n_el, s->n_modes, R, s->jdet, s->mass,
                       s->mat, (const float*)dofs, d_part);
  int rc = EDGE_OK;
  std::vector<double> h(2 * R * nblk);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(h.data(), d_part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    rc = edge::fail(EDGE_ERR_HIP, "energy reduction");
  (void)hipFree(d_part);
  if (rc) return rc;
  for (int r = 0; r < R; r++) {
    double k = 0, e = 0;
    for (unsigned b = 0; b < nblk; b++) {
      k += h[(size_t)b * 2 * R + r];
      e += h[(size_t)b * 2 * R + R + r];
    }
    if (kinetic) kinetic[r] = k;
    if (strain) strain[r] = e;
  }
  return EDGE_OK;
}

namespace {
// Per-block L1, squared L2 and Linf errors per (quantity, run) at the
// quadrature points of every element against a sum of plane waves
// q(x, t) = sum_w r_w sin(k_w . x - omega_w t + phase_w) (per run:
// waves [R][n_waves][14] = r[9], k[3], omega, phase).  Thread t keeps entry
// t % (9 R) for elements t / (9 R), t / (9 R) + lanes, ... of the block's
// range; partials summed in a fixed order.
template <typename T>
__global__ __launch_bounds__(256) void k_err(int64_t n_el, int n_md, int n_crs, int n_pts, const double* __restrict__ geo,
                                             const double* __restrict__ jdet, const double* __restrict__ pts,
                                             const double* __restrict__ wts, const double* __restrict__ basis,
                                             int n_waves, const double* __restrict__ waves, double time,
                                             const T* __restrict__ dofs, double* __restrict__ part) {
  __shared__ double s1[256], s2[256], si[256];
  const int nqr = 9 * n_crs, lanes = blockDim.x / nqr;
  const int tid = threadIdx.x, qr = tid % nqr, sub = tid / nqr, q = qr / n_crs, r = qr % n_crs;
  const int64_t per = (n_el + gridDim.x - 1) / gridDim.x;
  const int64_t e0 = blockIdx.x * per, e1 = min(n_el, e0 + per);
  double l1 = 0, l2 = 0, li = 0;
  if (sub < lanes) {
    for (int64_t el = e0 + sub; el < e1; el += lanes) {
      const double* gm = geo + el * 12;
      const double j = fabs(jdet[el]);
      const T* a = dofs + ((size_t)el * 9 + q) * n_md * n_crs + r;
      for (int p = 0; p < n_pts; p++) {
        double num = 0;
        for (int n = 0; n < n_md; n++) num += basis[p * n_md + n] * (double)a[(size_t)n * n_crs];
        double x[3];
        for (int d = 0; d < 3; d++)
          x[d] = gm[d] + pts[3 * p] * gm[3 + d] + pts[3 * p + 1] * gm[6 + d] + pts[3 * p + 2] * gm[9 + d];
        double ref = 0;
        for (int w = 0; w < n_waves; w++) {
          const double* wv = waves + ((size_t)r * n_waves + w) * 14;
          ref += wv[q] * sin(wv[9] * x[0] + wv[10] * x[1] + wv[11] * x[2] - wv[12] * time + wv[13]);
        }
        const double d = fabs(num - ref), wj = wts[p] * j;
        l1 += wj * d;
        l2 += wj * d * d;
        li = fmax(li, d);
      }
    }
  }
  s1[tid] = l1;
  s2[tid] = l2;
  si[tid] = li;
  __syncthreads();
  if (tid < nqr) {
    double a1 = 0, a2 = 0, ai = 0;
    for (int k = 0; k < lanes; k++) {
      a1 += s1[k * nqr + tid];
      a2 += s2[k * nqr + tid];
      ai = fmax(ai, si[k * nqr + tid]);
    }
    part[((size_t)blockIdx.x * 3 + 0) * nqr + tid] = a1;
    part[((size_t)blockIdx.x * 3 + 1) * nqr + tid] = a2;
    part[((size_t)blockIdx.x * 3 + 2) * nqr + tid] = ai;
  }
}
}  // namespace

extern "C" int edge_el_error_norms(edge_el_solver* s, const void* dofs, int32_t n1d, int32_t n_waves,
                                   const double* waves, double time, double* l1, double* l2sq, double* linf,
                                   void* stream) {
  if (!s) return edge::fail(EDGE_ERR_ARG, "solver handle missing");
  if (!s->jdet || !s->geo) return edge::fail(EDGE_ERR_UNSUPPORTED, "error norms need a solver built from geometry");
  if (!dofs || n_waves < 0 || (n_waves > 0 && !waves) || n1d < 0 || n1d > 32)
    return edge::fail(EDGE_ERR_ARG, "error-norm arguments");
  const int R = s->cfg.n_runs, nqr = 9 * R, N = s->n_modes;
  if (nqr > 256) return edge::fail(EDGE_ERR_UNSUPPORTED, "error norms: at most 28 runs");
  const int n1 = n1d ? n1d : s->cfg.order + 1, np = n1 * n1 * n1;
  std::vector<double> pts(3 * np), w(np), basis((size_t)np * N);
  int rc = edge_dg_quad_tet(n1, pts.data(), w.data());
  if (rc == EDGE_OK) rc = edge_dg_eval_basis(s->cfg.order, np, pts.data(), basis.data());
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  const unsigned nblk = 1024;
  double *d_pts = nullptr, *d_w = nullptr, *d_b = nullptr, *d_wv = nullptr, *d_part = nullptr;
  const size_t nwv = (size_t)R * n_waves * 14;
  auto ok = [](hipError_t e) { return e == hipSuccess; };
  std::vector<double> h((size_t)3 * nqr * nblk);
  if (!ok(hipMalloc(&d_pts, sizeof(double) * 3 * np)) || !ok(hipMalloc(&d_w, sizeof(double) * np)) ||
      !ok(hipMalloc(&d_b, sizeof(double) * np * N)) || !ok(hipMalloc(&d_wv, sizeof(double) * (nwv ? nwv : 1))) ||
      !ok(hipMalloc(&d_part, sizeof(double) * h.size())) ||
      !ok(hipMemcpyAsync(d_pts, pts.data(), sizeof(double) * 3 * np, hipMemcpyHostToDevice, st)) ||
      !ok(hipMemcpyAsync(d_w, w.data(), sizeof(double) * np, hipMemcpyHostToDevice, st)) ||
      !ok(hipMemcpyAsync(d_b, basis.data(), sizeof(double) * np * N, hipMemcpyHostToDevice, st)) ||
      (nwv && !ok(hipMemcpyAsync(d_wv, waves, sizeof(double) * nwv, hipMemcpyHostToDevice, st)))) {
    rc = edge::fail(EDGE_ERR_HIP, "error norms: buffers");
  } else {
    if (s->cfg.precision == 64)
      hipLaunchKernelGGL(k_err<double>, dim3(nblk), dim3(256), 0, st, s->n_el, N, R, np, s->geo, s->jdet, d_pts, d_w,
                         d_b, n_waves, d_wv, time, (const double*)dofs, d_part);
    else
      hipLaunchKernelGGL(k_err<float>, dim3(nblk), dim3(256), 0, st, s->n_el, N, R, np, s->geo, s->jdet, d_pts, d_w,
                         d_b, n_waves, d_wv, time, (const float*)dofs, d_part);
    if (!ok(hipGetLastError()) ||
        !ok(hipMemcpyAsync(h.data(), d_part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st)) ||
        !ok(hipStreamSynchronize(st)))
      rc = edge::fail(EDGE_ERR_HIP, "error-norm reduction");
  }
  (void)hipFree(d_pts);
  (void)hipFree(d_w);
  (void)hipFree(d_b);
  (void)hipFree(d_wv);
  (void)hipFree(d_part);
  if (rc) return rc;
  for (int i = 0; i < nqr; i++) {
    double a1 = 0, a2 = 0, ai = 0;
    for (unsigned b = 0; b < nblk; b++) {
      a1 += h[((size_t)b * 3 + 0) * nqr + i];
      a2 += h[((size_t)b * 3 + 1) * nqr + i];
      ai = std::max(ai, h[((size_t)b * 3 + 2) * nqr + i]);
    }
    if (l1) l1[i] = a1;
    if (l2sq) l2sq[i] = a2;
    if (linf) linf[i] = ai;
  }
  return EDGE_OK;
}
