// FP64 L2 norms that allocate nothing: edge_el_l2_norms with its per-block
// partial sums in a file-scope device array instead of a per-call hipMalloc,
// and the block sums formed on the device in the same order as the host sum
// of edge_el_l2_norms (so both return the same bits).
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops).  The array lives on every GPU the library
// runs on until the process ends; calls on one GPU share it, so calls that
// overlap on one GPU (other streams, other threads) race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "common.hpp"
#include "kernels.hpp"

namespace {

constexpr int kMaxRuns = 16;  // the largest run count a solver takes
constexpr size_t kPartDoubles = (size_t)edge::kL2Blocks * 2 * 9 * kMaxRuns;
constexpr size_t kScratchDoubles = kPartDoubles + 2 * 9 * kMaxRuns;  // 2.25 MiB + the sums

// part[kL2Blocks][2][9 R], then sums[2][9 R]
__device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};  // device address of `scratch` per device

// This device's address of `scratch` (looked up once per device, with that
// device current).
int scratch_ptr(double** out) {
  int dev = -1;
  EDGE_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return edge::fail(EDGE_ERR_UNSUPPORTED, "device index out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev]) {
    double* p = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return edge::fail(EDGE_ERR_HIP, "hipGetSymbolAddress(scratch) failed");
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return EDGE_OK;
}

// sums[i] = sum_k part[k][i] for i < 2 nqr, k in block order (the host loop
// of edge_el_l2_norms: plain additions from 0, k ascending)
__global__ __launch_bounds__(256) void k_l2_sum(const double* __restrict__ part, unsigned nblk, int nqr,
                                                double* __restrict__ sums) {
  for (int i = threadIdx.x; i < 2 * nqr; i += blockDim.x) {
    double acc = 0.0;
    for (unsigned k = 0; k < nblk; k++) acc += part[(size_t)k * 2 * nqr + i];
    sums[i] = acc;
  }
}

}  // namespace

extern "C" int edge_el_l2_norms_noalloc(edge_el_solver* s, const void* a, int prec_a, const void* b, int prec_b,
                                        double* diff_sq, double* ref_sq, void* stream) {
  if (!s) return edge::fail(EDGE_ERR_ARG, "solver handle missing");
  edge::SolverRef r;
  const int rc0 = edge::solver_ref(s, &r);
  if (rc0 != EDGE_OK) return rc0;
  if (!r.jdet) return edge::fail(EDGE_ERR_UNSUPPORTED, "norms need a solver built from geometry");
  if (!a || (prec_a != 32 && prec_a != 64) || (b && prec_b != 32 && prec_b != 64))
    return edge::fail(EDGE_ERR_ARG, "norm arguments");
  if (r.n_runs < 1 || r.n_runs > kMaxRuns)
    return edge::fail(EDGE_ERR_ARG, "run count above the static scratch's maximum (16)");
  double* base = nullptr;
  const int rc1 = scratch_ptr(&base);
  if (rc1 != EDGE_OK) return rc1;
  hipStream_t st = (hipStream_t)stream;
  const int nqr = 9 * r.n_runs;
  const unsigned nblk = edge::kL2Blocks;
  double* part = base;
  double* sums = base + kPartDoubles;
  // every partial of the nblk blocks is written (empty ranges write zeros),
  // so nothing from an earlier call is read
  const int rc = edge::l2_partials(s, a, prec_a, b, prec_b, part, nblk, st);
  if (rc != EDGE_OK) return rc;
  hipLaunchKernelGGL(k_l2_sum, dim3(1), dim3(256), 0, st, (const double*)part, nblk, nqr, sums);
  EDGE_HIP_TRY(hipGetLastError());
  if (diff_sq) EDGE_HIP_TRY(hipMemcpyAsync(diff_sq, sums, sizeof(double) * nqr, hipMemcpyDeviceToHost, st));
  if (ref_sq) EDGE_HIP_TRY(hipMemcpyAsync(ref_sq, sums + nqr, sizeof(double) * nqr, hipMemcpyDeviceToHost, st));
  EDGE_HIP_TRY(hipStreamSynchronize(st));
  return EDGE_OK;
}
