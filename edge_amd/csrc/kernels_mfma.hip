// MFMA-tiled element kernels.
#include "kernels_mfma.hpp"

namespace edge {

bool mfma_supported(int, int, int) { return false; }

template <typename T>
int launch_local_mfma(const SolverDev<T>&, int64_t, int64_t, T, T*, T*, hipStream_t) {
  return fail(EDGE_ERR_UNSUPPORTED, "MFMA kernels not built");
}
template <typename T>
int launch_neigh_mfma(const SolverDev<T>&, int64_t, int64_t, const T*, T*, hipStream_t) {
  return fail(EDGE_ERR_UNSUPPORTED, "MFMA kernels not built");
}
template int launch_local_mfma<double>(const SolverDev<double>&, int64_t, int64_t, double, double*, double*, hipStream_t);
template int launch_local_mfma<float>(const SolverDev<float>&, int64_t, int64_t, float, float*, float*, hipStream_t);
template int launch_neigh_mfma<double>(const SolverDev<double>&, int64_t, int64_t, const double*, double*, hipStream_t);
template int launch_neigh_mfma<float>(const SolverDev<float>&, int64_t, int64_t, const float*, float*, hipStream_t);

}  // namespace edge
