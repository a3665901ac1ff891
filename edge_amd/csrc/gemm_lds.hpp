// Small GEMMs on LDS tiles for the generic (one workgroup per element)
// kernels: threads of a kThreads-wide workgroup spread over the outputs.
// Row-major with the run axis innermost, as src/linalg/Matrix.h:296-380.
#pragma once

#include <hip/hip_runtime.h>

namespace edge {

constexpr int kThreads = 256;

// C[m][n][r] (+)= sum_k A[m][k][r] B[k][n]  (A, C in LDS; B row-major, ld = ldb)
template <typename T, int R>
__device__ inline void gemm_ac(int m, int n, int k, int lda, int ldb, int ldc, bool beta, const T* A, const T* B,
                               T* C) {
  for (int i = threadIdx.x; i < m * n * R; i += kThreads) {
    const int r = i % R, in = (i / R) % n, im = i / (R * n);
    T acc = beta ? C[(im * ldc + in) * R + r] : T(0);
    for (int ik = 0; ik < k; ik++) acc += A[(im * lda + ik) * R + r] * B[ik * ldb + in];
    C[(im * ldc + in) * R + r] = acc;
  }
}

// C[m][n][r] (+)= sum_k A[m][k] B[k][n][r]
template <typename T, int R>
__device__ inline void gemm_bc(int m, int n, int k, int lda, int ldb, int ldc, bool beta, const T* A, const T* B,
                               T* C) {
  for (int i = threadIdx.x; i < m * n * R; i += kThreads) {
    const int r = i % R, in = (i / R) % n, im = i / (R * n);
    T acc = beta ? C[(im * ldc + in) * R + r] : T(0);
    for (int ik = 0; ik < k; ik++) acc += A[im * lda + ik] * B[(ik * ldb + in) * R + r];
    C[(im * ldc + in) * R + r] = acc;
  }
}

}  // namespace edge
