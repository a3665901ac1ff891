// Internal helpers shared by the library's translation units.
#pragma once

#include <cstdint>
#include <string>

#include "../../include/edge_amd.h"

namespace edge {

// Records an error message for edge_last_error() and returns the code.
int fail(int code, const std::string& msg);

constexpr int n_modes(int order) { return order * (order + 1) * (order + 2) / 6; }
constexpr int n_face_modes(int order) { return order * (order + 1) / 2; }
// sub-faces of a fault face: (2 O - 1)^2 (src/sc/cexprs.inc:51-64)
constexpr int n_sub_faces(int order) { return (2 * order - 1) * (2 * order - 1); }

constexpr int kMinOrder = 2;
constexpr int kMaxOrder = 5;

// Reference-element matrices of one order in FP64 (host tables).
struct DgTables {
  int order, n, nf, nsf;
  const double* mass;
  const double* stiffT;
  const double* stiffV;
  const double* fluxL;
  const double* fluxN;
  const double* fluxT;
  const double* massF;     // [NF] diagonal face mass
  const double* sfMean;    // [NSF][NF] means of the face basis over the fault sub-faces
  const double* sfGather;  // [NF][NSF] least-squares inverse of sfMean
};
bool dg_tables(int order, DgTables* out);

}  // namespace edge

#define EDGE_HIP_TRY(expr)                                                                   \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return ::edge::fail(EDGE_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)
