// exports.cpp -- extern "C" entry points phist_D* (double) and phist_S*
// (float) of libphist_kernels_amd.so, instantiated from exports_def.inc.
#include "objects.hpp"
#include <cmath>

using namespace phx;

namespace {
// alpha == 0 or an empty matrix only scales y (crsmat_module.F90:1508-1511)
template <typename T>
inline bool spmv_trivial(const SparseMat<T>* A, T alpha) {
  return alpha == T(0) || A->nnz == 0;
}
}  // namespace

#define ST double
#define P(name) phist_D##name
#define H(name) phist_D##name
#include "exports_def.inc"
#undef ST
#undef P
#undef H

#define ST float
#define P(name) phist_S##name
#define H(name) phist_S##name
#include "exports_def.inc"
#undef ST
#undef P
#undef H
