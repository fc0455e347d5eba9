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

// colour classes of a CARP work object made by phist_[DS]carp_setup(_rc)
extern "C" void phist_amd_Dcarp_colors(void* work, int* ncolors, int64_t* color_ptr, int32_t* rows, int* iflag) {
  guard(iflag, "phist_amd_Dcarp_colors", [&] {
    carp_colors<double>(work, ncolors, color_ptr, rows);
    *iflag = 0;
  });
}
extern "C" void phist_amd_Scarp_colors(void* work, int* ncolors, int64_t* color_ptr, int32_t* rows, int* iflag) {
  guard(iflag, "phist_amd_Scarp_colors", [&] {
    carp_colors<float>(work, ncolors, color_ptr, rows);
    *iflag = 0;
  });
}
