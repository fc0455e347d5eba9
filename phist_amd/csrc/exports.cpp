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

// fused_spmv_mvdot that allocates nothing: the dot results stay on the device
// and the reduction partials live in the library's fixed device scratch
extern "C" void phist_Dfused_spmv_mvdot_noalloc(double alpha, phist_Dconst_sparseMat_ptr A, phist_Dconst_mvec_ptr V,
                                                double beta, phist_Dmvec_ptr W, double* WdotW_dev, double* VdotW_dev,
                                                int* iflag) {
  guard(iflag, "phist_Dfused_spmv_mvdot_noalloc", [&] {
    const SparseMat<double>* a = SP<double>(A);
    const Mvec<double>* v = MV<double>(V);
    Mvec<double>* w = MV<double>(W);
    PHX_REQUIRE(v && w, PHIST_INVALID_INPUT, "V and W are required");
    PHX_REQUIRE(v->nvec <= 64, PHIST_INVALID_INPUT, "at most 64 columns (V has %d)", v->nvec);
    PHX_REQUIRE(w->nrows > 0, PHIST_INVALID_INPUT, "empty block");
    NoAllocScope scope;
    if (spmv_trivial(a, alpha)) {
      blas_scale_cols<double>(w, &beta, true);
      if (VdotW_dev) blas_dot_dev<double>(v, w, VdotW_dev);
      if (WdotW_dev) blas_dot_dev<double>(w, w, WdotW_dev);
    } else {
      SpmvRequest rq;
      rq.has_beta = beta != 0.0;
      rq.dot_ww = WdotW_dev != nullptr;
      rq.dot_vw = VdotW_dev != nullptr;
      rq.dot_ww_dev = WdotW_dev;
      rq.dot_vw_dev = VdotW_dev;
      spmv_launch<double>(a, alpha, nullptr, v, beta, w, rq, nullptr, nullptr);
    }
    *iflag = 0;
  });
}
