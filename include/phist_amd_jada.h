/* phist_amd_jada.h -- eigensolver and shifted linear solver entry points of
 * libphist_kernels_amd.so, with PHIST's own option and operator structures.
 *
 * The structures below have the member order and types of the reference's
 * definitions so that callers compiled against PHIST's headers can pass
 * theirs unchanged:
 *   phist_jadaOpts        src/jada/phist_jadaOpts.h:33-199
 *   enums                 src/tools/phist_enums.h:25-109
 *   phist_[DS]linearOp    src/core/phist_operator_decl.h:14-74
 */
#ifndef PHIST_AMD_JADA_H
#define PHIST_AMD_JADA_H

#include "phist_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum phist_EeigSort {
  phist_NO_EIGSORT = 0,
  phist_LM = 1,
  phist_SM = 2,
  phist_LR = 3,
  phist_SR = 4,
  phist_TARGET = 5,
  phist_INVALID_EIGSORT
} phist_EeigSort;

typedef enum phist_EeigExtr { phist_STANDARD = 0, phist_HARMONIC = 1, phist_INVALID_EIGEXTR } phist_EeigExtr;

typedef enum phist_ElinSolv {
  phist_NO_LINSOLV = 0,
  phist_GMRES = 1,
  phist_MINRES = 2,
  phist_QMR = 3,
  phist_BICGSTAB = 4,
  phist_CARP_CG = 5,
  phist_PCG = 6,
  phist_IDRS = 7,
  phist_USER_LINSOLV = 98,
  phist_INVALID_LINSOLV
} phist_ElinSolv;

typedef enum phist_EmatSym {
  phist_GENERAL = 0,
  phist_HERMITIAN = 1,
  phist_COMPLEX_SYMMETRIC = 2,
  phist_PATTERN_SYMMETRIC = 3,
  phist_INVALID_MATSYM
} phist_EmatSym;

typedef enum phist_Eprecon {
  phist_NO_PRECON = 0,
  phist_IFPACK,
  phist_ML,
  phist_MUELU,
  phist_AMESOS2,
  phist_USER_PRECON = 98,
  phist_INVALID_PRECON
} phist_Eprecon;

/* src/jada/phist_jadaOpts.h:33-199 */
typedef struct phist_jadaOpts {
  int numEigs;
  phist_EeigSort which;
  double convTol;
  int relConvTol;
  phist_EmatSym symmetry;
  phist_EeigExtr how;
  int maxIters;
  int blockSize;
  int lookAhead;
  int minBas;
  int maxBas;
  void* v0;
  int arno;
  double initialShift_r;
  double initialShift_i;
  int initialShiftIters;
  phist_ElinSolv innerSolvType;
  int innerSolvBlockSize;
  int innerSolvMaxBas;
  int innerSolvMaxIters;
  double innerSolvBaseTol;
  int innerSolvRobust;
  int innerSolvStopAfterFirstConverged;
  int innerSolvMaxProjectionSpace;
  void* preconOp;
  phist_Eprecon preconType;
  char preconOpts[1024];
  int preconSkewProject;
  int preconUpdate;
  void* customSolver;
  void (*customSolver_run)(void* customSolverData, void const* A_op, void const* B_op, void const* Qtil,
                           void const* BQtil, const double sigma_r[], const double sigma_i[], void const* res,
                           const int resIndex[], const double tol[], int maxIter, void* t, int robust,
                           int abortAfterFirstConvergedInBlock, int* iflag);
  void (*customSolver_run1)(void* customSolverData, void const* A_op, void const* B_op, void const* Qtil,
                            void const* BQtil, const double sigma, double sigma_i, void const* res, double tol,
                            int maxIter, void* t, int robust, int* iflag);
} phist_jadaOpts;

/* src/jada/phist_jadaOpts.cpp:19-58 (the reference's default values) */
void phist_jadaOpts_setDefaults(phist_jadaOpts* opts);

/* src/core/phist_operator_decl.h:14-74 */
struct phist_DlinearOp {
  void const* A;
  void* aux;
  phist_const_map_ptr range_map;
  phist_const_map_ptr domain_map;
  int use_transpose;
  double* shifts;
  void (*apply)(double alpha, const void* A, phist_Dconst_mvec_ptr X, double beta, phist_Dmvec_ptr Y, int* iflag);
  void (*applyT)(double alpha, const void* A, phist_Dconst_mvec_ptr X, double beta, phist_Dmvec_ptr Y, int* iflag);
  void (*apply_shifted)(double alpha, const void* A, double const* sigma, phist_Dconst_mvec_ptr X, double beta,
                        phist_Dmvec_ptr Y, int* iflag);
  void (*fused_apply_mvTmv)(double alpha, const void* A, phist_Dconst_mvec_ptr V, double beta, phist_Dmvec_ptr W,
                            phist_DsdMat_ptr WtW, phist_DsdMat_ptr VtW, int* iflag);
  void (*update)(const void* A, void* aux, double sigma, phist_Dconst_mvec_ptr Vkern, phist_Dconst_mvec_ptr BVkern,
                 int* iflag);
  void (*destroy)(struct phist_DlinearOp* me, int* iflag);
};
typedef phist_DlinearOp* phist_DlinearOp_ptr;

struct phist_SlinearOp {
  void const* A;
  void* aux;
  phist_const_map_ptr range_map;
  phist_const_map_ptr domain_map;
  int use_transpose;
  float* shifts;
  void (*apply)(float alpha, const void* A, phist_Sconst_mvec_ptr X, float beta, phist_Smvec_ptr Y, int* iflag);
  void (*applyT)(float alpha, const void* A, phist_Sconst_mvec_ptr X, float beta, phist_Smvec_ptr Y, int* iflag);
  void (*apply_shifted)(float alpha, const void* A, float const* sigma, phist_Sconst_mvec_ptr X, float beta,
                        phist_Smvec_ptr Y, int* iflag);
  void (*fused_apply_mvTmv)(float alpha, const void* A, phist_Sconst_mvec_ptr V, float beta, phist_Smvec_ptr W,
                            phist_SsdMat_ptr WtW, phist_SsdMat_ptr VtW, int* iflag);
  void (*update)(const void* A, void* aux, float sigma, phist_Sconst_mvec_ptr Vkern, phist_Sconst_mvec_ptr BVkern,
                 int* iflag);
  void (*destroy)(struct phist_SlinearOp* me, int* iflag);
};
typedef phist_SlinearOp* phist_SlinearOp_ptr;

/* src/core/phist_operator_decl.h:78 (apply, apply_shifted with +sigma_j as
   phist_operator_def.hpp:84-106, fused_apply_mvTmv; applyT returns -99) */
void phist_DlinearOp_wrap_sparseMat(phist_DlinearOp_ptr op, phist_Dconst_sparseMat_ptr A, int* iflag);
void phist_SlinearOp_wrap_sparseMat(phist_SlinearOp_ptr op, phist_Sconst_sparseMat_ptr A, int* iflag);

/* Block Jacobi-Davidson QR for a few exterior eigenpairs (subspacejada,
   src/jada/phist_subspacejada_def.hpp:69-1097): partial Schur form A*Q = Q*R.
   Implemented for opts.symmetry == phist_HERMITIAN (or COMPLEX_SYMMETRIC in
   real arithmetic), B_op == NULL, opts.how == phist_STANDARD, which in
   {LM, SM, LR, SR}; the correction equations are solved with blocked MINRES
   (innerSolvType MINRES or GMRES) or skipped (NO_LINSOLV); other settings
   return -99.  ev holds numEigs+blockSize-1 values, resNorm numEigs. */
void phist_Dsubspacejada(phist_Dconst_linearOp_ptr AB_op, phist_Dconst_linearOp_ptr B_op, phist_jadaOpts opts,
                         phist_Dmvec_ptr Q, phist_DsdMat_ptr R, double* ev, double* resNorm, int* nConv, int* nIter,
                         int* iflag);
void phist_Ssubspacejada(phist_Sconst_linearOp_ptr AB_op, phist_Sconst_linearOp_ptr B_op, phist_jadaOpts opts,
                         phist_Smvec_ptr Q, phist_SsdMat_ptr R, float* ev, float* resNorm, int* nConv, int* nIter,
                         int* iflag);

/* ---- extensions (no reference counterpart) ---- */

/* Precision of the correction-equation solves of phist_Dsubspacejada when the
   operator wraps a sparse matrix of this library: 0 (default) FP64, 1 FP32
   (an FP32 copy of the matrix is made once and kept with the matrix; the
   outer iteration, orthogonalisation and residuals stay FP64).  The
   environment variable PHIST_AMD_JADA_INNER_FP32 sets the initial value. */
void phist_amd_jada_set_inner_fp32(int enable, int* iflag);

/* Shifted systems (A + shifts[j] I) x_j = b_j, j < nsys <= 64, solved
   together by blocked MINRES (A symmetric), starting from x = 0, until
   ||b_j - (A + shifts[j]) x_j|| <= tol[j] ||b_j|| (the MINRES estimate) or
   maxIter iterations; converged systems stop updating.  With fp32 != 0 the
   iteration runs in FP32 on a cached FP32 copy of A (B and X stay FP64) and
   is refined in FP64: between FP32 solves the true residuals are formed in
   FP64 and the systems still above tol solve for their correction again
   (at most maxIter iterations in all); relres[j] is then the true FP64
   residual, nIter[j] the iterations of all rounds.  In FP64, nIter[j] and
   relres[j] (the MINRES estimate) are returned per system; iflag > 0 is the
   number of systems that did not converge. */
void phist_amd_DblockedMINRES_shifted(phist_Dconst_sparseMat_ptr A, int nsys, const double shifts[],
                                      phist_Dconst_mvec_ptr B, phist_Dmvec_ptr X, const double tol[], int maxIter,
                                      int fp32, int nIter[], double relres[], int* iflag);

/* Any number nsys of shifted systems (A + shifts[j] I) x_j = b_j (A symmetric; B and X hold nsys
   columns) solved by blocked MINRES in a block of `slots` columns (1..64): when a system converges or
   reaches maxIter its solution is stored in X(:,j) and its slot takes the next system while the
   others keep iterating -- the replace-converged loop of
   src/jada/phist_jadaCorrectionSolver_def.hpp:428-466 (there over the correction equations; here over
   plain shifted solves).  Same stopping rule, precision switch and outputs as
   phist_amd_DblockedMINRES_shifted; iflag > 0 = number of unconverged systems. */
void phist_amd_DblockedMINRES_shifted_refill(phist_Dconst_sparseMat_ptr A, int nsys, const double shifts[],
                                             phist_Dconst_mvec_ptr B, phist_Dmvec_ptr X, const double tol[],
                                             int maxIter, int fp32, int slots, int nIter[], double relres[],
                                             int* iflag);

/* Restarted GMRES(restart) on nsys <= 64 shifted systems (A + shifts[j] I) x_j = b_j of a general
   (nonsymmetric) matrix of this library, one Krylov basis per system as in
   src/krylov/phist_blockedgmres_def.hpp:453-1010 (restartedBlockedGMRES :1029-1200); X holds the initial
   guess.  Converged when ||r|| / ||r0|| < tol[j] or ||r|| < 1000 eps; maxIter bounds the Arnoldi steps
   per system.  nIter[j] = steps taken, relres[j] = final true residual over the initial one; fp32 != 0
   iterates in FP32 on the matrix's cached FP32 copy.  iflag = number of unconverged systems. */
void phist_amd_DblockedGMRES_shifted(phist_Dconst_sparseMat_ptr A, int nsys, const double shifts[],
                                     phist_Dconst_mvec_ptr B, phist_Dmvec_ptr X, const double tol[], int maxIter,
                                     int restart, int fp32, int nIter[], double relres[], int* iflag);

#ifdef __cplusplus
}
#endif

#endif
