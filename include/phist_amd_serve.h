/* phist_amd_serve.h -- a continuous-batching service loop for shifted solves
 * and eigenvalue requests on one GPU (no counterpart in the reference).
 *
 * Shifted systems (A + s_j I) x_j = b_j arriving in separate requests are
 * solved together: one blocked MINRES batch per (matrix, precision) with a
 * fixed number of slots; when a system converges its slot is refilled with
 * the next queued system while the others keep iterating (the slot swap of
 * src/jada/phist_jadaCorrectionSolver_def.hpp:428-466, applied across
 * requests).  Eigenvalue requests (symmetric block Jacobi-Davidson) advance
 * one outer iteration per step, interleaved with the solves.  Work is
 * admitted first-come first-served while its device work space fits a
 * byte budget.  Single host thread; the caller drives the loop with
 * phist_amd_server_step and keeps every buffer of a request alive until
 * the request is done.
 */
#ifndef PHIST_AMD_SERVE_H
#define PHIST_AMD_SERVE_H

#include "phist_amd_jada.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct phist_amd_server_s* phist_amd_server_ptr;

/* request states reported by phist_amd_server_status */
#define PHIST_AMD_REQ_QUEUED 0
#define PHIST_AMD_REQ_RUNNING 1
#define PHIST_AMD_REQ_DONE 2
#define PHIST_AMD_REQ_DONE_UNCONVERGED 3

/* hbm_budget_bytes bounds the work space of admitted requests (solver
   blocks, not the caller's matrices and vectors); slots = systems per MINRES
   batch (1..64). */
void phist_amd_server_create(phist_amd_server_ptr* srv, int64_t hbm_budget_bytes, int slots, int* iflag);
void phist_amd_server_delete(phist_amd_server_ptr srv, int* iflag);

/* nrhs shifted systems with right-hand sides B(:,j) and solutions X(:,j)
   (x starts at 0); tolerance relative to ||b_j||; fp32 != 0 iterates in FP32
   on the matrix's cached FP32 copy. */
void phist_amd_server_submit_solve(phist_amd_server_ptr srv, phist_Dconst_sparseMat_ptr A, int nrhs,
                                   const double shifts[], phist_Dconst_mvec_ptr B, phist_Dmvec_ptr X, double tol,
                                   int maxIter, int fp32, int64_t* id, int* iflag);

/* phist_Dsubspacejada on a matrix of this library (same restrictions);
   Q, R, ev, resNorm as there; fp32 != 0 solves the correction equations in
   FP32. The options are copied. */
void phist_amd_server_submit_jada(phist_amd_server_ptr srv, phist_Dconst_sparseMat_ptr A, const phist_jadaOpts* opts,
                                  int fp32, phist_Dmvec_ptr Q, phist_DsdMat_ptr R, double* ev, double* resNorm,
                                  int64_t* id, int* iflag);

/* admit queued work, then one MINRES iteration of every batch and one outer
   iteration of every running eigenvalue request; pending = requests not yet
   done */
void phist_amd_server_step(phist_amd_server_ptr srv, int* pending, int* iflag);

/* state of a request and, once done, its latency (seconds from submission to
   completion) and iterations (MINRES: the largest over its systems; JD:
   outer iterations) */
void phist_amd_server_status(phist_amd_server_ptr srv, int64_t id, int* state, double* latency_s, int* iters,
                             int* iflag);

/* forget a finished request (its buffers may then be reused) */
void phist_amd_server_release(phist_amd_server_ptr srv, int64_t id, int* iflag);

#ifdef __cplusplus
}
#endif

#endif
