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
#define PHIST_AMD_REQ_CANCELLED 4
/* a request that failed reports its error code (< 0) as its state, e.g.
   -44 when its work space can never fit the budget */

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

/* phist_Dsubspacejada on a matrix of this library (same restrictions: options
   it does not support fail here with -99); Q, R, ev, resNorm as there;
   fp32 != 0 solves the correction equations in FP32. The options are copied. */
void phist_amd_server_submit_jada(phist_amd_server_ptr srv, phist_Dconst_sparseMat_ptr A, const phist_jadaOpts* opts,
                                  int fp32, phist_Dmvec_ptr Q, phist_DsdMat_ptr R, double* ev, double* resNorm,
                                  int64_t* id, int* iflag);

/* admit queued work, then advance every MINRES batch and every running
   eigenvalue request; pending = requests not yet done.  An eigenvalue request
   advances one outer iteration per step.  A batch advances by at least one
   MINRES iteration per step: by more while eigenvalue requests run (as many
   as fit in the time their outer iterations took in the previous step, at
   most PHIST_AMD_SERVE_MAX_REPS, default 4), and each of those may queue
   several iterations on the device (PHIST_AMD_MINRES_ITERS_PER_CALL; default
   2 for blocks of at most 4M elements, else 1).  The repeat count comes from
   wall-clock timings, so how far a solve advances per step is not
   deterministic; the results do not depend on it (a converged system stops
   updating).  Admission is first come, first served: a request waiting for
   memory holds back the ones behind it, and batches with no system in
   flight are dropped to make room for it. */
void phist_amd_server_step(phist_amd_server_ptr srv, int* pending, int* iflag);

/* state of a request and, once done, its latency (seconds from submission to
   completion) and iterations (MINRES: the largest over its systems; JD:
   outer iterations) */
void phist_amd_server_status(phist_amd_server_ptr srv, int64_t id, int* state, double* latency_s, int* iters,
                             int* iflag);

/* withdraw a queued or running request (state PHIST_AMD_REQ_CANCELLED):
   its systems leave their slots, an eigenvalue request frees its work space;
   X, Q, R, ev, resNorm hold whatever was written so far.  A finished request
   is left unchanged.  No counterpart in the reference. */
void phist_amd_server_cancel(phist_amd_server_ptr srv, int64_t id, int* iflag);

/* forget a finished or cancelled request (its buffers may then be reused) */
void phist_amd_server_release(phist_amd_server_ptr srv, int64_t id, int* iflag);

#ifdef __cplusplus
}
#endif

#endif
