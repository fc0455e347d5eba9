/*
 * phist_amd.h -- C ABI of the MI355X kernel library for PHIST
 * (libphist_kernels_amd.so).
 *
 * The library is a PHIST kernel backend selectable next to `builtin`: it
 * exports the symbols PHIST's solvers link against, with the reference's
 * names, argument meaning, ownership rules and `int* iflag` error protocol.
 * Data live in HBM: mvecs are device-resident row-major blocks (element
 * (i,j) at v[i*lda+j]), sparse matrices are stored as SELL-C-sigma on the
 * device, sdMats are column-major and mirrored on host and device.
 *
 * Reference interface files (relative to the PHIST source tree):
 *   src/kernels/phist_kernels.h          type-independent functions
 *   src/kernels/phist_kernels_decl.h     sparseMat / mvec / sdMat functions
 *   src/kernels/phist_kernels_fused_decl.h, _common_decl.h, _carp_decl.h
 *   src/kernels/phist_kernel_flags.h     iflag input flags
 *   src/tools/phist_defs.h:52-74         return codes
 *   src/kernels/builtin/phist_typedefs.h index types
 *   src/kernels/phist_void_aliases*.h    opaque handle types
 *   src/core/phist_orthog_decl.h, phist_chol_QR_decl.h, phist_svqb_decl.h,
 *   src/core/phist_sdFact_decl.h         orthogonalisation entry points
 *
 * Per-type declarations (phist_D* for double, phist_S* for float) are in
 * phist_amd_decl.h, included twice below.
 */
#ifndef PHIST_AMD_H
#define PHIST_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- index types (builtin/phist_typedefs.h:28-31) ---- */
typedef int32_t phist_lidx;
typedef int64_t phist_gidx;

/* ---- return codes (src/tools/phist_defs.h:52-74) ---- */
#define PHIST_SUCCESS            0
#define PHIST_WARNING            1
#define PHIST_FUNCTIONAL_ERROR  -1
#define PHIST_MPI_ERROR        -33
#define PHIST_MEM_ALLOC_FAILED -44
#define PHIST_INVALID_INPUT    -55
#define PHIST_INTEGER_OVERFLOW -66
#define PHIST_CAUGHT_EXCEPTION -77
#define PHIST_BAD_CAST         -88
#define PHIST_NOT_IMPLEMENTED  -99
#define PHIST_DEPRECATED        99

/* ---- input flags (src/kernels/phist_kernel_flags.h) ---- */
#define PHIST_IFLAG_DEFAULT                  0
#define PHIST_KERNELS_QUIET                 64
#define PHIST_SPARSEMAT_PERM_LOCAL           1
#define PHIST_SPARSEMAT_PERM_GLOBAL          2
#define PHIST_SPARSEMAT_DIST2_COLOR          4
#define PHIST_SPARSEMAT_OPT_SINGLESPMVM      8
#define PHIST_SPARSEMAT_OPT_BLOCKSPMVM      16
#define PHIST_SPARSEMAT_OPT_CARP            32
#define PHIST_SPARSEMAT_QUIET               64
#define PHIST_SPARSEMAT_OWN_MAPS           128
#define PHIST_SPARSEMAT_OVERLAP_COMMUNICATION 256
#define PHIST_MVEC_REPLICATE_DEVICE_MEM      1
#define PHIST_ROBUST_REDUCTIONS              1
#define PHIST_NO_GLOBAL_REDUCTION            2
#define PHIST_SPMVM_ONLY_LOCAL            1024
#define PHIST_SPMVM_VECTOR                2048
#define PHIST_SPMVM_OVERLAP               4096
#define PHIST_SPMVM_TASK                  8192
#define PHIST_SDMAT_RUN_ON_HOST           1024
#define PHIST_SDMAT_RUN_ON_DEVICE         2048
#define PHIST_SDMAT_RUN_ON_HOST_AND_DEVICE (PHIST_SDMAT_RUN_ON_HOST|PHIST_SDMAT_RUN_ON_DEVICE)
#define PHIST_MVEC_RUN_ON_DEVICE PHIST_SDMAT_RUN_ON_DEVICE

/* ---- opaque handles (phist_void_aliases.h:17-35) ---- */
typedef void* phist_comm_ptr;
typedef const void* phist_const_comm_ptr;
typedef void* phist_map_ptr;
typedef const void* phist_const_map_ptr;
typedef void* phist_context_ptr;
typedef const void* phist_const_context_ptr;

/* This backend uses no MPI; like PHIST's MPI-free backends it aliases the
 * communicator type to int (src/kernels/phist_mpi_kernels.h). */
typedef int MPI_Comm;

/* per-type opaque objects (phist_void_aliases_decl.h:14-35) */
typedef struct phist_Dmvec_s phist_Dmvec;
typedef phist_Dmvec* phist_Dmvec_ptr;
typedef const phist_Dmvec* phist_Dconst_mvec_ptr;
typedef struct phist_DsdMat_s phist_DsdMat;
typedef phist_DsdMat* phist_DsdMat_ptr;
typedef const phist_DsdMat* phist_Dconst_sdMat_ptr;
typedef struct phist_DsparseMat_s phist_DsparseMat;
typedef phist_DsparseMat* phist_DsparseMat_ptr;
typedef const phist_DsparseMat* phist_Dconst_sparseMat_ptr;
typedef struct phist_DlinearOp phist_DlinearOp; /* defined in phist_amd_jada.h */
typedef const phist_DlinearOp* phist_Dconst_linearOp_ptr;

typedef struct phist_Smvec_s phist_Smvec;
typedef phist_Smvec* phist_Smvec_ptr;
typedef const phist_Smvec* phist_Sconst_mvec_ptr;
typedef struct phist_SsdMat_s phist_SsdMat;
typedef phist_SsdMat* phist_SsdMat_ptr;
typedef const phist_SsdMat* phist_Sconst_sdMat_ptr;
typedef struct phist_SsparseMat_s phist_SsparseMat;
typedef phist_SsparseMat* phist_SsparseMat_ptr;
typedef const phist_SsparseMat* phist_Sconst_sparseMat_ptr;
typedef struct phist_SlinearOp phist_SlinearOp; /* defined in phist_amd_jada.h */
typedef const phist_SlinearOp* phist_Sconst_linearOp_ptr;

/* row and element callbacks (phist_kernels.h:197-202) */
typedef int (*phist_sparseMat_rowFunc)(phist_gidx row, phist_lidx* nnz, phist_gidx* cols, void* vals, void* arg);
typedef int (*phist_sparseMat_rowFuncConstructor)(void* arg, void** work);
typedef int (*phist_mvec_elemFunc)(phist_gidx row, phist_lidx col, void* val, void* arg);

/* ---- type-independent functions (phist_kernels.h:79-185) ---- */

/* phist_kernels.h:79. Selects the HIP device (HIP_VISIBLE_DEVICES / current
 * device), creates the library's stream and seeds the global KISS state. */
void phist_kernels_init(int* argc, char*** argv, int* iflag);
/* phist_kernels.h:84 */
void phist_kernels_finalize(int* iflag);
/* phist_kernels.h:87-99: a single-process communicator (rank 0 of 1). */
void phist_comm_create(phist_comm_ptr* comm, int* iflag);
void phist_comm_delete(phist_comm_ptr comm, int* iflag);
void phist_comm_get_rank(phist_const_comm_ptr comm, int* rank, int* iflag);
void phist_comm_get_size(phist_const_comm_ptr comm, int* size, int* iflag);
void phist_comm_get_mpi_comm(phist_const_comm_ptr comm, MPI_Comm* mpiComm, int* iflag);
/* phist_kernels.h:102-132: contiguous row maps. */
void phist_map_create(phist_map_ptr* map, phist_const_comm_ptr comm, phist_gidx nglob, int* iflag);
void phist_map_delete(phist_map_ptr map, int* iflag);
void phist_map_get_comm(phist_const_map_ptr map, phist_const_comm_ptr* comm, int* iflag);
void phist_map_get_local_length(phist_const_map_ptr map, phist_lidx* nloc, int* iflag);
void phist_map_get_global_length(phist_const_map_ptr map, phist_gidx* nglob, int* iflag);
void phist_map_get_ilower(phist_const_map_ptr map, phist_gidx* ilower, int* iflag);
void phist_map_get_iupper(phist_const_map_ptr map, phist_gidx* iupper, int* iflag);
void phist_maps_compatible(phist_const_map_ptr map1, phist_const_map_ptr map2, int* iflag);
/* phist_kernels.h:154-162 */
void phist_context_create(phist_context_ptr* ctx, phist_const_map_ptr row_map, phist_const_map_ptr col_map,
                          phist_const_map_ptr range_map, phist_const_map_ptr domain_map, int* iflag);
void phist_context_delete(phist_context_ptr vctx, int* iflag);
/* phist_kernels.h:172-178: STREAM benchmarks, run on the GPU's HBM (GB/s). */
void phist_bench_stream_load(double* mean_bw, double* max_bw, int* iflag);
void phist_bench_stream_store(double* mean_bw, double* max_bw, int* iflag);
void phist_bench_stream_copy(double* mean_bw, double* max_bw, int* iflag);
void phist_bench_stream_triad(double* mean_bw, double* max_bw, int* iflag);

#ifdef __cplusplus
}
#endif

/* ---- per-type functions ---- */
#define PHIST_AMD_ST double
#define PHIST_AMD_MT double
#define PHIST_AMD_SUBR(name) phist_D##name
#define PHIST_AMD_TYPE(name) phist_D##name
#include "phist_amd_decl.h"
#undef PHIST_AMD_ST
#undef PHIST_AMD_MT
#undef PHIST_AMD_SUBR
#undef PHIST_AMD_TYPE

#define PHIST_AMD_ST float
#define PHIST_AMD_MT float
#define PHIST_AMD_SUBR(name) phist_S##name
#define PHIST_AMD_TYPE(name) phist_S##name
#include "phist_amd_decl.h"
#undef PHIST_AMD_ST
#undef PHIST_AMD_MT
#undef PHIST_AMD_SUBR
#undef PHIST_AMD_TYPE

#endif /* PHIST_AMD_H */
