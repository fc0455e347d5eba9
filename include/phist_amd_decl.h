/*
 * phist_amd_decl.h -- per-type declarations of libphist_kernels_amd.so.
 *
 * Included by phist_amd.h once per scalar type with PHIST_AMD_ST (scalar),
 * PHIST_AMD_MT (magnitude), PHIST_AMD_SUBR(name) (symbol) and
 * PHIST_AMD_TYPE(name) (handle type) defined: phist_D* (double) and phist_S*
 * (float).  Each group cites the reference declaration it replaces; names,
 * argument order and meaning, and iflag behaviour are the reference's.
 * No include guard on purpose.
 */
#define ST_ PHIST_AMD_ST
#define MT_ PHIST_AMD_MT
#define SUBR_ PHIST_AMD_SUBR
#define T_ PHIST_AMD_TYPE

#ifdef __cplusplus
extern "C" {
#endif

/* phist_kernels_decl.h:79 -- 0 if the type is available */
void SUBR_(type_avail)(int* iflag);

/* phist_kernels_decl.h:98-118 -- MatrixMarket input ("coordinate real
 * general|symmetric"); HB and binary formats return -99. */
void SUBR_(sparseMat_read_mm)(T_(sparseMat_ptr)* A, phist_const_comm_ptr comm, const char* filename, int* iflag);
void SUBR_(sparseMat_read_hb)(T_(sparseMat_ptr)* A, phist_const_comm_ptr comm, const char* filename, int* iflag);
void SUBR_(sparseMat_read_bin)(T_(sparseMat_ptr)* A, phist_const_comm_ptr comm, const char* filename, int* iflag);
void SUBR_(sparseMat_read_mm_with_context)(T_(sparseMat_ptr)* A, phist_const_context_ptr ctx, const char* filename, int* iflag);
void SUBR_(sparseMat_read_hb_with_context)(T_(sparseMat_ptr)* A, phist_const_context_ptr ctx, const char* filename, int* iflag);
void SUBR_(sparseMat_read_bin_with_context)(T_(sparseMat_ptr)* A, phist_const_context_ptr ctx, const char* filename, int* iflag);

/* phist_kernels_decl.h:127-158 -- maps, context and nnz counts */
void SUBR_(sparseMat_get_row_map)(T_(const_sparseMat_ptr) A, phist_const_map_ptr* map, int* iflag);
void SUBR_(sparseMat_get_col_map)(T_(const_sparseMat_ptr) A, phist_const_map_ptr* map, int* iflag);
void SUBR_(sparseMat_get_domain_map)(T_(const_sparseMat_ptr) A, phist_const_map_ptr* map, int* iflag);
void SUBR_(sparseMat_get_range_map)(T_(const_sparseMat_ptr) A, phist_const_map_ptr* map, int* iflag);
void SUBR_(sparseMat_get_context)(T_(const_sparseMat_ptr) A, phist_const_context_ptr* ctx, int* iflag);
void SUBR_(sparseMat_local_nnz)(T_(const_sparseMat_ptr) A, int64_t* local_nnz, int* iflag);
void SUBR_(sparseMat_global_nnz)(T_(const_sparseMat_ptr) A, int64_t* global_nnz, int* iflag);

/* phist_kernels_decl.h:166-190 -- object lifetime.  mvec_create allocates a
 * device-resident row-major block (HBM) and a lazily created pinned host
 * mirror; sdMat_create allocates host and device copies. */
void SUBR_(mvec_create)(T_(mvec_ptr)* V, phist_const_map_ptr map, int nvec, int* iflag);
void SUBR_(sdMat_create)(T_(sdMat_ptr)* M, int nrows, int ncols, phist_const_comm_ptr comm, int* iflag);
void SUBR_(sparseMat_delete)(T_(sparseMat_ptr) A, int* iflag);
void SUBR_(mvec_delete)(T_(mvec_ptr) V, int* iflag);
void SUBR_(sdMat_delete)(T_(sdMat_ptr) M, int* iflag);

/* phist_kernels_decl.h:196-270 -- queries and raw views.  extract_view
 * returns the host mirror unless *iflag == PHIST_MVEC_RUN_ON_DEVICE on entry,
 * which returns the device pointer; lda is the row stride in elements. */
void SUBR_(mvec_get_map)(T_(const_mvec_ptr) V, phist_const_map_ptr* map, int* iflag);
void SUBR_(mvec_num_vectors)(T_(const_mvec_ptr) V, int* nvec, int* iflag);
void SUBR_(mvec_get_data)(T_(const_mvec_ptr) V, ST_* data, phist_lidx lda, int output_row_major, int* iflag);
void SUBR_(mvec_extract_view)(T_(mvec_ptr) V, ST_** V_raw, phist_lidx* lda, int* iflag);
void SUBR_(mvec_extract_const_view)(T_(const_mvec_ptr) V, ST_ const** V_raw, phist_lidx* lda, int* iflag);
void SUBR_(sdMat_get_nrows)(T_(const_sdMat_ptr) M, int* nrows, int* iflag);
void SUBR_(sdMat_get_ncols)(T_(const_sdMat_ptr) M, int* ncols, int* iflag);
void SUBR_(sdMat_extract_view)(T_(sdMat_ptr) M, ST_** M_raw, phist_lidx* lda, int* iflag);
void SUBR_(sdMat_extract_const_view)(T_(const_sdMat_ptr) M, ST_ const** M_raw, phist_lidx* lda, int* iflag);

/* phist_kernels_decl.h:296-312 -- host/device mirror synchronisation */
void SUBR_(mvec_to_device)(T_(mvec_ptr) V, int* iflag);
void SUBR_(mvec_from_device)(T_(mvec_ptr) V, int* iflag);
void SUBR_(sdMat_to_device)(T_(sdMat_ptr) M, int* iflag);
void SUBR_(sdMat_from_device)(T_(sdMat_ptr) M, int* iflag);

/* phist_kernels_decl.h:324-343 -- raw binary I/O and map conversion */
void SUBR_(mvec_write_bin)(T_(const_mvec_ptr) V, const char* filename, int* iflag);
void SUBR_(mvec_read_bin)(T_(mvec_ptr) V, const char* filename, int* iflag);
void SUBR_(sdMat_write_bin)(T_(const_sdMat_ptr) M, const char* filename, int* iflag);
void SUBR_(sdMat_read_bin)(T_(sdMat_ptr) M, const char* filename, int* iflag);
void SUBR_(mvec_to_mvec)(T_(const_mvec_ptr) v_in, T_(mvec_ptr) v_out, int* iflag);

/* phist_kernels_decl.h:368-438 -- views (non-owning; *Vblock reused if non-NULL)
 * and block copies */
void SUBR_(mvec_view_block)(T_(mvec_ptr) V, T_(mvec_ptr)* Vblock, int jmin, int jmax, int* iflag);
void SUBR_(mvec_get_block)(T_(const_mvec_ptr) V, T_(mvec_ptr) Vblock, int jmin, int jmax, int* iflag);
void SUBR_(mvec_set_block)(T_(mvec_ptr) V, T_(const_mvec_ptr) Vblock, int jmin, int jmax, int* iflag);
void SUBR_(sdMat_view_block)(T_(sdMat_ptr) M, T_(sdMat_ptr)* Mblock, int imin, int imax, int jmin, int jmax, int* iflag);
void SUBR_(sdMat_get_block)(T_(const_sdMat_ptr) M, T_(sdMat_ptr) Mblock, int imin, int imax, int jmin, int jmax, int* iflag);
void SUBR_(sdMat_set_block)(T_(sdMat_ptr) M, T_(const_sdMat_ptr) Mblock, int imin, int imax, int jmin, int jmax, int* iflag);
void SUBR_(sdMat_set_data)(T_(sdMat_ptr) M, const ST_* data_in, phist_lidx lda_in, int input_row_major, int* iflag);
void SUBR_(sdMat_get_data)(T_(const_sdMat_ptr) M, ST_* data_out, phist_lidx lda_in, int output_row_major, int* iflag);

/* phist_kernels_decl.h:448-486 -- fills.  mvec_random / sdMat_random draw from
 * PHIST's global KISS stream (src/tools/phist_random.c) with the same
 * skip-ahead accounting as kernels_common_impl_def.h:256-405. */
void SUBR_(mvec_put_value)(T_(mvec_ptr) V, ST_ value, int* iflag);
void SUBR_(mvec_put_func)(T_(mvec_ptr) V, phist_mvec_elemFunc elemFunPtr, void* last_arg, int* iflag);
void SUBR_(mvec_random)(T_(mvec_ptr) V, int* iflag);
void SUBR_(sdMat_random)(T_(sdMat_ptr) V, int* iflag);
void SUBR_(sdMat_put_value)(T_(sdMat_ptr) V, ST_ value, int* iflag);
void SUBR_(sdMat_identity)(T_(sdMat_ptr) M, int* iflag);
void SUBR_(mvec_print)(T_(const_mvec_ptr) V, int* iflag);
void SUBR_(sdMat_print)(T_(const_sdMat_ptr) M, int* iflag);

/* phist_kernels_decl.h:496-548 -- BLAS-1 on blocks (reductions complete on return) */
void SUBR_(mvec_norm2)(T_(const_mvec_ptr) V, MT_* vnrm, int* iflag);
void SUBR_(mvec_normalize)(T_(mvec_ptr) V, MT_* vnrm, int* iflag);
void SUBR_(mvec_scale)(T_(mvec_ptr) V, ST_ scalar, int* iflag);
void SUBR_(mvec_vscale)(T_(mvec_ptr) V, const ST_* scalar, int* iflag);
void SUBR_(mvec_add_mvec)(ST_ alpha, T_(const_mvec_ptr) X, ST_ beta, T_(mvec_ptr) Y, int* iflag);
void SUBR_(mvec_vadd_mvec)(const ST_ alpha[], T_(const_mvec_ptr) X, ST_ beta, T_(mvec_ptr) Y, int* iflag);
void SUBR_(mvec_times_mvec_elemwise)(ST_ alpha, T_(const_mvec_ptr) V, T_(mvec_ptr) W, int* iflag);
void SUBR_(mvec_dot_mvec)(T_(const_mvec_ptr) V, T_(const_mvec_ptr) W, ST_* vw, int* iflag);

/* phist_kernels_decl.h:557-585 -- tall-skinny GEMMs */
void SUBR_(mvecT_times_mvec)(ST_ alpha, T_(const_mvec_ptr) V, T_(const_mvec_ptr) W, ST_ beta, T_(sdMat_ptr) C, int* iflag);
void SUBR_(mvec_times_sdMat)(ST_ alpha, T_(const_mvec_ptr) V, T_(const_sdMat_ptr) C, ST_ beta, T_(mvec_ptr) W, int* iflag);
void SUBR_(mvec_times_sdMat_inplace)(T_(mvec_ptr) V, T_(const_sdMat_ptr) M, int* iflag);
void SUBR_(mvec_times_sdMat_add_mvec_times_sdMat)(T_(const_mvec_ptr) V, T_(const_sdMat_ptr) C, T_(mvec_ptr) W,
                                                  T_(const_sdMat_ptr) D, int* iflag);

/* phist_kernels_decl.h:597-632 -- small dense algebra on sdMats */
void SUBR_(sdMat_add_sdMat)(ST_ alpha, T_(const_sdMat_ptr) A, ST_ beta, T_(sdMat_ptr) B, int* iflag);
void SUBR_(sdMatT_add_sdMat)(ST_ alpha, T_(const_sdMat_ptr) A, ST_ beta, T_(sdMat_ptr) B, int* iflag);
void SUBR_(sdMat_times_sdMat)(ST_ alpha, T_(const_sdMat_ptr) V, T_(const_sdMat_ptr) W, ST_ beta, T_(sdMat_ptr) C, int* iflag);
void SUBR_(sdMatT_times_sdMat)(ST_ alpha, T_(const_sdMat_ptr) V, T_(const_sdMat_ptr) W, ST_ beta, T_(sdMat_ptr) C, int* iflag);
void SUBR_(sdMat_times_sdMatT)(ST_ alpha, T_(const_sdMat_ptr) V, T_(const_sdMat_ptr) W, ST_ beta, T_(sdMat_ptr) C, int* iflag);

/* phist_kernels_decl.h:644-667 -- sparse matrix times block vector on the
 * device SELL-C-sigma matrix; shifts[] is a host array of nvec values. */
void SUBR_(sparseMat_times_mvec_communicate)(T_(const_sparseMat_ptr) A, T_(const_mvec_ptr) x, int* iflag);
void SUBR_(sparseMat_times_mvec)(ST_ alpha, T_(const_sparseMat_ptr) A, T_(const_mvec_ptr) x, ST_ beta, T_(mvec_ptr) y, int* iflag);
void SUBR_(sparseMatT_times_mvec)(ST_ alpha, T_(const_sparseMat_ptr) A, T_(const_mvec_ptr) x, ST_ beta, T_(mvec_ptr) y, int* iflag);
void SUBR_(sparseMat_times_mvec_add_mvec)(ST_ alpha, T_(const_sparseMat_ptr) A, ST_ shift, T_(const_mvec_ptr) x,
                                          ST_ beta, T_(mvec_ptr) y, int* iflag);
void SUBR_(sparseMat_times_mvec_vadd_mvec)(ST_ alpha, T_(const_sparseMat_ptr) A, const ST_ shifts[], T_(const_mvec_ptr) x,
                                           ST_ beta, T_(mvec_ptr) y, int* iflag);

/* phist_kernels_decl.h:686 -- returns -99 like builtin, so callers use CholQR */
void SUBR_(mvec_QR)(T_(mvec_ptr) V, T_(sdMat_ptr) R, int* iflag);

/* phist_kernels_decl.h:700-728 -- construction from a row function: rows are
 * filled in order, each row stably sorted by column (builtin order), then the
 * device SELL-C-sigma copy is built. */
void SUBR_(sparseMat_create_fromRowFunc)(T_(sparseMat_ptr)* A, phist_const_comm_ptr comm, phist_gidx nrows, phist_gidx ncols,
                                         phist_lidx maxnne, phist_sparseMat_rowFunc rowFunPtr, void* last_arg, int* iflag);
void SUBR_(sparseMat_create_fromRowFuncAndContext)(T_(sparseMat_ptr)* vA, phist_const_context_ptr ctx, phist_lidx maxnne,
                                                   phist_sparseMat_rowFunc rowFunPtr, void* last_arg, int* iflag);
void SUBR_(sparseMat_create_fromRowFuncWithConstructor)(T_(sparseMat_ptr)* A, phist_const_comm_ptr comm, phist_gidx nrows,
                                                        phist_gidx ncols, phist_lidx maxnne, phist_sparseMat_rowFunc rowFunPtr,
                                                        phist_sparseMat_rowFuncConstructor rowFunConstructorPtr,
                                                        void* last_arg, int* iflag);
void SUBR_(sparseMat_create_fromRowFuncWithConstructorAndContext)(T_(sparseMat_ptr)* vA, phist_const_context_ptr ctx,
                                                                  phist_lidx maxnne, phist_sparseMat_rowFunc rowFunPtr,
                                                                  phist_sparseMat_rowFuncConstructor rowFunConstructorPtr,
                                                                  void* last_arg, int* iflag);

/* phist_kernels_decl.h:736-737 -- builtin-only gather/scatter of column blocks */
void SUBR_(mvec_gather_mvecs)(T_(mvec_ptr) V, T_(const_mvec_ptr) W[], int nblocks, int* iflag);
void SUBR_(mvec_scatter_mvecs)(T_(const_mvec_ptr) V, T_(mvec_ptr) W[], int nblocks, int* iflag);

/* phist_kernels_common_decl.h:28-75 -- derived helpers */
void SUBR_(sparseMat_global_nrows)(T_(sparseMat_ptr) A, phist_gidx* s, int* iflag);
void SUBR_(sparseMat_global_ncols)(T_(sparseMat_ptr) A, phist_gidx* s, int* iflag);
void SUBR_(mvec_my_length)(T_(const_mvec_ptr) V, phist_lidx* len, int* iflag);
void SUBR_(mvec_global_length)(T_(const_mvec_ptr) V, phist_gidx* len, int* iflag);
void SUBR_(mvec_get_comm)(T_(const_mvec_ptr) V, phist_const_comm_ptr* comm, int* iflag);
void SUBR_(mvec_clone_shape)(T_(mvec_ptr)* V, T_(const_mvec_ptr) V_in, int* iflag);
void SUBR_(mvec_set_data)(T_(mvec_ptr) V, const ST_* data, phist_lidx lda, int input_row_major, int* iflag);

/* phist_kernels_fused_decl.h:45-92 -- fused kernels.  fused_spmv_mvdot makes
 * one pass over A, V and W (the builtin fallback makes three); WdotW / VdotW
 * may be NULL.  With PHIST_NO_GLOBAL_REDUCTION the local result is returned,
 * which on one process is the global one. */
void SUBR_(fused_spmv_pair)(ST_ alpha, ST_ const shifts1[], T_(const_sparseMat_ptr) A1, ST_ const shifts2[],
                            T_(const_sparseMat_ptr) A2, T_(const_mvec_ptr) X, ST_ beta, T_(mvec_ptr) Y, int* iflag);
void SUBR_(fused_spmv_mvdot)(ST_ alpha, T_(const_sparseMat_ptr) A, T_(const_mvec_ptr) V, ST_ beta, T_(mvec_ptr) W,
                             ST_* WdotW, ST_* VdotW, int* iflag);
void SUBR_(fused_spmv_mvTmv)(ST_ alpha, T_(const_sparseMat_ptr) A, T_(const_mvec_ptr) V, ST_ beta, T_(mvec_ptr) W,
                             T_(sdMat_ptr) WtW, T_(sdMat_ptr) VtW, int* iflag);
void SUBR_(fused_spmv_mvdot_mvadd)(ST_ alpha, T_(const_sparseMat_ptr) A, T_(const_mvec_ptr) V, ST_ beta, T_(mvec_ptr) W,
                                   ST_ gamma, ST_ delta, T_(mvec_ptr) U, ST_* lWdotW, ST_* lVdotW, int* iflag);
void SUBR_(fused_mvsdi_mvTmv)(ST_ alpha, T_(const_mvec_ptr) V, T_(mvec_ptr) W, T_(const_sdMat_ptr) C, ST_ beta,
                              T_(sdMat_ptr) D, int* iflag);
void SUBR_(fused_mvsd_mvTmv)(ST_ alpha, T_(const_mvec_ptr) V, T_(const_sdMat_ptr) C, ST_ beta, T_(mvec_ptr) W,
                             T_(sdMat_ptr) WtW, int* iflag);

/* phist_kernels_carp_decl.h:35-102 -- CARP sweeps (Kaczmarz), run on the
 * device by multicolouring: rows of one colour share no entry of x and update
 * in parallel; a sweep runs the colours forward, then backward.  Column j of
 * X uses shift sigma[j] and relaxation omega[j]; Rhs may be NULL (b = 0).
 * The augmented variants (_aug, _aug_rc) return -99, the reference's answer
 * of kernel libraries without them (src/kernels/common/kernels_no_carp.cpp). */
void SUBR_(carp_setup)(T_(const_sparseMat_ptr) A, int numShifts, ST_ const sigma[], void** work, int* iflag);
void SUBR_(carp_sweep)(T_(const_sparseMat_ptr) A, ST_ const sigma[], T_(const_mvec_ptr) Rhs, T_(mvec_ptr) X,
                       void* const work, MT_ const* omega, int* iflag);
void SUBR_(carp_sweep_aug)(T_(const_sparseMat_ptr) A, ST_ const sigma[], T_(const_mvec_ptr) Q, T_(const_mvec_ptr) Rhs,
                           T_(mvec_ptr) X, T_(sdMat_ptr) q, void* const work, MT_ const* omega, int* iflag);
void SUBR_(carp_setup_rc)(T_(const_sparseMat_ptr) A, int numShifts, MT_ const sigma_r[], MT_ const sigma_i[],
                          void** work, int* iflag);
void SUBR_(carp_sweep_rc)(T_(const_sparseMat_ptr) A, MT_ const sigma_r[], MT_ const sigma_i[], T_(const_mvec_ptr) Rhs,
                          T_(mvec_ptr) X_r, T_(mvec_ptr) X_i, void* const work, MT_ const* omega, int* iflag);
void SUBR_(carp_sweep_aug_rc)(T_(const_sparseMat_ptr) A, MT_ const sigma_r[], MT_ const sigma_i[], T_(const_mvec_ptr) Q,
                              T_(const_mvec_ptr) Rhs, T_(mvec_ptr) X_r, T_(mvec_ptr) X_i, T_(sdMat_ptr) q_r,
                              T_(sdMat_ptr) q_i, void* const work, MT_ const* omega, int* iflag);
void SUBR_(carp_destroy)(T_(const_sparseMat_ptr) A, void* work, int* iflag);

/* ---- orthogonalisation (src/core), run on the device ---- */

/* src/core/phist_sdFact_decl.h:22-47 -- host FP64 factorisations of sdMats:
 * rank-revealing pivoted Cholesky (M = R'R, R(:,perm) upper triangular),
 * permuted triangular solves, and the SVQB factor B (scaled eigenvectors). */
void SUBR_(sdMat_cholesky)(T_(sdMat_ptr) M, int* perm, int* rank, MT_ rankTol, int* iflag);
void SUBR_(sdMat_backwardSubst_sdMat)(const T_(sdMat_ptr) R, int* perm, int rank, T_(sdMat_ptr) X, int* iflag);
void SUBR_(sdMat_forwardSubst_sdMat)(const T_(sdMat_ptr) R, int* perm, int rank, T_(sdMat_ptr) X, int* iflag);
void SUBR_(sdMat_qb)(T_(sdMat_ptr) B, T_(sdMat_ptr) B_1, MT_ nrmsV[], int* rank, MT_ rankTol, int* iflag);

/* src/core/phist_chol_QR_decl.h:30-38 -- CholQR: Gram on the device, pivoted
 * Cholesky on the host, fused in-place V*R^-1 on the device; returns
 * iflag = ncols - rank (null space randomised and re-orthogonalised). */
void SUBR_(chol_QRp)(T_(mvec_ptr) V, T_(sdMat_ptr) R, int perm[], int* iflag);
void SUBR_(chol_QR)(T_(mvec_ptr) V, T_(sdMat_ptr) R, int* iflag);

/* src/core/phist_svqb_decl.h:30 -- SVQB (Stathopoulos & Wu) */
void SUBR_(svqb)(T_(mvec_ptr) V, T_(sdMat_ptr) B, MT_* e, int* iflag);

/* src/core/phist_orthog_decl.h:64 -- W = Q*R1 + V*R2 with Q'Q = I, Q orthogonal
 * to V (which must have orthonormal columns), by block classical Gram-Schmidt
 * and CholQR sweeps (the orthogrr scheme).  B must be NULL (no B-inner
 * product: -99 otherwise).  rankVW returns the rank of [V W]. */
void SUBR_(orthog)(T_(const_mvec_ptr) V, T_(mvec_ptr) W, T_(const_linearOp_ptr) B, T_(sdMat_ptr) R1, T_(sdMat_ptr) R2,
                   int numSweeps, int* rankVW, int* iflag);

#ifdef __cplusplus
}
#endif

#undef ST_
#undef MT_
#undef SUBR_
#undef T_
