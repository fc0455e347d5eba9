/*
 * phist_amd_ext.h -- extensions of libphist_kernels_amd.so.
 *
 * None of these has a counterpart in the reference's kernel interface.  They
 * are for tuning, measurement and inspection (benchmarks, parity tests) and
 * for building test problems; solvers written against phist_amd.h never need
 * them.
 */
#ifndef PHIST_AMD_EXT_H
#define PHIST_AMD_EXT_H

#include "phist_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* SELL-C-sigma parameters for matrices created after the call
 * (C a power of two, 0 = default 32; sigma 0 = chosen per matrix: 1, which
 * keeps the rows in their own order, when that pads the chunks by at most
 * 25 %, else 256). */
void phist_amd_sparseMat_set_format(int C, int sigma, int* iflag);
/* upper bound for the persistent grid of the streaming kernels (0 = one
 * resident wave of workgroups on every CU) */
void phist_amd_set_grid_blocks(int nblocks, int* iflag);
/* workgroups per CU of the SpMV's persistent grid (0 = as many as fit;
 * n < 0 = the built-in table, the default: 2 for rows of >= 24 nonzeros and
 * blocks of >= 4 columns, whose gathers then keep a wide band's x rows
 * L2-resident, 3 for such rows and narrower blocks, 6 for shorter rows) */
void phist_amd_set_spmv_blocks_per_cu(int n, int* iflag);
/* bytes of a block row each lane of the SpMV gathers per nonzero (0 = the
 * default, one 16-byte vector; sizeof(scalar) = one scalar per lane; 32 or 64
 * = two or four 16-byte vectors per lane, taken when the block width and
 * leading dimension are multiples of that many scalars).  Every choice gives
 * the same result bit for bit: a row's sum is always formed by the lane that
 * owns its columns, in the row's entry order. */
void phist_amd_set_spmv_lane_bytes(int bytes, int* iflag);
/* entries of a row whose val/col loads and x gathers the SpMV issues
 * together before their multiply-adds (0 = the default: the matrix's tuned
 * value, else 8; 8, 16 or 32).  16 and 32 also request the next batch's
 * val/col before the current batch's gathers; they apply to row groups of at
 * most 8 lanes gathering at most 32 bytes each, without per-column x
 * factors, and fall back to 8 otherwise.  Every choice
 * gives the same result bit for bit. */
void phist_amd_set_spmv_depth(int slots, int* iflag);
/* block columns per SpMV launch: ncols > 0 = at most that many (a multiple of
 * the scalars each lane holds, else ignored), 0 = as many as 64 lane groups
 * hold, -1 = the default: the matrix's tuned value, else 0.  Each further
 * launch reads the matrix again; wide rows of x (256-byte rows at 32 FP64
 * columns) gather half as many bytes per launch in two blocks of 16.  Every
 * choice gives the same result bit for bit (column dots to rounding). */
void phist_amd_set_spmv_col_block(int ncols, int* iflag);
/* x staged in LDS: mode 1 = each workgroup first copies the x rows its tile
 * of SELL rows gathers from (the tile's column intervals, found once per
 * matrix and tile height) into LDS with 16-byte loads, then gathers from
 * there; used where every tile's window fits the workgroup's share of LDS
 * (stencil-like matrices: a 3D Laplacian tile gathers from three short
 * intervals) and the block rows are whole 16-byte vectors (at most 8 lanes
 * per row, no per-column x factors, no Gram epilogue), else ignored.  0 = off,
 * -1 = the default: the matrix's tuned value, else off.  Every choice gives
 * the same result bit for bit. */
void phist_amd_set_spmv_stage(int mode, int* iflag);
/* results of host reductions and small device-to-host copies (dots, norms,
 * sdMat downloads): enable = 1 (the default, PHIST_AMD_FAST_SYNC) lets the
 * device write them with a sequence number into coherent pinned memory and
 * the host wait for that number; 0 copies them and synchronises the stream.
 * The values are the same either way. */
void phist_amd_set_fast_sync(int enable, int* iflag);
/* blocked GMRES with the per-system Gram-Schmidt sweeps: enable = 1 (the
 * default, PHIST_AMD_GMRES_DEVICE) runs the Givens rotations on the device
 * and lets the host enqueue the next Arnoldi step before the previous one's
 * convergence flags arrive; 0 computes them on the host after each step.
 * The two agree to rounding (the rotations' hypot may differ in the last bit). */
void phist_amd_set_gmres_device(int enable, int* iflag);
/* SpMV launches so far (this process) that staged x in LDS */
void phist_amd_spmv_staged_launches(int64_t* count, int* iflag);
/* SpMV autotuner.  For each block width nvecs[0..n_nvec) (1..256, at most 16
 * widths) it times the product y = A*x on the device over the gather width
 * per lane (one scalar, or one, two or four 16-byte vectors: how many lanes
 * share a row), the workgroups per CU (1, 2, 3, 4, 6, 8), for rows of at
 * least 16 entries the depth (8, 16 or 32, phist_amd_set_spmv_depth), for more
 * than 16 columns launches of 16 (and of 32 above 32) columns
 * (phist_amd_set_spmv_col_block), where the matrix allows it x staged in LDS
 * per workgroup tile (phist_amd_set_spmv_stage), and with
 * flags & 1 also over the SELL-C-sigma layout (C = 16, 32, 64; sigma = 1,
 * 256; the device copy is rebuilt from the host CRS for each).  The best
 * layout (least total time over the widths) is kept, and the best launch
 * settings of each width are stored on the matrix: later SpMVs of that width
 * use them unless phist_amd_set_spmv_lane_bytes / _blocks_per_cu / _depth / _col_block / _stage set global
 * values.  Choices are cached for the process by matrix signature (scalar
 * type, rows, entries, longest row, mean distance of entries from the
 * diagonal, block width); flags & 2 measures again instead of using the
 * cache; flags & 4 (D only) also tunes the matrix's FP32 copy used by the
 * mixed-precision solvers (launch settings, same widths).  Outputs may be
 * NULL: default_ms[i] / tuned_ms[i] = fastest of three products with the
 * library's untuned choice on the incoming layout / with the chosen settings;
 * settings[7*i..7*i+6] = C, sigma, lane bytes, workgroups per CU, depth, column block,
 * staging (1 = x staged in LDS).  Products
 * are identical bit for bit under every setting. */
void phist_Damd_sparseMat_tune(phist_DsparseMat_ptr A, int n_nvec, const int* nvecs, int flags, double* default_ms,
                               double* tuned_ms, int* settings, int* iflag);
void phist_Samd_sparseMat_tune(phist_SsparseMat_ptr A, int n_nvec, const int* nvecs, int flags, double* default_ms,
                               double* tuned_ms, int* settings, int* iflag);
/* the device-resident MINRES iteration (blocked MINRES, the eigensolver's
 * correction equations, the service loop): iters_per_call iterations are
 * queued per host step, never more than an unfinished system may still take
 * (0, the default: 2 for blocks of at most 4M elements, else 1); enable != 0 (default 0) replays runs of graph_len (>= 1,
 * default 4) iterations from captured HIP graphs (one per layout of the
 * rotating blocks, captured on first use and again when the projection
 * basis changes) instead of launching their kernels one by one.  Results are
 * the same bit for bit either way.  stats: graphs captured and graph
 * launches since initialisation. */
void phist_amd_set_minres_graphs(int enable, int iters_per_call, int graph_len, int* iflag);
void phist_amd_minres_graph_stats(int64_t* captures, int64_t* launches, int* iflag);

/* the library's HIP stream (hipStream_t), device synchronisation, and an
 * event timer on that stream (ms between start and stop) */
void phist_amd_get_stream(void** stream, int* iflag);
void phist_amd_synchronize(int* iflag);
void phist_amd_timer_start(int* iflag);
void phist_amd_timer_stop(float* ms, int* iflag);
void phist_amd_device_info(char* name, int name_len, int* num_cu, int* iflag);
/* per-kernel timer: with enable = 1 the main kernel of every SpMV,
 * tall-skinny GEMM and BLAS-1 call is bracketed by HIP events on the library
 * stream (synchronising after it); read returns the summed milliseconds and
 * the number of timed launches since the last enable.  0 stops it; 2 pauses
 * and 3 resumes it without clearing the totals (to leave a call out). */
void phist_amd_kernel_timer(int enable, int* iflag);
void phist_amd_kernel_timer_read(double* total_ms, int64_t* count, int* iflag);
/* known-size gather: every seg_bytes-byte segment (16..256, a power of two)
 * of a 1 GiB device array read exactly once in a scrambled order, one 16-byte
 * load per lane; mean and best GB/s of 20 launches.  A counter profile of it
 * calibrates the L2 fetch counters for gathers of that size. */
void phist_amd_bench_gather(int seg_bytes, double* mean_bw, double* max_bw, int* iflag);
/* the device's streaming ceiling as this library's kernels can reach it:
 * kind 0 reads, kind 1 copies a 1 GiB array with 16-byte accesses, four in
 * flight per thread, each workgroup over a contiguous part; mean and best GB/s of 20
 * launches (bytes read + written).  phist_bench_stream_* are the reference's
 * plain STREAM kernels (one 8-byte element per thread and step). */
void phist_amd_bench_hbm(int kind, double* mean_bw, double* max_bw, int* iflag);

/* phist_Dfused_spmv_mvdot that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * W = alpha*A*V + beta*W as phist_Dfused_spmv_mvdot, with the column dots
 * WdotW[j] = W_j.W_j and VdotW[j] = V_j.W_j written to DEVICE arrays of nvec
 * doubles (either may be NULL), asynchronously on the library stream
 * (phist_amd_get_stream): the call does not wait for the device.  Same kernels
 * and summation order as phist_Dfused_spmv_mvdot, so W and the dots equal
 * its results bit for bit.  The reduction partials live in a fixed device
 * array of 2 MiB per GPU, held until the process ends (enough for 64 columns
 * on the persistent grid of up to 2047 workgroups); a call needing more, a
 * block of more than 64 columns or an empty block fails with -55
 * (PHIST_INVALID_INPUT) before anything is launched.  Every call on one GPU
 * uses the same array: calls ordered on the library stream are safe, calls
 * that overlap on one GPU (several host threads) race. */
void phist_Dfused_spmv_mvdot_noalloc(double alpha, phist_Dconst_sparseMat_ptr A, phist_Dconst_mvec_ptr V,
                                     double beta, phist_Dmvec_ptr W, double* WdotW, double* VdotW, int* iflag);

/* the global KISS stream: reset to the initial state, jump ahead by n draws,
 * draw n doubles in [-1,1) */
void phist_amd_kiss_reset(int* iflag);
void phist_amd_kiss_skip(uint64_t n, int* iflag);
void phist_amd_kiss_draw(int64_t n, double* out, int* iflag);

/* the host's symmetric eigensolvers (no device work): a (n x n, column-major,
 * lda) is overwritten by the eigenvectors, w gets the eigenvalues ascending;
 * method 0 = cyclic Jacobi (SVQB), 1 = Householder + implicit QL (the
 * eigensolver's Rayleigh-Ritz problems) */
void phist_amd_dense_sym_eig(double* a, int n, int lda, double* w, int method, int* iflag);

/* host copies of a matrix's index data.  get_crs: builtin-order CRS with
 * 1-based row_offset (nrows+1) and col_idx (nnz).  get_sell: C, sigma,
 * chunk count and padded slot count; arrays may be NULL to query sizes only
 * (chunk_ptr nchunks+1, chunk_w nchunks, perm nchunks*C, col/val nslots). */
void phist_Damd_sparseMat_get_crs(phist_Dconst_sparseMat_ptr A, int64_t* row_offset, int32_t* col_idx, double* val,
                                  int* iflag);
void phist_Samd_sparseMat_get_crs(phist_Sconst_sparseMat_ptr A, int64_t* row_offset, int32_t* col_idx, float* val,
                                  int* iflag);
void phist_Damd_sparseMat_get_sell(phist_Dconst_sparseMat_ptr A, int* C, int* sigma, int64_t* nchunks,
                                   int64_t* nslots, int64_t* chunk_ptr, int32_t* chunk_w, int32_t* perm,
                                   int32_t* col, double* val, int* iflag);
void phist_Samd_sparseMat_get_sell(phist_Sconst_sparseMat_ptr A, int* C, int* sigma, int64_t* nchunks,
                                   int64_t* nslots, int64_t* chunk_ptr, int32_t* chunk_w, int32_t* perm,
                                   int32_t* col, float* val, int* iflag);

/* test-problem row functions (phist_sparseMat_rowFunc); the _f variants
 * write float values. */
typedef struct {
  int64_t n;          /* rows = columns */
  int32_t nnz_per_row;/* including the diagonal */
  int64_t halfband;   /* off-diagonal columns within +-halfband */
  double diag;        /* diagonal value */
  uint64_t seed;
} phist_amd_banded_params;
typedef struct {
  int64_t nx, ny, nz;
} phist_amd_laplace3d_params;
int phist_amd_banded_random_rowfunc(phist_gidx row, phist_lidx* nnz, phist_gidx* cols, void* vals, void* arg);
int phist_amd_banded_random_rowfunc_f(phist_gidx row, phist_lidx* nnz, phist_gidx* cols, void* vals, void* arg);
int phist_amd_laplace3d_rowfunc(phist_gidx row, phist_lidx* nnz, phist_gidx* cols, void* vals, void* arg);
int phist_amd_laplace3d_rowfunc_f(phist_gidx row, phist_lidx* nnz, phist_gidx* cols, void* vals, void* arg);
int phist_amd_rowfunc_shared_ctor(void* arg, void** work);

/* Heisenberg/XXZ chain of L spins in the sector of nup up spins
   (drivers/matfuncs/SpinChainSZ.c:189-310, called by the reference driver
   with {L, L/2, 0} and jz = jxy = 1): C(L, nup) rows, at most L+1 entries. */
typedef struct {
  int32_t L;
  int32_t nup;
  int32_t obc; /* 0: periodic chain, 1: open */
  int32_t pad_;
  double jz;
  double jxy;
} phist_amd_spinchain_sz_params;
void phist_amd_spinchain_sz_dims(const phist_amd_spinchain_sz_params* p, phist_gidx* nrows, phist_lidx* maxnne,
                                 int* iflag);
int phist_amd_spinchain_sz_rowfunc(phist_gidx row, phist_lidx* nnz, phist_gidx* cols, void* vals, void* arg);
int phist_amd_spinchain_sz_rowfunc_f(phist_gidx row, phist_lidx* nnz, phist_gidx* cols, void* vals, void* arg);

/* tight-binding honeycomb (graphene) sheet of W x L sites
   (drivers/matfuncs/graphene.c:38-245): nearest neighbours -1, with
   long_range also -0.1/2.78 (second) and -0.095/2.78 (third neighbours);
   on-site gamma*(u-1/2) with u from a per-row seeded stream. */
typedef struct {
  int64_t W;
  int64_t L;
  int32_t pbc_w;
  int32_t pbc_l;
  int32_t long_range;
  int32_t pad_;
  double gamma;
  uint64_t seed;
} phist_amd_honeycomb_params;
void phist_amd_honeycomb_dims(const phist_amd_honeycomb_params* p, phist_gidx* nrows, phist_lidx* maxnne,
                              int* iflag);
int phist_amd_honeycomb_rowfunc(phist_gidx row, phist_lidx* nnz, phist_gidx* cols, void* vals, void* arg);
int phist_amd_honeycomb_rowfunc_f(phist_gidx row, phist_lidx* nnz, phist_gidx* cols, void* vals, void* arg);

/* colour classes of a CARP work object (from phist_[DS]carp_setup or _rc):
   ncolors colours; rows rows[color_ptr[c] .. color_ptr[c+1]) have colour c,
   ascending.  color_ptr needs ncolors+1 entries, rows nrows; either may be
   NULL (query ncolors first).  A sweep visits colours 0..ncolors-1, then
   ncolors-1..0. */
void phist_amd_Dcarp_colors(void* work, int* ncolors, int64_t* color_ptr, int32_t* rows, int* iflag);
void phist_amd_Scarp_colors(void* work, int* ncolors, int64_t* color_ptr, int32_t* rows, int* iflag);

#ifdef __cplusplus
}
#endif

#endif /* PHIST_AMD_EXT_H */
