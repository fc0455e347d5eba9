// amd_common.hpp -- internal object model of libphist_kernels_amd.
//
// Objects behind the opaque handles of include/phist_amd.h:
//   Comm       single-process communicator (rank 0 of 1)
//   Map        contiguous row distribution (nglob rows, all local)
//   Context    row/col/range/domain maps of a sparse matrix
//   Mvec<T>    device-resident row-major block vector with a lazily created
//              pinned host mirror; views share the root's storage
//   SdMat<T>   column-major small dense matrix mirrored on host and device
//   SparseMat<T>  SELL-C-sigma on the device + the builtin-order CRS on the host
//
// Every device kernel runs on the library's stream (rt().stream); every entry
// point returns only after the results it hands to the host are complete.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <vector>
#include <string>
#include <stdexcept>

#include "phist_amd_ext.h"

namespace phx {

// ----------------------------------------------------------------------------
// error handling: every entry point converts exceptions to iflag codes
// ----------------------------------------------------------------------------
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const char* fmt, ...);

#define PHX_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      ::phx::fail(PHIST_FUNCTIONAL_ERROR, "%s:%d: HIP error %s in %s", __FILE__, __LINE__, \
                  hipGetErrorString(e_), #call);                                       \
  } while (0)

#define PHX_REQUIRE(cond, code, ...)          \
  do {                                        \
    if (!(cond)) ::phx::fail(code, __VA_ARGS__); \
  } while (0)

// Runs body, maps exceptions to *iflag and prints the message unless quiet.
template <class F>
inline void guard(int* iflag, const char* fn, F&& body) {
  try {
    body();
  } catch (const Error& e) {
    if (iflag) *iflag = e.code;
    if (e.code < 0 && e.code != PHIST_NOT_IMPLEMENTED) fprintf(stderr, "[phist_amd] %s: %s\n", fn, e.what());
  } catch (const std::bad_alloc&) {
    if (iflag) *iflag = PHIST_MEM_ALLOC_FAILED;
    fprintf(stderr, "[phist_amd] %s: host allocation failed\n", fn);
  } catch (const std::exception& e) {
    if (iflag) *iflag = PHIST_CAUGHT_EXCEPTION;
    fprintf(stderr, "[phist_amd] %s: %s\n", fn, e.what());
  }
}

// ----------------------------------------------------------------------------
// runtime: device, stream, workspace
// ----------------------------------------------------------------------------
struct Runtime {
  bool initialized = false;
  int device = 0;
  int num_cu = 256;
  int64_t lds_per_block = 65536;  // LDS a workgroup may allocate (static + dynamic)
  int num_xcd = 8;
  hipStream_t stream = nullptr;
  hipEvent_t ev_start = nullptr, ev_stop = nullptr;
  // device scratch for block partials of reductions (grown on demand)
  void* d_work = nullptr;
  size_t d_work_bytes = 0;
  // pinned host scratch for reduction results
  void* h_work = nullptr;
  size_t h_work_bytes = 0;
  // results of host reductions published by the device: coherent pinned
  // memory, a sequence word at offset 0 and the values from kSignalData on
  // (signal_buffer); fast_sync = 0 takes the copy + stream synchronisation
  static constexpr size_t kSignalBytes = 64 * 1024, kSignalData = 256;
  void* h_sig = nullptr;
  void* d_sig = nullptr;
  uint32_t sig_seq = 0;
  int fast_sync = 1;
  // pinned staging slots of stage_upload_2d, each with the event of its last copy
  static constexpr int kStageSlots = 64;
  static constexpr size_t kStageSlotBytes = 64 * 1024;
  char* stage_base = nullptr;
  hipEvent_t stage_ev[kStageSlots] = {};
  bool stage_rec[kStageSlots] = {};
  int stage_next = 0;
  // default SELL parameters for newly built matrices (0 = automatic)
  int sell_C = 0, sell_sigma = 0;
  // maximum persistent grid size for streaming kernels (0 = automatic)
  int grid_blocks = 0;
  // workgroups per CU for the SpMV: the x gathers of a +-band thrash L2 when
  // too many rows are in flight, two workgroups per CU measured best
  int spmv_blocks_per_cu = -1;  // < 0: spmv_table_blocks_per_cu
  // bytes of a block row each SpMV lane gathers per nonzero (0: one 16-byte vector)
  int spmv_lane_bytes = 0;
  // entries per row whose val/col loads and x gathers the SpMV issues
  // together (8 or 32; < 0: the matrix's tuned value, else 8)
  int spmv_depth = -1;
  // y of the SpMV written with non-temporal stores, which keeps L2 for the
  // gathered x rows: 1 always, 0 never, -1 (default) when y spans at least
  // kNtYBytes (more than the L2s hold, so its next reader cannot find it
  // there) and its rows are shorter than a 128-byte line (measured: rows of
  // 32 and 64 bytes gain 1 %, rows of 128 and 256 bytes lose 4-5 %)
  int spmv_nt_y = -1;
  // block columns per SpMV launch: > 0 at most that many, 0 as many as 64
  // lane groups hold, -1 (default) the matrix's tuned value
  int spmv_col_block = -1;
  // x rows of each workgroup tile staged in LDS before the tile's gathers
  // (matrices whose tiles gather from a few short column intervals): 1 on
  // where the matrix allows it, 0 off, -1 (default) the matrix's tuned value
  int spmv_stage = -1;
  int64_t spmv_staged_launches = 0;  // SpMV launches that staged x in LDS (a statistic)
  static constexpr int64_t kNtYBytes = int64_t(64) << 20;
  // Gram products with both operands >= 16 columns run on the matrix cores
  bool gram_mfma = true;
  // ... when the wider side has at least gram_mfma_wide columns, or an
  // operand is a column view of a wider block, or a self Gram has >= 8
  // (both sides zero-padded to 16 on the matrix cores; measured on 2.7M rows:
  // the VALU kernels win only for compact blocks of <= 8 columns)
  int gram_mfma_wide = 9;
  // workgroups per CU of the Gram kernels (each wave leaves one partial tile,
  // so this also bounds the finalize work)
  int gram_blocks_per_cu = 8;
  // row-mixing products (V*C, in-place V*M) with k, m >= 8 run on the matrix cores
  bool mix_mfma = true;
  // ... for at least this many output columns (and 8..64 inputs); narrow
  // outputs are zero-padded to 16 (measured faster than the VALU kernels on
  // 2.7M rows for 1..4 output columns from 8..44 inputs, FP64 and FP32)
  int mix_mfma_min_m = 1;
  bool minres_device = true;  // MINRES scalars on the device for this library's matrices
  // blocked GMRES: Gram-Schmidt with the per-system sweeps (gmres_cgs_dev)
  // where the block allows, else with Grams and row mixes of the whole basis
  bool gmres_cgs = true;
  // ... and with the sweeps, the Givens rotations on the device (the host
  // runs one Arnoldi step ahead, reading the previous step's active flags)
  bool gmres_device = true;
  // service loop: most MINRES iterations per batch and step while eigenvalue requests run
  int serve_max_reps = 4;  // (rate 40/s mix: p99 92.4 ms at 1, 84.3 at 4, 84.6 at 8)
  // device MINRES: iterations enqueued per iterate() call (at most; fewer when
  // no active system can take that many more), and whether runs of
  // minres_graph_len iterations are replayed from captured HIP graphs instead
  // of launching their kernels one by one (measured no faster on MI355X for
  // 3k..262k rows, so off by default)
  int minres_iters_per_call = 0;  // 0: 2 for blocks of <= 4M elements (host-bound), else 1
  bool minres_graphs = false;
  int minres_graph_len = 4;
  int64_t graph_captures = 0, graph_launches = 0;
  // device_work while a graph is captured: every request is answered with
  // work_arena (the capture records, nothing runs), the largest request is
  // kept in work_need so that the capturer can check its arena afterwards
  bool capturing = false;
  void* work_arena = nullptr;
  size_t work_need = 0;
  // the allocation-free entry points: address of the fixed device scratch
  // array on each device (hipGetSymbolAddress at initialisation), and while
  // such a call runs, the buffer device_work answers with (never allocating;
  // larger requests are rejected) -- host_work is refused meanwhile
  static constexpr int kMaxDevices = 64;
  void* noalloc_scratch[kMaxDevices] = {};
  void* fixed_work = nullptr;
  size_t fixed_work_bytes = 0;
  bool gram_rows = true;  // narrow unaligned Grams: several columns per lane (k_gram_rows)
  bool gram_sets = true;  // self Grams of 2, 4 or 8 columns: row sets on the matrix cores (k_gram_mfma_sets)
  // optional timing of the main kernel of each hot-path call (HIP events on
  // the library stream around that launch only)
  bool ktimer_on = false;
  hipEvent_t kt_start = nullptr, kt_stop = nullptr;
  double ktimer_ms = 0.0;
  int64_t ktimer_count = 0;
};
Runtime& rt();
void ensure_init();
void* device_work(size_t bytes);
void* host_work(size_t bytes);
// the published-result buffer (host and device addresses of the same coherent
// pinned memory), allocated on first use; false where fast_sync is off or
// host reductions are refused (allocation-free calls, graph capture)
bool signal_buffer(void** host, void** dev);
// waits until the device has written seq to the buffer's sequence word: a
// short spin, then a stream synchronisation (after which the word must hold seq)
void wait_signal(uint32_t seq);
// copies bytes from device memory to host memory through the published-result
// buffer (a one-wave kernel writes them and a new sequence number; the host
// waits for the number); false, with nothing copied, where that buffer is
// unavailable or too small or the source is not 4-byte aligned
bool fetch_to_host(void* host, const void* dev, size_t bytes);
// asynchronous form for loops that run ahead of their results: enqueues the
// copy of bytes (at most kSignalSlotBytes) into published slot `slot` (1 or
// 2; fetch_to_host uses slot 0) and returns its sequence number (0: the
// buffer is unavailable, nothing enqueued); published_data waits for that
// number or a later one and returns the slot's host address.  A slot must be
// read before it is written again.
constexpr size_t kSignalSlotBytes = 4096;
uint32_t publish_async(const void* dev, size_t bytes, int slot);
const void* published_data(uint32_t seq, int slot);
void sync_stream();
// Fixed device scratch of the allocation-free entry points: kNoAllocScratchBytes
// per device, held until the process ends.  noalloc_scratch() returns this
// device's address of it (resolved at initialisation).  NoAllocScope makes
// device_work answer every request from it for the scope's lifetime (a
// request larger than the array fails with PHIST_INVALID_INPUT before
// anything is launched) and host_work fail.
constexpr size_t kNoAllocScratchBytes = (size_t)2 << 20;
void* noalloc_scratch();
struct NoAllocScope {
  NoAllocScope();
  ~NoAllocScope();
};
// Queues a 2-D host-to-device copy (cols columns of row_bytes bytes, pitches
// in bytes) from a pinned staging slot holding a copy of the source, so the
// caller may change or free its array at once; false when the block does not
// fit a slot (nothing queued)
bool stage_upload_2d(void* dst, size_t dpitch, const void* src, size_t spitch, size_t row_bytes, size_t cols);

// Caching allocators for block vectors and small dense matrices.  Released
// device blocks go back to a size-class cache and are handed out again
// without synchronising: every use of a block is ordered on the library
// stream.  Pinned host blocks are cached after a stream synchronisation
// (a copy from or into them may still be queued).  Exceeding the cache caps
// or a failed allocation returns cached blocks to the runtime.
void* dev_alloc(size_t bytes);
void dev_release(void* p, size_t bytes);
void* pinned_alloc(size_t bytes);
void pinned_release(void* p, size_t bytes);
void pool_trim();
// brackets one kernel launch when the kernel timer is on
// resident workgroups per CU of a kernel launched with `threads` threads
// and lds_bytes of dynamic LDS (the runtime's occupancy query, asked once
// per kernel, workgroup size, LDS size and device; at least 1)
int occupancy_per_cu(const void* kernel, int threads, size_t lds_bytes);
void ktimer_begin();
void ktimer_end();  // synchronises the event and accumulates the time

// ----------------------------------------------------------------------------
// communicator, map, context
// ----------------------------------------------------------------------------
struct Comm {
  int rank = 0, size = 1;
};
Comm* default_comm();

struct Map {
  const Comm* comm = nullptr;
  int64_t nglob = 0;
  int32_t nloc = 0;
  int64_t ilower = 0;
};

struct Context {
  const Map* row_map = nullptr;
  const Map* col_map = nullptr;
  const Map* range_map = nullptr;
  const Map* domain_map = nullptr;
};

// ----------------------------------------------------------------------------
// block vectors
// ----------------------------------------------------------------------------
template <typename T>
struct Mvec {
  const Map* map = nullptr;
  int64_t nrows = 0;
  int nvec = 0;       // columns of this (view of the) block
  int64_t lda = 0;    // row stride of the root allocation (elements)
  Mvec* root = this;  // owner of the storage
  int jofs = 0;       // first column inside the root
  T* dev_base = nullptr;   // root only
  T* host_base = nullptr;  // root only, pinned, created on demand
  int nviews = 0;          // live views of this root
  bool is_view() const { return root != this; }
  T* dev() const { return root->dev_base + jofs; }
  T* host();  // creates the root mirror on first use
};

// ----------------------------------------------------------------------------
// small dense matrices (column-major, host + device copies)
// ----------------------------------------------------------------------------
template <typename T>
struct SdMat {
  int nrows = 0, ncols = 0;
  int ld = 0;             // column stride of the root (elements)
  SdMat* root = this;
  int iofs = 0, jofs = 0;
  T* host_base = nullptr;  // root only
  T* dev_base = nullptr;   // root only
  int nviews = 0;
  bool is_view() const { return root != this; }
  T* host() const { return root->host_base + (int64_t)jofs * root->ld + iofs; }
  T* dev() const { return root->dev_base + (int64_t)jofs * root->ld + iofs; }
  T& h(int i, int j) const { return host()[(int64_t)j * root->ld + i]; }
};
template <typename T> void sd_to_device(const SdMat<T>* M);
template <typename T> void sd_from_device(const SdMat<T>* M);
// queues the device-to-host copy only: the host copy is valid after the next stream synchronisation
template <typename T> void sd_from_device_async(const SdMat<T>* M);

// ----------------------------------------------------------------------------
// sparse matrices
// ----------------------------------------------------------------------------
// launch settings of the SpMV for one matrix and block width: bytes each
// lane gathers per nonzero (0 = one 16-byte vector) and workgroups per CU
// (0 = as many as fit)
// the depth the SpMV runs for a requested depth (8, 16 or 32), the scalars
// per lane vpl, the scalar size and the block width: 16 and 32 need at most
// 32 gathered bytes per lane and at most 8 lanes per row; otherwise 8
inline int spmv_effective_depth(int depth, int vpl, int scalar_bytes, int ncols) {
  int groups = (std::min(ncols, 64 * vpl) + vpl - 1) / vpl, lpr = 1;
  while (lpr < groups) lpr <<= 1;
  if (lpr > 8 || vpl * scalar_bytes > 32) return 8;
  return depth >= 32 ? 32 : (depth >= 16 ? 16 : 8);
}

struct SpmvTuned {
  int lane_bytes = 0;
  int blocks_per_cu = 0;
  int depth = 8;  // entries per row whose loads are issued together (8 or 32)
  int col_block = 0;  // block columns per launch (0: as many as 64 lane groups hold)
  int stage = 0;      // x rows of each tile staged in LDS
  double ms = 0.0;  // measured time per product when chosen
};

// The x rows a workgroup tile of rpb consecutive SELL rows gathers from, as
// column intervals, for the SpMV that stages them in LDS (spmv.hip): tile t
// owns intervals [ivptr[t], ivptr[t+1]), interval k copies x rows
// iv[3k] .. iv[3k] + iv[3k+2] - 1 to window rows iv[3k+1] ..; lcol[slot] is
// the window row of the column of SELL slot `slot` (padding included).  ok
// when every tile needs at most kMaxIv intervals and kMaxRows window rows.
struct SellStage {
  static constexpr int kMaxIv = 64;
  static constexpr int kMaxRows = 65535;
  bool ok = false;
  int max_rows = 0;  // the widest window over the tiles
  int max_iv = 0;    // the most intervals of a tile
  int32_t* d_ivptr = nullptr;
  int32_t* d_iv = nullptr;
  uint16_t* d_lcol = nullptr;
};

template <typename T>
struct SparseMat {
  Map* row_map = nullptr;  // owned
  Map* col_map = nullptr;  // owned (same layout as row_map for square matrices)
  Map* range_map = nullptr;
  Map* domain_map = nullptr;
  Context ctx;
  int64_t nrows = 0, ncols = 0, nnz = 0;
  int32_t max_row_len = 0;
  // builtin-order CRS on the host (0-based row_ptr and col, each row stably
  // sorted by column); kept for index-parity checks and host-side consumers
  std::vector<int64_t> crs_ptr;
  std::vector<int32_t> crs_col;
  std::vector<T> crs_val;
  // A == A' entry by entry (-1: not yet checked; sparse_is_symmetric fills it)
  mutable int symmetric = -1;
  // SELL-C-sigma on the device
  int C = 0, sigma = 0;
  int64_t nchunks = 0;
  int64_t nnz_padded = 0;
  int64_t* d_chunk_ptr = nullptr;  // [nchunks+1] element offset of each chunk
  int32_t* d_chunk_w = nullptr;    // [nchunks] chunk width (longest row)
  int32_t* d_perm = nullptr;       // [nchunks*C] SELL row -> matrix row (-1: padding)
  int32_t* d_col = nullptr;        // [nnz_padded] column of each slot
  T* d_val = nullptr;              // [nnz_padded] value of each slot
  bool perm_identity = true;
  // host copies of the SELL index arrays (for inspection)
  std::vector<int64_t> h_chunk_ptr;
  std::vector<int32_t> h_chunk_w;
  std::vector<int32_t> h_perm;
  // FP32 copy of an FP64 matrix for mixed-precision solvers (made on first
  // use, freed with the matrix)
  SparseMat<float>* fp32_twin = nullptr;
  // SpMV launch settings the autotuner chose for this matrix, by block width
  // (tune.cpp); the SpMV uses them unless a global setting overrides them
  std::map<int, SpmvTuned> tuned;
  // LDS staging tables by tile height (sell_stage; made on first use, freed
  // with the SELL copy)
  mutable std::map<int, SellStage> stage;
};

// the staging tables of A for tiles of rpb SELL rows (built on first use:
// host work over every slot, then one upload); nullptr when some tile
// gathers from more intervals or window rows than SellStage allows, or when
// they are not built yet inside an allocation-free call or a graph capture
template <typename T> const SellStage* sell_stage(const SparseMat<T>* A, int rpb);
// LDS a staged SpMV window may take per workgroup (two workgroups per CU still fit)
constexpr int kStageLdsBytes = 64 << 10;
// the staging tables for an SpMV launch of nc columns with vpl scalars per
// lane (spmv.hip), with log2 of the 16-byte pieces per window row and the
// window's LDS bytes; nullptr where that launch cannot stage
template <typename T>
const SellStage* spmv_stage_for(const SparseMat<T>* A, int nc, int vpl, int* pshift, size_t* lds_bytes);

// the cached FP32 copy of A (same SELL layout, values rounded to float)
const SparseMat<float>* sparse_fp32_twin(const SparseMat<double>* A);

// build the SELL-C-sigma device copy of a matrix whose host CRS is filled
template <typename T> void sell_build(SparseMat<T>* A, int C, int sigma);
template <typename T> void sparse_free(SparseMat<T>* A);
// replace the device SELL copy by one with chunk height C and window sigma
// (the FP32 twin and the tuned launch settings are dropped with the old one)
template <typename T> void sell_rebuild(SparseMat<T>* A, int C, int sigma);
// SpMV autotuner (tune.cpp): see phist_amd_ext.h
template <typename T>
void sparse_tune(SparseMat<T>* A, int n_nvec, const int* nvecs, int flags, double* default_ms, double* tuned_ms,
                 int* settings);

// handle casts
template <typename T> struct Handles;
template <> struct Handles<double> {
  using mvec = phist_Dmvec; using sdMat = phist_DsdMat; using sparseMat = phist_DsparseMat;
};
template <> struct Handles<float> {
  using mvec = phist_Smvec; using sdMat = phist_SsdMat; using sparseMat = phist_SsparseMat;
};
template <typename T> inline Mvec<T>* MV(typename Handles<T>::mvec* p) { return reinterpret_cast<Mvec<T>*>(p); }
template <typename T> inline const Mvec<T>* MV(const typename Handles<T>::mvec* p) { return reinterpret_cast<const Mvec<T>*>(p); }
template <typename T> inline SdMat<T>* SD(typename Handles<T>::sdMat* p) { return reinterpret_cast<SdMat<T>*>(p); }
template <typename T> inline const SdMat<T>* SD(const typename Handles<T>::sdMat* p) { return reinterpret_cast<const SdMat<T>*>(p); }
template <typename T> inline SparseMat<T>* SP(typename Handles<T>::sparseMat* p) { return reinterpret_cast<SparseMat<T>*>(p); }
template <typename T> inline const SparseMat<T>* SP(const typename Handles<T>::sparseMat* p) { return reinterpret_cast<const SparseMat<T>*>(p); }

// ----------------------------------------------------------------------------
// device kernel launchers (spmv.hip, blas1.hip, tsgemm.hip)
// ----------------------------------------------------------------------------
struct SpmvRequest {
  bool has_shift = false, has_beta = false;
  bool dot_ww = false, dot_vw = false;
  bool axpy_u = false;  // U = gamma*W + delta*U after W is formed
  // when set, the dots stay on the device: dots_dev[0..nvec) = ww, dots_dev[nvec..2nvec) = vw
  // (both written whichever were requested; at most 64 column groups)
  void* dots_dev = nullptr;
  // when set (instead of dots_dev), the dots stay on the device in two
  // arrays of nvec entries: dot_ww_dev = ww, dot_vw_dev = vw (either may be null)
  void* dot_ww_dev = nullptr;
  void* dot_vw_dev = nullptr;
  // when set (with has_shift), the shifts are read on the device from
  // shifts_dev[0..nvec) instead of the host array
  const void* shifts_dev = nullptr;
  // when set, x is read as x(i, c) * xscale[c * xscale_stride] (a per-column
  // factor on the device, converted to the scalar type)
  const double* xscale = nullptr;
  int64_t xscale_stride = 1;
  // when >= 0, columns from ncols_active on are neither read nor written
  // (their dots are 0); the launch is the one for all columns, so the
  // other columns' results do not depend on it
  int ncols_active = -1;
  // Gram epilogue (at most 32 columns, no other epilogue): W'W and V'W of
  // the new block written to these device arrays (column-major, column
  // stride *_ld); with no_store the new block itself is not written (y may
  // then be null)
  void* gram_ww_dev = nullptr;
  int gram_ww_ld = 0;
  void* gram_vw_dev = nullptr;
  int gram_vw_ld = 0;
  // launch settings for this call only (the autotuner's trials): lane bytes
  // (>= 0) and workgroups per CU (>= 0); -1 = the usual choice
  int lane_bytes = -1;
  int blocks_per_cu = -1;
  int depth = -1;
  int col_block = -1;
  int stage = -1;
  bool no_store = false;
};

// limits of the blocked GMRES Gram-Schmidt sweeps (systems per block, basis
// vectors per system held in registers)
constexpr int kGmresCgsMaxKk = 64, kGmresCgsMaxNl = 32;

// device state of the blocked GMRES's Givens recurrences for kk systems and
// basis size m (double): H (per system (m+1) x m, column-major), cs, sn (m
// per system), g (m+1), and per system ||r0||, tol, |g_{i+1}|/||r0||, steps
// this cycle, status (1 iterating, 0 converged, 3 failed) and iterations
// before this cycle
struct GmresDevState {
  double *H, *cs, *sn, *g, *normR0, *tol, *relres;
  int *steps, *status, *iters0;
};

// per-column state of the device-resident MINRES recurrences (Paige & Saunders)
struct MinresState {
  double beta1, beta, c1, s1, c2, s2, phibar, relres, tol;
  int iters, max_iter, active, pad;
  // the Lanczos vectors are kept unnormalised (the block that received A v):
  // v = block * sc[...]; with iteration counter p, sc[p % 3] belongs to the
  // current v, sc[(p + 2) % 3] to the previous one, and the rotation of
  // iteration p writes sc[(p + 1) % 3] for the next
  double sc[3];
};

// y = alpha*(A + diag-shifts)*x + beta*y over column range of the given views;
// optional column dots of the new y (ww) and of x with the new y (vw),
// written to host arrays; optional U update (fused_spmv_mvdot_mvadd).
template <typename T>
void spmv_launch(const SparseMat<T>* A, T alpha, const T* shifts_host, const Mvec<T>* x, T beta, Mvec<T>* y,
                 const SpmvRequest& req, T* dot_ww_host, T* dot_vw_host, T gamma = T(0), T delta = T(0),
                 Mvec<T>* U = nullptr);

// BLAS-1 on blocks
template <typename T> void blas_put_value(Mvec<T>* V, T value);
template <typename T> void blas_scale_cols(Mvec<T>* V, const T* scal_host, bool uniform);
template <typename T> void blas_axpby(const T* alpha_host, bool alpha_uniform, const Mvec<T>* X, T beta, Mvec<T>* Y);
template <typename T> void blas_elemwise(T alpha, const Mvec<T>* V, Mvec<T>* W);
template <typename T> void blas_dot(const Mvec<T>* V, const Mvec<T>* W, T* out_host);
// column dots left on the device (out_dev[0..nvec)); no host synchronisation
template <typename T> void blas_dot_dev(const Mvec<T>* V, const Mvec<T>* W, T* out_dev);
// V(:, c) *= 1/sqrt(n2_dev[c]) (0 where n2_dev[c] <= 0), squared norms on the device
template <typename T> void blas_scale_inv_norm_dev(Mvec<T>* V, const T* n2_dev);
template <typename T> void blas_copy(const Mvec<T>* X, Mvec<T>* Y);

// CARP: multicoloured Kaczmarz sweeps (carp.hip)
template <typename T> void* carp_setup(const SparseMat<T>* A, int nshifts);
template <typename T> void carp_destroy(void* work);
template <typename T>
void carp_sweep(const SparseMat<T>* A, void* work, const T* sr, const T* si, const Mvec<T>* B, Mvec<T>* Xr,
                Mvec<T>* Xi, const T* omega);
template <typename T> void carp_colors(void* work, int* ncolors, int64_t* color_ptr, int32_t* rows);

// tall-skinny GEMMs; the sdMat results are written to the device copy and
// copied to the host copy before returning
// to_host = false leaves C only on the device (its host mirror is not refreshed)
template <typename T>
void gemm_gram(T alpha, const Mvec<T>* V, const Mvec<T>* W, T beta, SdMat<T>* C, bool to_host = true);
template <typename T> void gemm_times_sd(T alpha, const Mvec<T>* V, const SdMat<T>* C, T beta, Mvec<T>* W);
template <typename T> void gemm_times_sd_inplace(Mvec<T>* V, const SdMat<T>* M);
// On the device, one workgroup: G (k x k, k <= 64, symmetric positive definite)
// = R'R with R upper triangular, Rinv = R^-1 (device copy of an sdMat of the
// same shape).  *d_flag (device) becomes 1 when a pivot is not positive or
// below 64*eps times the largest diagonal entry of G (Rinv is then the
// identity), else 0.
template <typename T> void chol_inv_dev(const SdMat<T>* G, SdMat<T>* Rinv, T* d_flag);
// W -= V*C and D = V'W (new W) in one sweep where the shapes allow; D on the device only
template <typename T> void gemm_sub_gram_dev(const Mvec<T>* V, const SdMat<T>* C, Mvec<T>* W, SdMat<T>* D);
template <typename T>
void gemm_times_sd_add_times_sd(const Mvec<T>* V, const SdMat<T>* C, Mvec<T>* W, const SdMat<T>* D);
template <typename T>
void gemm_fused_mvsdi_mvTmv(T alpha, const Mvec<T>* V, Mvec<T>* W, const SdMat<T>* C, T beta, SdMat<T>* D);
template <typename T>
void gemm_fused_mvsd_mvTmv(T alpha, const Mvec<T>* V, const SdMat<T>* C, T beta, Mvec<T>* W, SdMat<T>* WtW);

// KISS random numbers (kiss.cpp)
void kiss_reset();
// draws n values of the global stream into out (double), advancing it by n
void kiss_draw(int64_t n, double* out);
void kiss_skip(uint64_t n);

}  // namespace phx
