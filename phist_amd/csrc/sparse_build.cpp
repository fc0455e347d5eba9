// sparse_build.cpp -- host-side index work for sparse matrices.
//
// 1. The builtin-order CRS: rows are collected in row order, each row's
//    entries stably sorted by column (crsmat_module.F90:2058-2312 fills rows
//    in order and calls sort_global_cols :484-516, a stable merge ranking; on
//    one process sort_rows_local_nonlocal :563-630 re-sorts the same keys and
//    keeps that order).  MatrixMarket input buckets entries by row in file
//    order, mirroring symmetric off-diagonal entries directly after the
//    original (read_mm_data, crsmat_module.F90:1760-1830), then sorts rows
//    the same way.  Explicit zeros and duplicate columns are kept, as builtin
//    keeps them.
// 2. SELL-C-sigma for the device: rows are stably sorted by descending length
//    inside windows of sigma rows, grouped into chunks of C rows, each chunk
//    padded to its longest row and stored column-major (slot (r, j) of chunk c
//    at chunk_ptr[c] + j*C + r).  Padding slots hold value 0 and point at the
//    row's own index (or 0 if that is not a valid column), so they never
//    address outside x.  The row permutation stays inside the matrix: block
//    vectors keep their natural order, y is written through perm.
// 3. LDS staging tables (sell_stage): for tiles of consecutive SELL rows, the
//    sorted columns the tile gathers from (padding slots included), merged
//    into intervals across gaps of at most kStageGap rows, and each slot's
//    row in the tile's window of intervals laid end to end.
#include "amd_common.hpp"
#include <algorithm>
#include <atomic>
#include <mutex>
#include <numeric>
#include <omp.h>

namespace phx {

namespace {
// stable insertion sort for short rows, std::stable_sort for long ones
template <typename T>
void sort_row(int64_t n, int64_t* col, T* val, std::vector<int64_t>& idx, std::vector<int64_t>& tc, std::vector<T>& tv) {
  if (n <= 1) return;
  if (n <= 64) {
    for (int64_t i = 1; i < n; i++) {
      int64_t c = col[i];
      T v = val[i];
      int64_t j = i - 1;
      while (j >= 0 && col[j] > c) {
        col[j + 1] = col[j];
        val[j + 1] = val[j];
        j--;
      }
      col[j + 1] = c;
      val[j + 1] = v;
    }
    return;
  }
  idx.resize(n);
  tc.resize(n);
  tv.resize(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return col[a] < col[b]; });
  for (int64_t k = 0; k < n; k++) {
    tc[k] = col[idx[k]];
    tv[k] = val[idx[k]];
  }
  std::copy(tc.begin(), tc.begin() + n, col);
  std::copy(tv.begin(), tv.begin() + n, val);
}
}  // namespace

// rows given in fill order as (ptr, int64 cols, vals); produces the sorted CRS
template <typename T>
void crs_finish(SparseMat<T>* A, std::vector<int64_t>& ptr, std::vector<int64_t>& gcol, std::vector<T>& gval) {
  const int64_t n = A->nrows;
  for (int64_t k = 0; k < (int64_t)gcol.size(); k++)
    PHX_REQUIRE(gcol[k] >= 0 && gcol[k] < A->ncols, PHIST_INVALID_INPUT,
                "column index %lld outside [0,%lld)", (long long)gcol[k], (long long)A->ncols);
#pragma omp parallel
  {
    std::vector<int64_t> idx, tc;
    std::vector<T> tv;
#pragma omp for schedule(dynamic, 4096)
    for (int64_t i = 0; i < n; i++) sort_row<T>(ptr[i + 1] - ptr[i], gcol.data() + ptr[i], gval.data() + ptr[i], idx, tc, tv);
  }
  A->nnz = ptr[n];
  A->crs_ptr.swap(ptr);
  A->crs_col.resize(A->nnz);
  int32_t maxlen = 0;
#pragma omp parallel for reduction(max : maxlen) schedule(static)
  for (int64_t i = 0; i < n; i++) {
    maxlen = std::max<int32_t>(maxlen, (int32_t)(A->crs_ptr[i + 1] - A->crs_ptr[i]));
    for (int64_t k = A->crs_ptr[i]; k < A->crs_ptr[i + 1]; k++) A->crs_col[k] = (int32_t)gcol[k];
  }
  A->max_row_len = maxlen;
  A->crs_val.swap(gval);
  std::vector<int64_t>().swap(gcol);
}

template <typename T>
void sell_build(SparseMat<T>* A, int C, int sigma) {
  Runtime& r = rt();
  if (C <= 0) C = r.sell_C > 0 ? r.sell_C : 32;
  if (sigma <= 0) sigma = r.sell_sigma;
  PHX_REQUIRE((C & (C - 1)) == 0 && C <= 1024, PHIST_INVALID_INPUT, "SELL chunk height %d is not a power of two", C);
  const int64_t n = A->nrows;
  std::vector<int32_t> len(n);
  bool uniform = true;
  for (int64_t i = 0; i < n; i++) {
    len[i] = (int32_t)(A->crs_ptr[i + 1] - A->crs_ptr[i]);
    if (i && len[i] != len[0]) uniform = false;
  }
  if (sigma <= 0) {
    // rows in their own order keep the locality of the matrix's numbering
    // (neighbouring rows gather neighbouring x rows); sort by length only
    // when that order pads the chunks by more than a quarter
    int64_t slots = 0;
    for (int64_t c0 = 0; c0 < n; c0 += C) {
      int32_t w = 0;
      for (int64_t i = c0; i < std::min<int64_t>(n, c0 + C); i++) w = std::max(w, len[i]);
      slots += (int64_t)w * C;
    }
    sigma = (double)slots <= 1.25 * (double)std::max<int64_t>(1, A->nnz) ? 1 : 256;
  }
  const int64_t nchunks = (n + C - 1) / C;
  const int64_t nslots = nchunks * C;
  std::vector<int32_t> perm(nslots, -1);
  std::iota(perm.begin(), perm.begin() + n, 0);
  if (sigma > 1 && !uniform) {
    for (int64_t w0 = 0; w0 < n; w0 += sigma) {
      int64_t w1 = std::min<int64_t>(n, w0 + sigma);
      std::stable_sort(perm.begin() + w0, perm.begin() + w1, [&](int32_t a, int32_t b) { return len[a] > len[b]; });
    }
  }
  bool ident = true;
  for (int64_t s = 0; s < n; s++)
    if (perm[s] != s) { ident = false; break; }
  std::vector<int32_t> cw(nchunks, 0);
  std::vector<int64_t> cp(nchunks + 1, 0);
  for (int64_t c = 0; c < nchunks; c++) {
    int32_t w = 0;
    for (int rr = 0; rr < C; rr++) {
      int32_t row = perm[c * C + rr];
      if (row >= 0) w = std::max(w, len[row]);
    }
    cw[c] = w;
    cp[c + 1] = cp[c] + (int64_t)w * C;
  }
  const int64_t nslot_el = cp[nchunks];
  std::vector<int32_t> col(nslot_el);
  std::vector<T> val(nslot_el);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t c = 0; c < nchunks; c++) {
    for (int rr = 0; rr < C; rr++) {
      int32_t row = perm[c * C + rr];
      int32_t pad_col = (row >= 0 && row < A->ncols) ? row : 0;
      int64_t p0 = row >= 0 ? A->crs_ptr[row] : 0;
      int32_t l = row >= 0 ? len[row] : 0;
      for (int32_t j = 0; j < cw[c]; j++) {
        int64_t slot = cp[c] + (int64_t)j * C + rr;
        if (j < l) {
          col[slot] = A->crs_col[p0 + j];
          val[slot] = A->crs_val[p0 + j];
        } else {
          col[slot] = pad_col;
          val[slot] = T(0);
        }
      }
    }
  }
  A->C = C;
  A->sigma = (sigma > 1 && !uniform) ? sigma : 1;
  A->nchunks = nchunks;
  A->nnz_padded = nslot_el;
  A->perm_identity = ident;
  hipStream_t st = r.stream;
  PHX_HIP(hipMalloc(&A->d_chunk_ptr, (nchunks + 1) * sizeof(int64_t)));
  PHX_HIP(hipMalloc(&A->d_chunk_w, std::max<int64_t>(1, nchunks) * sizeof(int32_t)));
  PHX_HIP(hipMalloc(&A->d_perm, std::max<int64_t>(1, nslots) * sizeof(int32_t)));
  PHX_HIP(hipMalloc(&A->d_col, std::max<int64_t>(1, nslot_el) * sizeof(int32_t)));
  PHX_HIP(hipMalloc(&A->d_val, std::max<int64_t>(1, nslot_el) * sizeof(T)));
  PHX_HIP(hipMemcpyAsync(A->d_chunk_ptr, cp.data(), (nchunks + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
  if (nchunks) PHX_HIP(hipMemcpyAsync(A->d_chunk_w, cw.data(), nchunks * sizeof(int32_t), hipMemcpyHostToDevice, st));
  if (nslots) PHX_HIP(hipMemcpyAsync(A->d_perm, perm.data(), nslots * sizeof(int32_t), hipMemcpyHostToDevice, st));
  if (nslot_el) {
    PHX_HIP(hipMemcpyAsync(A->d_col, col.data(), nslot_el * sizeof(int32_t), hipMemcpyHostToDevice, st));
    PHX_HIP(hipMemcpyAsync(A->d_val, val.data(), nslot_el * sizeof(T), hipMemcpyHostToDevice, st));
  }
  PHX_HIP(hipStreamSynchronize(st));  // host staging vectors die here
  A->h_chunk_ptr.swap(cp);
  A->h_chunk_w.swap(cw);
  A->h_perm.swap(perm);
}

namespace {
// unused rows between two gathered columns that one interval still spans
constexpr int32_t kStageGap = 4;

template <typename T>
void stage_free(const SparseMat<T>* A) {
  for (auto& kv : A->stage) {
    for (void* p : {(void*)kv.second.d_ivptr, (void*)kv.second.d_iv, (void*)kv.second.d_lcol})
      if (p) PHX_HIP(hipFree(p));
  }
  A->stage.clear();
}
}  // namespace

template <typename T>
const SellStage* sell_stage(const SparseMat<T>* A, int rpb) {
  static std::mutex mu;
  std::lock_guard<std::mutex> lock(mu);
  auto found = A->stage.find(rpb);
  if (found != A->stage.end()) return found->second.ok ? &found->second : nullptr;
  // building the tables allocates device memory: not inside an allocation-free call or a graph capture
  if (rt().fixed_work || rt().capturing) return nullptr;
  SellStage& S = A->stage[rpb];  // not ok until the tables are up
  const int64_t nslots = A->nchunks * A->C;
  const int64_t ntiles = rpb > 0 ? (nslots + rpb - 1) / rpb : 0;
  if (ntiles == 0 || A->nnz_padded == 0 || A->crs_ptr.empty() || ntiles >= INT32_MAX) return nullptr;
  const int C = A->C;
  const int64_t* cp = A->h_chunk_ptr.data();
  const int32_t* cw = A->h_chunk_w.data();
  const int32_t* perm = A->h_perm.data();
  const int64_t* rp = A->crs_ptr.data();
  const int32_t* ci = A->crs_col.data();
  std::vector<std::vector<int32_t>> tiv(ntiles);
  std::vector<uint16_t> lcol(A->nnz_padded, 0);
  std::atomic<bool> fail{false};
  int max_rows = 0, max_iv = 0;
#pragma omp parallel
  {
    std::vector<int32_t> cols;
    int my_rows = 0, my_iv = 0;
#pragma omp for schedule(dynamic, 64)
    for (int64_t t = 0; t < ntiles; t++) {
      if (fail.load(std::memory_order_relaxed)) continue;
      const int64_t s0 = t * rpb, s1 = std::min(nslots, s0 + rpb);
      // the column of entry j of SELL row s (the row's own index, or 0, for padding)
      auto col_of = [&](int64_t s, int32_t j) -> int32_t {
        const int32_t row = perm[s];
        const int64_t len = row >= 0 ? rp[row + 1] - rp[row] : 0;
        if (j < len) return ci[rp[row] + j];
        return (row >= 0 && row < A->ncols) ? row : 0;
      };
      cols.clear();
      for (int64_t sr = s0; sr < s1; sr++) {
        const int32_t w = cw[sr / C];
        for (int32_t j = 0; j < w; j++) cols.push_back(col_of(sr, j));
      }
      std::sort(cols.begin(), cols.end());
      cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
      std::vector<int32_t>& iv = tiv[t];  // (source column, window row, length) per interval
      int32_t rows = 0;
      for (int32_t c : cols) {
        if (!iv.empty() && c <= iv[iv.size() - 3] + iv.back() - 1 + kStageGap) {
          const int32_t grow = c - (iv[iv.size() - 3] + iv.back() - 1);
          iv.back() += grow;
          rows += grow;
        } else {
          iv.insert(iv.end(), {c, rows, 1});
          rows += 1;
        }
        if (rows > SellStage::kMaxRows || (int)iv.size() / 3 > SellStage::kMaxIv) break;
      }
      const int niv = (int)iv.size() / 3;
      if (rows > SellStage::kMaxRows || niv > SellStage::kMaxIv) {
        fail.store(true, std::memory_order_relaxed);
        continue;
      }
      my_rows = std::max(my_rows, rows);
      my_iv = std::max(my_iv, niv);
      for (int64_t sr = s0; sr < s1; sr++) {
        const int64_t ch = sr / C;
        const int r = (int)(sr % C);
        for (int32_t j = 0; j < cw[ch]; j++) {
          const int32_t c = col_of(sr, j);
          int lo = 0, hi = niv - 1;  // the last interval starting at or before c
          while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (iv[3 * mid] <= c) lo = mid; else hi = mid - 1;
          }
          lcol[cp[ch] + (int64_t)j * C + r] = (uint16_t)(iv[3 * lo + 1] + (c - iv[3 * lo]));
        }
      }
    }
#pragma omp critical
    {
      max_rows = std::max(max_rows, my_rows);
      max_iv = std::max(max_iv, my_iv);
    }
  }
  if (fail.load()) return nullptr;
  std::vector<int32_t> ivptr(ntiles + 1, 0), ivs;
  for (int64_t t = 0; t < ntiles; t++) ivptr[t + 1] = ivptr[t] + (int32_t)(tiv[t].size() / 3);
  ivs.reserve((size_t)ivptr[ntiles] * 3);
  for (auto& v : tiv) ivs.insert(ivs.end(), v.begin(), v.end());
  hipStream_t st = rt().stream;
  PHX_HIP(hipMalloc(&S.d_ivptr, ivptr.size() * sizeof(int32_t)));
  PHX_HIP(hipMalloc(&S.d_iv, std::max<size_t>(1, ivs.size()) * sizeof(int32_t)));
  PHX_HIP(hipMalloc(&S.d_lcol, lcol.size() * sizeof(uint16_t)));
  PHX_HIP(hipMemcpyAsync(S.d_ivptr, ivptr.data(), ivptr.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
  if (!ivs.empty())
    PHX_HIP(hipMemcpyAsync(S.d_iv, ivs.data(), ivs.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
  PHX_HIP(hipMemcpyAsync(S.d_lcol, lcol.data(), lcol.size() * sizeof(uint16_t), hipMemcpyHostToDevice, st));
  PHX_HIP(hipStreamSynchronize(st));  // host vectors die here
  S.max_rows = max_rows;
  S.max_iv = max_iv;
  S.ok = true;
  return &S;
}

template <typename T>
void sell_rebuild(SparseMat<T>* A, int C, int sigma) {
  sync_stream();
  stage_free(A);
  if (A->fp32_twin) {  // made again from the new layout on first use
    sparse_free(A->fp32_twin);
    A->fp32_twin = nullptr;
  }
  for (void* p : {(void*)A->d_chunk_ptr, (void*)A->d_chunk_w, (void*)A->d_perm, (void*)A->d_col, (void*)A->d_val})
    if (p) PHX_HIP(hipFree(p));
  A->d_chunk_ptr = nullptr;
  A->d_chunk_w = nullptr;
  A->d_perm = nullptr;
  A->d_col = nullptr;
  A->d_val = nullptr;
  A->tuned.clear();  // launch settings belong to a layout
  sell_build(A, C, sigma);
}

template <typename T>
void sparse_free(SparseMat<T>* A) {
  if (!A) return;
  sync_stream();
  if (A->fp32_twin) sparse_free(A->fp32_twin);
  if (A->d_chunk_ptr) PHX_HIP(hipFree(A->d_chunk_ptr));
  if (A->d_chunk_w) PHX_HIP(hipFree(A->d_chunk_w));
  if (A->d_perm) PHX_HIP(hipFree(A->d_perm));
  if (A->d_col) PHX_HIP(hipFree(A->d_col));
  if (A->d_val) PHX_HIP(hipFree(A->d_val));
  stage_free(A);
  if (A->col_map && A->col_map != A->row_map) delete A->col_map;
  if (A->domain_map && A->domain_map != A->row_map && A->domain_map != A->col_map) delete A->domain_map;
  delete A->row_map;
  delete A;
}

const SparseMat<float>* sparse_fp32_twin(const SparseMat<double>* A) {
  if (A->fp32_twin) return A->fp32_twin;
  auto* F = new SparseMat<float>();
  F->nrows = A->nrows;
  F->ncols = A->ncols;
  F->nnz = A->nnz;
  F->max_row_len = A->max_row_len;
  F->row_map = new Map(*A->row_map);
  F->col_map = A->col_map == A->row_map ? F->row_map : new Map(*A->col_map);
  F->range_map = F->row_map;
  F->domain_map = F->col_map;
  F->ctx.row_map = F->row_map;
  F->ctx.col_map = F->col_map;
  F->ctx.range_map = F->range_map;
  F->ctx.domain_map = F->domain_map;
  F->crs_ptr = A->crs_ptr;
  F->crs_col = A->crs_col;
  F->crs_val.resize(A->crs_val.size());
  for (size_t i = 0; i < A->crs_val.size(); i++) F->crs_val[i] = (float)A->crs_val[i];
  sell_build(F, A->C, A->sigma);
  const_cast<SparseMat<double>*>(A)->fp32_twin = F;
  return F;
}

// ---- construction entry points ----
namespace {
template <typename T>
SparseMat<T>* new_matrix(int64_t nrows, int64_t ncols, const Comm* comm) {
  PHX_REQUIRE(nrows >= 0 && nrows <= INT32_MAX && ncols >= 0 && ncols <= INT32_MAX, PHIST_INVALID_INPUT,
              "matrix dimensions %lld x %lld out of range", (long long)nrows, (long long)ncols);
  auto* A = new SparseMat<T>();
  A->nrows = nrows;
  A->ncols = ncols;
  A->row_map = new Map{comm ? comm : default_comm(), nrows, (int32_t)nrows, 0};
  A->col_map = ncols == nrows ? A->row_map : new Map{A->row_map->comm, ncols, (int32_t)ncols, 0};
  A->range_map = A->row_map;
  A->domain_map = A->col_map;
  A->ctx.row_map = A->row_map;
  A->ctx.col_map = A->col_map;
  A->ctx.range_map = A->range_map;
  A->ctx.domain_map = A->domain_map;
  return A;
}
}  // namespace

template <typename T>
SparseMat<T>* sparse_from_rowfunc(const Comm* comm, int64_t nrows, int64_t ncols, int32_t maxnne,
                                  phist_sparseMat_rowFunc f, phist_sparseMat_rowFuncConstructor ctor, void* arg,
                                  bool parallel) {
  ensure_init();
  PHX_REQUIRE(f, PHIST_INVALID_INPUT, "NULL row function");
  PHX_REQUIRE(maxnne >= 0, PHIST_INVALID_INPUT, "negative maxnne");
  SparseMat<T>* A = new_matrix<T>(nrows, ncols, comm);
  try {
    std::vector<int64_t> ptr(nrows + 1, 0);
    std::vector<int64_t> gcol((size_t)nrows * maxnne);
    std::vector<T> gval((size_t)nrows * maxnne);
    std::vector<int32_t> cnt(nrows, 0);
    int bad = 0;
    // rows write into fixed slots of width maxnne, then are compacted
    auto fill = [&](int64_t i, void* work) {
      phist_lidx nne = 0;
      f((phist_gidx)i, &nne, (phist_gidx*)(gcol.data() + (size_t)i * maxnne), (void*)(gval.data() + (size_t)i * maxnne),
        work);
      if (nne < 0 || nne > maxnne) bad = 1;
      cnt[i] = nne;
    };
    if (parallel && ctor) {
#pragma omp parallel
      {
        void* work = nullptr;
        ctor(arg, &work);
#pragma omp for schedule(static)
        for (int64_t i = 0; i < nrows; i++) fill(i, work);
        ctor(arg, &work);
      }
    } else {
      void* work = arg;
      if (ctor) {
        work = nullptr;
        ctor(arg, &work);
      }
      for (int64_t i = 0; i < nrows; i++) fill(i, work);
      if (ctor) ctor(arg, &work);
    }
    PHX_REQUIRE(!bad, PHIST_INVALID_INPUT, "row function returned more than maxnne=%d entries", maxnne);
    for (int64_t i = 0; i < nrows; i++) ptr[i + 1] = ptr[i] + cnt[i];
    if (ptr[nrows] != (int64_t)nrows * maxnne) {  // compact in place (rows only move left)
      for (int64_t i = 0; i < nrows; i++) {
        const size_t src = (size_t)i * maxnne;
        for (int32_t k = 0; k < cnt[i]; k++) {
          gcol[ptr[i] + k] = gcol[src + k];
          gval[ptr[i] + k] = gval[src + k];
        }
      }
      gcol.resize(ptr[nrows]);
      gval.resize(ptr[nrows]);
    }
    crs_finish(A, ptr, gcol, gval);
    sell_build(A, 0, 0);
  } catch (...) {
    sparse_free(A);
    throw;
  }
  return A;
}

template <typename T>
SparseMat<T>* sparse_from_mm(const Comm* comm, const char* filename) {
  ensure_init();
  PHX_REQUIRE(filename, PHIST_INVALID_INPUT, "NULL file name");
  FILE* fp = fopen(filename, "r");
  PHX_REQUIRE(fp, PHIST_INVALID_INPUT, "cannot open %s", filename);
  std::string text;
  {
    char buf[1 << 16];
    size_t k;
    while ((k = fread(buf, 1, sizeof buf, fp)) > 0) text.append(buf, k);
    fclose(fp);
  }
  const char* p = text.c_str();
  const char* end = p + text.size();
  auto next_line = [&](const char*& q) {
    while (q < end && *q != '\n') q++;
    if (q < end) q++;
  };
  bool symmetric;
  if (text.compare(0, 45, "%%MatrixMarket matrix coordinate real general") == 0)
    symmetric = false;
  else if (text.compare(0, 47, "%%MatrixMarket matrix coordinate real symmetric") == 0)
    symmetric = true;
  else
    fail(PHIST_INVALID_INPUT, "%s: unsupported MatrixMarket header", filename);
  next_line(p);
  while (p < end && *p == '%') next_line(p);
  char* e = nullptr;
  long long nr = strtoll(p, &e, 10), nc = strtoll(e, &e, 10), nl = strtoll(e, &e, 10);
  PHX_REQUIRE(e != p && nr >= 0 && nc >= 0 && nl >= 0, PHIST_INVALID_INPUT, "%s: bad size line", filename);
  p = e;
  SparseMat<T>* A = new_matrix<T>(nr, nc, comm);
  try {
    std::vector<int64_t> ri, ci;
    std::vector<T> vv;
    ri.reserve(symmetric ? 2 * nl : nl);
    ci.reserve(ri.capacity());
    vv.reserve(ri.capacity());
    for (long long k = 0; k < nl; k++) {
      long long i = strtoll(p, &e, 10);
      long long j = strtoll(e, &e, 10);
      double a = strtod(e, &e);
      PHX_REQUIRE(e != p, PHIST_INVALID_INPUT, "%s: truncated entry list", filename);
      p = e;
      PHX_REQUIRE(i >= 1 && i <= nr && j >= 1 && j <= nc, PHIST_INVALID_INPUT, "%s: entry (%lld,%lld) out of range",
                  filename, i, j);
      ri.push_back(i - 1);
      ci.push_back(j - 1);
      vv.push_back((T)a);
      if (symmetric && i != j) {
        ri.push_back(j - 1);
        ci.push_back(i - 1);
        vv.push_back((T)a);
      }
    }
    const int64_t ne = (int64_t)ri.size();
    std::vector<int64_t> ptr(nr + 1, 0);
    for (int64_t k = 0; k < ne; k++) ptr[ri[k] + 1]++;
    for (int64_t i = 0; i < nr; i++) ptr[i + 1] += ptr[i];
    std::vector<int64_t> fillp(ptr.begin(), ptr.end() - 1);
    std::vector<int64_t> gcol(ne);
    std::vector<T> gval(ne);
    for (int64_t k = 0; k < ne; k++) {
      int64_t q = fillp[ri[k]]++;
      gcol[q] = ci[k];
      gval[q] = vv[k];
    }
    crs_finish(A, ptr, gcol, gval);
    sell_build(A, 0, 0);
  } catch (...) {
    sparse_free(A);
    throw;
  }
  return A;
}

template void sell_build<double>(SparseMat<double>*, int, int);
template void sell_build<float>(SparseMat<float>*, int, int);
template const SellStage* sell_stage<double>(const SparseMat<double>*, int);
template const SellStage* sell_stage<float>(const SparseMat<float>*, int);
template void sell_rebuild<double>(SparseMat<double>*, int, int);
template void sell_rebuild<float>(SparseMat<float>*, int, int);
template void sparse_free<double>(SparseMat<double>*);
template void sparse_free<float>(SparseMat<float>*);
template SparseMat<double>* sparse_from_rowfunc<double>(const Comm*, int64_t, int64_t, int32_t, phist_sparseMat_rowFunc,
                                                        phist_sparseMat_rowFuncConstructor, void*, bool);
template SparseMat<float>* sparse_from_rowfunc<float>(const Comm*, int64_t, int64_t, int32_t, phist_sparseMat_rowFunc,
                                                      phist_sparseMat_rowFuncConstructor, void*, bool);
template SparseMat<double>* sparse_from_mm<double>(const Comm*, const char*);
template SparseMat<float>* sparse_from_mm<float>(const Comm*, const char*);

}  // namespace phx

// ---- index inspection (phist_amd_ext.h) ----
namespace {
template <typename T>
void get_crs(const phx::SparseMat<T>* A, int64_t* row_offset, int32_t* col_idx, T* val) {
  using namespace phx;
  PHX_REQUIRE(A, PHIST_INVALID_INPUT, "NULL matrix");
  if (row_offset)
    for (int64_t i = 0; i <= A->nrows; i++) row_offset[i] = A->crs_ptr[i] + 1;
  if (col_idx)
    for (int64_t k = 0; k < A->nnz; k++) col_idx[k] = A->crs_col[k] + 1;
  if (val)
    for (int64_t k = 0; k < A->nnz; k++) val[k] = A->crs_val[k];
}
template <typename T>
void get_sell(const phx::SparseMat<T>* A, int* C, int* sigma, int64_t* nchunks, int64_t* nslots, int64_t* chunk_ptr,
              int32_t* chunk_w, int32_t* perm, int32_t* col, T* val) {
  using namespace phx;
  PHX_REQUIRE(A, PHIST_INVALID_INPUT, "NULL matrix");
  if (C) *C = A->C;
  if (sigma) *sigma = A->sigma;
  if (nchunks) *nchunks = A->nchunks;
  if (nslots) *nslots = A->nnz_padded;
  if (chunk_ptr) memcpy(chunk_ptr, A->h_chunk_ptr.data(), sizeof(int64_t) * (A->nchunks + 1));
  if (chunk_w && A->nchunks) memcpy(chunk_w, A->h_chunk_w.data(), sizeof(int32_t) * A->nchunks);
  if (perm && A->nchunks) memcpy(perm, A->h_perm.data(), sizeof(int32_t) * A->nchunks * A->C);
  if (col && A->nnz_padded)
    PHX_HIP(hipMemcpy(col, A->d_col, sizeof(int32_t) * A->nnz_padded, hipMemcpyDeviceToHost));
  if (val && A->nnz_padded) PHX_HIP(hipMemcpy(val, A->d_val, sizeof(T) * A->nnz_padded, hipMemcpyDeviceToHost));
}
}  // namespace

extern "C" {
void phist_Damd_sparseMat_get_crs(phist_Dconst_sparseMat_ptr A, int64_t* row_offset, int32_t* col_idx, double* val,
                                  int* iflag) {
  phx::guard(iflag, __func__, [&] {
    get_crs(phx::SP<double>(A), row_offset, col_idx, val);
    *iflag = 0;
  });
}
void phist_Samd_sparseMat_get_crs(phist_Sconst_sparseMat_ptr A, int64_t* row_offset, int32_t* col_idx, float* val,
                                  int* iflag) {
  phx::guard(iflag, __func__, [&] {
    get_crs(phx::SP<float>(A), row_offset, col_idx, val);
    *iflag = 0;
  });
}
void phist_Damd_sparseMat_get_sell(phist_Dconst_sparseMat_ptr A, int* C, int* sigma, int64_t* nchunks,
                                   int64_t* nslots, int64_t* chunk_ptr, int32_t* chunk_w, int32_t* perm,
                                   int32_t* col, double* val, int* iflag) {
  phx::guard(iflag, __func__, [&] {
    get_sell(phx::SP<double>(A), C, sigma, nchunks, nslots, chunk_ptr, chunk_w, perm, col, val);
    *iflag = 0;
  });
}
void phist_Samd_sparseMat_get_sell(phist_Sconst_sparseMat_ptr A, int* C, int* sigma, int64_t* nchunks,
                                   int64_t* nslots, int64_t* chunk_ptr, int32_t* chunk_w, int32_t* perm,
                                   int32_t* col, float* val, int* iflag) {
  phx::guard(iflag, __func__, [&] {
    get_sell(phx::SP<float>(A), C, sigma, nchunks, nslots, chunk_ptr, chunk_w, perm, col, val);
    *iflag = 0;
  });
}
}  // extern "C"
