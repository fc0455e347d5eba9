// carp.hip -- Kaczmarz (CARP) sweeps on the device by multicolouring.
//
// PHIST's CARP kernels (phist_kernels_carp_decl.h:35-102) leave the
// parallelisation to the kernel library.  Here the rows are coloured so that
// two rows of one colour never touch the same entry of x (greedy distance-2
// colouring of the row graph, a row's support being its columns plus its own
// index, which the shift term updates); the rows of a colour then update in
// parallel, one thread per (row, column of X), and a sweep runs the colours
// forward and then backward.  Each colour's rows keep their entries in a
// chunked column-major copy (coalesced reads across the rows of a wave).
// The per-row arithmetic is that of builtin's
// kacz_sub_def.h:27-139 (shifted row product, row norm corrected for the
// shift, scaled projection with relaxation omega_j), with fused multiply-adds
// in a fixed place so that a sequential sweep in colour order reproduces it
// exactly.  RC: real matrix, complex shifts sigma_r + i sigma_i and complex
// solution X_r + i X_i (kacz_sub_def.h under KACZ_RC_VARIANT).
#include "amd_device.hpp"
#include <algorithm>
#include <vector>

namespace phx {

// Rows of one colour in chunks of kCarpChunk (in colour order), each chunk's
// entries stored column-major (entry k of chunk row rr at chunk_ptr + k*kCarpChunk
// + rr, padded to the chunk's longest row), so the rows a wave updates at once
// read their matrix entries in coalesced segments.
constexpr int kCarpChunk = 32;

template <typename T>
struct CarpWork {
  int64_t n = 0;
  int nshifts = 0;
  int ncolors = 0;
  std::vector<int64_t> color_ptr;    // rows of colour c: rows[color_ptr[c] .. color_ptr[c+1])
  std::vector<int64_t> color_chunk;  // first chunk of colour c
  std::vector<int32_t> h_rows;
  int32_t* rows = nullptr;       // device: rows in colour order
  int32_t* rlen = nullptr;       // device: their lengths
  int64_t* chunk_ptr = nullptr;  // device: first slot of every chunk
  int32_t* col = nullptr;        // device: chunked entries
  T* val = nullptr;
};

// shifts and relaxation factors of up to kCarpCols columns, passed by value
constexpr int kCarpCols = 64;
template <typename T>
struct CarpCoef {
  T sr[kCarpCols], si[kCarpCols], w[kCarpCols];
};

template <typename T, bool RC>
__global__ void __launch_bounds__(kBlock)
    k_kacz(const int32_t* __restrict__ rows, const int32_t* __restrict__ rlen, const int64_t* __restrict__ cptr,
           int64_t nr, const int32_t* __restrict__ ci, const T* __restrict__ va, int nvec, const CarpCoef<T> coef,
           const T* __restrict__ b, int64_t ldb, T* xr, T* xi, int64_t ldx) {
  const int64_t total = nr * nvec;
  for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < total; e += (int64_t)gridDim.x * kBlock) {
    const int64_t r = e / nvec;
    const int c = (int)(e - r * nvec);
    const int64_t i = rows[r];
    const int len = rlen[r];
    const int64_t k0 = cptr[r / kCarpChunk] + (r % kCarpChunk);
    const T s_r = coef.sr[c], s_i = RC ? coef.si[c] : T(0), w = coef.w[c];
    const T xri = xr[i * ldx + c];
    T tr, ti = T(0);
    if (RC) {
      const T xii = xi[i * ldx + c];
      tr = fma(s_i, xii, -s_r * xri);
      ti = fma(-s_i, xri, -s_r * xii);
    } else {
      tr = -s_r * xri;
    }
    T d = T(0), nrm = T(0);
    for (int q = 0; q < len; q++) {
      const int64_t k = k0 + (int64_t)q * kCarpChunk;
      const int64_t j = ci[k];
      const T a = va[k];
      tr = fma(a, xr[j * ldx + c], tr);
      if (RC) ti = fma(a, xi[j * ldx + c], ti);
      nrm = fma(a, a, nrm);
      if (j == i) d = a;
    }
    if (b) tr = tr - b[i * ldb + c];
    nrm = fma(s_r, s_r, fma(T(-2) * d, s_r, nrm));
    if (RC) nrm = fma(s_i, s_i, nrm);
    if (nrm == T(0)) nrm = T(1);
    tr = (tr * w) / nrm;
    if (RC) ti = (ti * w) / nrm;
    // projection: the shift term of x_i first, then every column of the row
    if (RC) {
      const T xii = xi[i * ldx + c];
      xr[i * ldx + c] = xri + fma(ti, s_i, tr * s_r);
      xi[i * ldx + c] = xii + fma(-tr, s_i, ti * s_r);
    } else {
      xr[i * ldx + c] = fma(tr, s_r, xri);
    }
    for (int q = 0; q < len; q++) {
      const int64_t k = k0 + (int64_t)q * kCarpChunk;
      const int64_t j = ci[k];
      const T a = va[k];
      xr[j * ldx + c] = fma(-tr, a, xr[j * ldx + c]);
      if (RC) xi[j * ldx + c] = fma(-ti, a, xi[j * ldx + c]);
    }
  }
}

// greedy distance-2 colouring in row order: a row takes the smallest colour
// not used by any row sharing one of its columns (its own index included)
static void color_rows(const std::vector<int64_t>& ptr, const std::vector<int32_t>& col, int64_t n,
                       std::vector<int>& color, int& ncolors) {
  // transpose structure: rows touching each column (own index included)
  std::vector<int64_t> cp(n + 1, 0);
  for (int64_t i = 0; i < n; i++) {
    cp[i + 1]++;
    for (int64_t k = ptr[i]; k < ptr[i + 1]; k++)
      if (col[k] != i) cp[(int64_t)col[k] + 1]++;
  }
  for (int64_t i = 0; i < n; i++) cp[i + 1] += cp[i];
  std::vector<int64_t> fill(cp.begin(), cp.end() - 1);
  std::vector<int32_t> rr(cp[n]);
  for (int64_t i = 0; i < n; i++) {
    rr[fill[i]++] = (int32_t)i;
    for (int64_t k = ptr[i]; k < ptr[i + 1]; k++)
      if (col[k] != i) rr[fill[col[k]]++] = (int32_t)i;
  }
  color.assign(n, -1);
  std::vector<int64_t> mark;  // mark[c] == i: colour c is taken next to row i
  ncolors = 0;
  auto forbid_column = [&](int64_t j, int64_t i) {
    for (int64_t q = cp[j]; q < cp[j + 1]; q++) {
      const int c = color[rr[q]];
      if (c >= 0) mark[c] = i;
    }
  };
  for (int64_t i = 0; i < n; i++) {
    forbid_column(i, i);
    for (int64_t k = ptr[i]; k < ptr[i + 1]; k++)
      if (col[k] != i) forbid_column(col[k], i);
    int c = 0;
    while (c < ncolors && mark[c] == i) c++;
    if (c == ncolors) {
      ncolors++;
      mark.push_back(-1);
    }
    color[i] = c;
  }
}

template <typename T>
void* carp_setup(const SparseMat<T>* A, int nshifts) {
  PHX_REQUIRE(A->nrows == A->ncols, PHIST_INVALID_INPUT, "CARP needs a square matrix");
  PHX_REQUIRE(nshifts >= 1, PHIST_INVALID_INPUT, "numShifts %d", nshifts);
  auto* W = new CarpWork<T>();
  W->n = A->nrows;
  W->nshifts = nshifts;
  std::vector<int> color;
  color_rows(A->crs_ptr, A->crs_col, W->n, color, W->ncolors);
  W->color_ptr.assign(W->ncolors + 1, 0);
  for (int64_t i = 0; i < W->n; i++) W->color_ptr[color[i] + 1]++;
  for (int c = 0; c < W->ncolors; c++) W->color_ptr[c + 1] += W->color_ptr[c];
  W->h_rows.resize(W->n);
  std::vector<int64_t> fill(W->color_ptr.begin(), W->color_ptr.end() - 1);
  for (int64_t i = 0; i < W->n; i++) W->h_rows[fill[color[i]]++] = (int32_t)i;
  // chunked entry layout, colour by colour (entries of a row in CRS order)
  std::vector<int32_t> rlen(W->n);
  std::vector<int64_t> cptr;
  W->color_chunk.assign(W->ncolors + 1, 0);
  int64_t slots = 0;
  for (int c = 0; c < W->ncolors; c++) {
    W->color_chunk[c] = (int64_t)cptr.size();
    for (int64_t r0 = W->color_ptr[c]; r0 < W->color_ptr[c + 1]; r0 += kCarpChunk) {
      const int64_t r1 = std::min<int64_t>(W->color_ptr[c + 1], r0 + kCarpChunk);
      int32_t wmax = 0;
      for (int64_t r = r0; r < r1; r++) {
        const int64_t i = W->h_rows[r];
        rlen[r] = (int32_t)(A->crs_ptr[i + 1] - A->crs_ptr[i]);
        wmax = std::max(wmax, rlen[r]);
      }
      cptr.push_back(slots);
      slots += (int64_t)wmax * kCarpChunk;
    }
  }
  W->color_chunk[W->ncolors] = (int64_t)cptr.size();
  cptr.push_back(slots);
  std::vector<int32_t> ccol((size_t)std::max<int64_t>(1, slots), 0);
  std::vector<T> cval((size_t)std::max<int64_t>(1, slots), T(0));
  for (int c = 0; c < W->ncolors; c++)
    for (int64_t r = W->color_ptr[c]; r < W->color_ptr[c + 1]; r++) {
      const int64_t rl = r - W->color_ptr[c];
      const int64_t base = cptr[W->color_chunk[c] + rl / kCarpChunk] + rl % kCarpChunk;
      const int64_t i = W->h_rows[r];
      for (int32_t q = 0; q < rlen[r]; q++) {
        ccol[base + (int64_t)q * kCarpChunk] = A->crs_col[A->crs_ptr[i] + q];
        cval[base + (int64_t)q * kCarpChunk] = A->crs_val[A->crs_ptr[i] + q];
      }
    }
  hipStream_t st = rt().stream;
  PHX_HIP(hipMalloc(&W->rows, sizeof(int32_t) * std::max<int64_t>(1, W->n)));
  PHX_HIP(hipMalloc(&W->rlen, sizeof(int32_t) * std::max<int64_t>(1, W->n)));
  PHX_HIP(hipMalloc(&W->chunk_ptr, sizeof(int64_t) * cptr.size()));
  PHX_HIP(hipMalloc(&W->col, sizeof(int32_t) * ccol.size()));
  PHX_HIP(hipMalloc(&W->val, sizeof(T) * cval.size()));
  if (W->n) {
    PHX_HIP(hipMemcpyAsync(W->rows, W->h_rows.data(), sizeof(int32_t) * W->n, hipMemcpyHostToDevice, st));
    PHX_HIP(hipMemcpyAsync(W->rlen, rlen.data(), sizeof(int32_t) * W->n, hipMemcpyHostToDevice, st));
  }
  PHX_HIP(hipMemcpyAsync(W->chunk_ptr, cptr.data(), sizeof(int64_t) * cptr.size(), hipMemcpyHostToDevice, st));
  PHX_HIP(hipMemcpyAsync(W->col, ccol.data(), sizeof(int32_t) * ccol.size(), hipMemcpyHostToDevice, st));
  PHX_HIP(hipMemcpyAsync(W->val, cval.data(), sizeof(T) * cval.size(), hipMemcpyHostToDevice, st));
  sync_stream();  // the host staging vectors die here
  return W;
}

template <typename T>
void carp_destroy(void* work) {
  auto* W = static_cast<CarpWork<T>*>(work);
  if (!W) return;
  sync_stream();
  if (W->rows) PHX_HIP(hipFree(W->rows));
  if (W->rlen) PHX_HIP(hipFree(W->rlen));
  if (W->chunk_ptr) PHX_HIP(hipFree(W->chunk_ptr));
  if (W->col) PHX_HIP(hipFree(W->col));
  if (W->val) PHX_HIP(hipFree(W->val));
  delete W;
}

// forward then backward sweep over the colours; B may be null (b = 0),
// si and Xi null for real shifts
template <typename T>
void carp_sweep(const SparseMat<T>* A, void* work, const T* sr, const T* si, const Mvec<T>* B, Mvec<T>* Xr,
                Mvec<T>* Xi, const T* omega) {
  auto* W = static_cast<CarpWork<T>*>(work);
  PHX_REQUIRE(W, PHIST_INVALID_INPUT, "CARP work object is NULL (call carp_setup first)");
  PHX_REQUIRE(W->n == A->nrows, PHIST_INVALID_INPUT, "CARP work object belongs to another matrix");
  const int nvec = Xr->nvec;
  PHX_REQUIRE(sr && omega, PHIST_INVALID_INPUT, "sigma and omega must hold one entry per column of X");
  PHX_REQUIRE(Xr->nrows == W->n && (!B || (B->nrows == W->n && B->nvec == nvec)), PHIST_INVALID_INPUT,
              "vector shapes do not match the matrix");
  PHX_REQUIRE(!Xi || (Xi->nrows == W->n && Xi->nvec == nvec && Xi->lda == Xr->lda), PHIST_INVALID_INPUT,
              "X_i must match X_r");
  hipStream_t st = rt().stream;
  ktimer_begin();  // whole sweep (all colour launches) when the kernel timer is on
  // columns are independent: blocks of up to kCarpCols columns sweep one after another
  for (int c0 = 0; c0 < nvec; c0 += kCarpCols) {
    const int nc = std::min(kCarpCols, nvec - c0);
    CarpCoef<T> coef{};
    for (int c = 0; c < nc; c++) {
      coef.sr[c] = sr[c0 + c];
      coef.si[c] = si ? si[c0 + c] : T(0);
      coef.w[c] = omega[c0 + c];
    }
    const T* bp = B ? B->dev() + c0 : nullptr;
    const int64_t ldb = B ? B->lda : 0;
    T* xr = Xr->dev() + c0;
    T* xi = Xi ? Xi->dev() + c0 : nullptr;
    auto run = [&](int c) {
      const int64_t r0 = W->color_ptr[c], nr = W->color_ptr[c + 1] - r0;
      if (nr == 0) return;
      const int64_t total = nr * nc;
      const int grid = (int)std::min<int64_t>((total + kBlock - 1) / kBlock, (int64_t)rt().num_cu * 8);
      const int64_t* cp = W->chunk_ptr + W->color_chunk[c];
      if (xi)
        hipLaunchKernelGGL((k_kacz<T, true>), dim3(grid), dim3(kBlock), 0, st, W->rows + r0, W->rlen + r0, cp, nr,
                           W->col, W->val, nc, coef, bp, ldb, xr, xi, Xr->lda);
      else
        hipLaunchKernelGGL((k_kacz<T, false>), dim3(grid), dim3(kBlock), 0, st, W->rows + r0, W->rlen + r0, cp, nr,
                           W->col, W->val, nc, coef, bp, ldb, xr, (T*)nullptr, Xr->lda);
      PHX_HIP(hipGetLastError());
    };
    for (int c = 0; c < W->ncolors; c++) run(c);
    for (int c = W->ncolors - 1; c >= 0; c--) run(c);
  }
  ktimer_end();
}

template <typename T>
void carp_colors(void* work, int* ncolors, int64_t* color_ptr, int32_t* rows) {
  auto* W = static_cast<CarpWork<T>*>(work);
  PHX_REQUIRE(W, PHIST_INVALID_INPUT, "NULL work object");
  *ncolors = W->ncolors;
  if (color_ptr) std::copy(W->color_ptr.begin(), W->color_ptr.end(), color_ptr);
  if (rows) std::copy(W->h_rows.begin(), W->h_rows.end(), rows);
}

#define PHX_INST(T)                                                                                           \
  template void* carp_setup<T>(const SparseMat<T>*, int);                                                     \
  template void carp_destroy<T>(void*);                                                                       \
  template void carp_sweep<T>(const SparseMat<T>*, void*, const T*, const T*, const Mvec<T>*, Mvec<T>*, Mvec<T>*, \
                              const T*);                                                                      \
  template void carp_colors<T>(void*, int*, int64_t*, int32_t*);
PHX_INST(double)
PHX_INST(float)
#undef PHX_INST

}  // namespace phx
