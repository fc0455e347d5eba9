// objects.cpp -- block vectors and small dense matrices: lifetime, views,
// host/device mirrors, data access, random and functional fills, printing and
// raw binary I/O.
//
// Reference semantics: src/kernels/phist_kernels_decl.h:166-486,
// src/kernels/builtin/mvec_module.F90 (views :3409-3467 -- a non-NULL
// *Vblock must itself be a view and is re-pointed; jmin < 0 or jmax < jmin is
// PHIST_BAD_CAST), src/kernels/builtin/sdmat_module.F90, and the common
// fills in src/kernels/common/kernels_common_impl_def.h:256-405.
#include "objects.hpp"
#include <algorithm>
#include <cmath>

namespace phx {

// ---------------------------------------------------------------------------
// block vectors
// ---------------------------------------------------------------------------
template <typename T>
T* Mvec<T>::host() {
  Mvec* r = root;
  if (!r->host_base) {
    size_t n = (size_t)std::max<int64_t>(1, r->nrows * r->lda);
    r->host_base = static_cast<T*>(pinned_alloc(n * sizeof(T)));
    memset(r->host_base, 0, n * sizeof(T));
  }
  return r->host_base + jofs;
}

namespace api {

template <typename T>
Mvec<T>* mvec_create(const Map* map, int nvec) {
  ensure_init();
  PHX_REQUIRE(map, PHIST_INVALID_INPUT, "NULL map");
  PHX_REQUIRE(nvec > 0, PHIST_INVALID_INPUT, "mvec needs at least one column (got %d)", nvec);
  auto* V = new Mvec<T>();
  V->map = map;
  V->nrows = map->nloc;
  V->nvec = nvec;
  V->lda = nvec;
  const size_t bytes = sizeof(T) * (size_t)std::max<int64_t>(1, V->nrows * nvec);
  try {
    V->dev_base = static_cast<T*>(dev_alloc(bytes));
  } catch (...) {
    delete V;
    throw;
  }
  PHX_HIP(hipMemsetAsync(V->dev_base, 0, bytes, rt().stream));
  return V;
}

template <typename T>
void mvec_delete(Mvec<T>* V) {
  if (!V) return;
  if (V->is_view()) {
    V->root->nviews--;
    delete V;
    return;
  }
  const size_t n = (size_t)std::max<int64_t>(1, V->nrows * V->lda);
  if (V->dev_base) dev_release(V->dev_base, n * sizeof(T));
  if (V->host_base) pinned_release(V->host_base, n * sizeof(T));
  delete V;
}

template <typename T>
void mvec_view_block(Mvec<T>* V, Mvec<T>** Vblock, int jmin, int jmax) {
  PHX_REQUIRE(V, PHIST_BAD_CAST, "NULL mvec");
  PHX_REQUIRE(jmin >= 0 && jmax >= jmin, PHIST_BAD_CAST, "view with jmin=%d, jmax=%d", jmin, jmax);
  PHX_REQUIRE(jmax < V->nvec, PHIST_INVALID_INPUT, "view columns %d..%d outside 0..%d", jmin, jmax, V->nvec - 1);
  Mvec<T>* view = *Vblock;
  if (view) {
    PHX_REQUIRE(view->is_view(), PHIST_BAD_CAST, "*Vblock is not a view");
    view->root->nviews--;
  } else {
    view = new Mvec<T>();
  }
  view->map = V->map;
  view->nrows = V->nrows;
  view->nvec = jmax - jmin + 1;
  view->root = V->root;
  view->jofs = V->jofs + jmin;
  view->lda = V->root->lda;
  view->root->nviews++;
  *Vblock = view;
}

template <typename T>
void mvec_get_block(const Mvec<T>* V, Mvec<T>* B, int jmin, int jmax) {
  PHX_REQUIRE(V && B, PHIST_BAD_CAST, "NULL mvec");
  PHX_REQUIRE(jmin >= 0 && jmax >= jmin && jmax < V->nvec, PHIST_INVALID_INPUT, "bad column range %d..%d", jmin, jmax);
  PHX_REQUIRE(B->nvec == jmax - jmin + 1 && B->nrows == V->nrows, PHIST_INVALID_INPUT, "block has wrong shape");
  Mvec<T> src = *V;
  src.root = V->root;
  src.jofs = V->jofs + jmin;
  src.nvec = jmax - jmin + 1;
  blas_copy<T>(&src, B);
}

template <typename T>
void mvec_set_block(Mvec<T>* V, const Mvec<T>* B, int jmin, int jmax) {
  PHX_REQUIRE(V && B, PHIST_BAD_CAST, "NULL mvec");
  PHX_REQUIRE(jmin >= 0 && jmax >= jmin && jmax < V->nvec, PHIST_INVALID_INPUT, "bad column range %d..%d", jmin, jmax);
  PHX_REQUIRE(B->nvec == jmax - jmin + 1 && B->nrows == V->nrows, PHIST_INVALID_INPUT, "block has wrong shape");
  Mvec<T> dst = *V;
  dst.root = V->root;
  dst.jofs = V->jofs + jmin;
  dst.nvec = jmax - jmin + 1;
  blas_copy<T>(B, &dst);
}

// Row-block copies between host memory (row pitch hp bytes) and device
// memory (row pitch dp bytes), `rows` rows of `width` bytes.  Contiguous
// blocks go as one linear copy; a strided side is packed on its own memory
// (host loop, or a device-to-device 2D copy through a staging buffer) so the
// host <-> device transfer is always linear: a 2D copy from host memory with
// millions of short rows is far slower than the link.
void copy_rows_h2d(void* dev, size_t dp, const void* host, size_t hp, size_t width, size_t rows) {
  hipStream_t st = rt().stream;
  const char* src = static_cast<const char*>(host);
  std::vector<char> pack;
  if (hp != width) {
    pack.resize(width * rows);
    for (size_t r = 0; r < rows; r++) memcpy(pack.data() + r * width, src + r * hp, width);
    src = pack.data();
  }
  if (dp == width) {
    PHX_HIP(hipMemcpyAsync(dev, src, width * rows, hipMemcpyHostToDevice, st));
  } else {
    void* stage = device_work(width * rows);
    PHX_HIP(hipMemcpyAsync(stage, src, width * rows, hipMemcpyHostToDevice, st));
    PHX_HIP(hipMemcpy2DAsync(dev, dp, stage, width, width, rows, hipMemcpyDeviceToDevice, st));
  }
  sync_stream();
}

void copy_rows_d2h(void* host, size_t hp, const void* dev, size_t dp, size_t width, size_t rows) {
  hipStream_t st = rt().stream;
  const void* src = dev;
  if (dp != width) {
    void* stage = device_work(width * rows);
    PHX_HIP(hipMemcpy2DAsync(stage, width, dev, dp, width, rows, hipMemcpyDeviceToDevice, st));
    src = stage;
  }
  if (hp == width) {
    if (!fetch_to_host(host, src, width * rows)) {
      PHX_HIP(hipMemcpyAsync(host, src, width * rows, hipMemcpyDeviceToHost, st));
      sync_stream();
    }
  } else {
    std::vector<char> pack(width * rows);
    if (!fetch_to_host(pack.data(), src, width * rows)) {
      PHX_HIP(hipMemcpyAsync(pack.data(), src, width * rows, hipMemcpyDeviceToHost, st));
      sync_stream();
    }
    char* dst = static_cast<char*>(host);
    for (size_t r = 0; r < rows; r++) memcpy(dst + r * hp, pack.data() + r * width, width);
  }
}

// copies between the device block and the host mirror (same layout)
template <typename T>
void mvec_to_device(Mvec<T>* V) {
  if (V->nrows == 0) return;
  T* h = V->host();
  copy_rows_h2d(V->dev(), V->lda * sizeof(T), h, V->lda * sizeof(T), V->nvec * sizeof(T), V->nrows);
}

template <typename T>
void mvec_from_device(Mvec<T>* V) {
  if (V->nrows == 0) return;
  T* h = V->host();
  copy_rows_d2h(h, V->lda * sizeof(T), V->dev(), V->lda * sizeof(T), V->nvec * sizeof(T), V->nrows);
}

template <typename T>
void mvec_get_data(const Mvec<T>* V, T* data, int64_t lda, int row_major) {
  PHX_REQUIRE(data, PHIST_INVALID_INPUT, "NULL data");
  PHX_REQUIRE(lda >= (row_major ? V->nvec : V->nrows), PHIST_INVALID_INPUT, "lda too small");
  if (V->nrows == 0) return;
  if (row_major) {
    copy_rows_d2h(data, lda * sizeof(T), V->dev(), V->lda * sizeof(T), V->nvec * sizeof(T), V->nrows);
  } else {
    std::vector<T> tmp((size_t)V->nrows * V->nvec);
    copy_rows_d2h(tmp.data(), V->nvec * sizeof(T), V->dev(), V->lda * sizeof(T), V->nvec * sizeof(T), V->nrows);
    for (int64_t i = 0; i < V->nrows; i++)
      for (int j = 0; j < V->nvec; j++) data[(int64_t)j * lda + i] = tmp[(size_t)i * V->nvec + j];
  }
}

template <typename T>
void mvec_set_data(Mvec<T>* V, const T* data, int64_t lda, int row_major) {
  PHX_REQUIRE(data, PHIST_INVALID_INPUT, "NULL data");
  PHX_REQUIRE(lda >= (row_major ? V->nvec : V->nrows), PHIST_INVALID_INPUT, "lda too small");
  if (V->nrows == 0) return;
  if (row_major) {
    copy_rows_h2d(V->dev(), V->lda * sizeof(T), data, lda * sizeof(T), V->nvec * sizeof(T), V->nrows);
  } else {
    std::vector<T> tmp((size_t)V->nrows * V->nvec);
    for (int64_t i = 0; i < V->nrows; i++)
      for (int j = 0; j < V->nvec; j++) tmp[(size_t)i * V->nvec + j] = data[(int64_t)j * lda + i];
    copy_rows_h2d(V->dev(), V->lda * sizeof(T), tmp.data(), V->nvec * sizeof(T), V->nvec * sizeof(T), V->nrows);
  }
}

// The element values are staged in a packed host array and uploaded: a column
// view of a wide block does not need the root's pinned host mirror (which
// for a basis of many columns is gigabytes of page-locked memory).
template <typename T>
void mvec_put_func(Mvec<T>* V, phist_mvec_elemFunc f, void* arg) {
  PHX_REQUIRE(f, PHIST_INVALID_INPUT, "NULL element function");
  if (V->nrows == 0) return;
  std::vector<T> h((size_t)V->nrows * V->nvec);
  const int64_t il = V->map->ilower;
  for (int64_t i = 0; i < V->nrows; i++)
    for (int j = 0; j < V->nvec; j++) f((phist_gidx)(il + i), (phist_lidx)j, (void*)&h[(size_t)i * V->nvec + j], arg);
  copy_rows_h2d(V->dev(), V->lda * sizeof(T), h.data(), V->nvec * sizeof(T), V->nvec * sizeof(T), V->nrows);
}

// Draws nrows*nvec values of the global stream (row-major) and then skips
// (gnrows - iupper)*nvec = nvec values, as kernels_common_impl_def.h:256-361
// does on a single process (pre_skip = ilower*nvec = 0).
// The values are uploaded from a packed host array (no root host mirror).
template <typename T>
void mvec_random(Mvec<T>* V) {
  const int64_t n = V->nrows * V->nvec;
  std::vector<double> r((size_t)std::max<int64_t>(1, n));
  kiss_draw(n, r.data());
  const int64_t iupper = V->map->ilower + V->map->nloc - 1;
  kiss_skip((uint64_t)((V->map->nglob - iupper) * V->nvec));
  if (n == 0) return;
  std::vector<T> h((size_t)n);
#pragma omp parallel for schedule(static) if (n > (1 << 16))
  for (int64_t e = 0; e < n; e++) h[(size_t)e] = (T)r[(size_t)e];
  copy_rows_h2d(V->dev(), V->lda * sizeof(T), h.data(), V->nvec * sizeof(T), V->nvec * sizeof(T), V->nrows);
}

template <typename T>
void mvec_print(const Mvec<T>* V) {
  std::vector<T> d((size_t)V->nrows * V->nvec);
  mvec_get_data(V, d.data(), V->nvec, 1);
  printf("mvec %lld x %d\n", (long long)V->nrows, V->nvec);
  for (int64_t i = 0; i < V->nrows; i++) {
    for (int j = 0; j < V->nvec; j++) printf(" %24.16e", (double)d[(size_t)i * V->nvec + j]);
    printf("\n");
  }
  fflush(stdout);
}

template <typename T>
void mvec_write_bin(const Mvec<T>* V, const char* fn) {
  FILE* f = fopen(fn, "wb");
  PHX_REQUIRE(f, PHIST_INVALID_INPUT, "cannot open %s", fn);
  std::vector<T> d((size_t)V->nrows * V->nvec);
  mvec_get_data(V, d.data(), V->nvec, 1);
  int64_t hdr[2] = {V->nrows, V->nvec};
  bool ok = fwrite(hdr, sizeof hdr, 1, f) == 1 && (d.empty() || fwrite(d.data(), sizeof(T), d.size(), f) == d.size());
  fclose(f);
  PHX_REQUIRE(ok, PHIST_FUNCTIONAL_ERROR, "write to %s failed", fn);
}

template <typename T>
void mvec_read_bin(Mvec<T>* V, const char* fn) {
  FILE* f = fopen(fn, "rb");
  PHX_REQUIRE(f, PHIST_INVALID_INPUT, "cannot open %s", fn);
  int64_t hdr[2] = {0, 0};
  bool ok = fread(hdr, sizeof hdr, 1, f) == 1 && hdr[0] == V->nrows && hdr[1] == V->nvec;
  std::vector<T> d((size_t)V->nrows * V->nvec);
  ok = ok && (d.empty() || fread(d.data(), sizeof(T), d.size(), f) == d.size());
  fclose(f);
  PHX_REQUIRE(ok, PHIST_INVALID_INPUT, "%s does not hold a %lld x %d block", fn, (long long)V->nrows, V->nvec);
  mvec_set_data(V, d.data(), V->nvec, 1);
}

// ---------------------------------------------------------------------------
// small dense matrices
// ---------------------------------------------------------------------------
template <typename T>
SdMat<T>* sdMat_create(int nrows, int ncols) {
  ensure_init();
  PHX_REQUIRE(nrows >= 0 && ncols >= 0, PHIST_INVALID_INPUT, "negative sdMat size");
  auto* M = new SdMat<T>();
  M->nrows = nrows;
  M->ncols = ncols;
  M->ld = std::max(1, nrows);
  const size_t n = (size_t)M->ld * std::max(1, ncols);
  M->host_base = static_cast<T*>(pinned_alloc(n * sizeof(T)));
  memset(M->host_base, 0, n * sizeof(T));
  M->dev_base = static_cast<T*>(dev_alloc(n * sizeof(T)));
  PHX_HIP(hipMemsetAsync(M->dev_base, 0, n * sizeof(T), rt().stream));
  return M;
}

template <typename T>
void sdMat_delete(SdMat<T>* M) {
  if (!M) return;
  if (M->is_view()) {
    M->root->nviews--;
    delete M;
    return;
  }
  const size_t n = (size_t)M->ld * std::max(1, M->ncols);
  if (M->dev_base) dev_release(M->dev_base, n * sizeof(T));
  if (M->host_base) pinned_release(M->host_base, n * sizeof(T));
  delete M;
}

template <typename T>
void sdMat_view_block(SdMat<T>* M, SdMat<T>** Mb, int imin, int imax, int jmin, int jmax) {
  PHX_REQUIRE(M, PHIST_BAD_CAST, "NULL sdMat");
  PHX_REQUIRE(imin >= 0 && imax >= imin && jmin >= 0 && jmax >= jmin, PHIST_BAD_CAST, "bad view range");
  PHX_REQUIRE(imax < M->nrows && jmax < M->ncols, PHIST_INVALID_INPUT, "view outside the matrix");
  SdMat<T>* v = *Mb;
  if (v) {
    PHX_REQUIRE(v->is_view(), PHIST_BAD_CAST, "*Mblock is not a view");
    v->root->nviews--;
  } else {
    v = new SdMat<T>();
  }
  v->nrows = imax - imin + 1;
  v->ncols = jmax - jmin + 1;
  v->root = M->root;
  v->ld = M->root->ld;
  v->iofs = M->iofs + imin;
  v->jofs = M->jofs + jmin;
  v->root->nviews++;
  *Mb = v;
}

}  // namespace api

template <typename T>
void sd_to_device(const SdMat<T>* M) {
  if (M->nrows == 0 || M->ncols == 0) return;
  const int ld = M->root->ld;
  const size_t pitch = ld * sizeof(T);
  if (stage_upload_2d(M->dev(), pitch, M->host(), pitch, M->nrows * sizeof(T), M->ncols)) return;
  // larger than a staging slot: copy from the host array and wait for it
  PHX_HIP(hipMemcpy2DAsync(M->dev(), pitch, M->host(), pitch, M->nrows * sizeof(T), M->ncols, hipMemcpyHostToDevice,
                           rt().stream));
  sync_stream();
}

template <typename T>
void sd_from_device(const SdMat<T>* M) {
  if (M->nrows == 0 || M->ncols == 0) return;
  // the column span [first entry, last entry] through the published-result buffer when it fits
  const int ld = M->root->ld;
  const size_t span = (size_t)ld * (M->ncols - 1) + M->nrows;
  if (M->nrows == ld) {
    if (fetch_to_host(M->host(), M->dev(), sizeof(T) * span)) return;
  } else if (sizeof(T) * span <= Runtime::kSignalBytes - Runtime::kSignalData - 2 * kSignalSlotBytes) {
    std::vector<T> tmp(span);
    if (fetch_to_host(tmp.data(), M->dev(), sizeof(T) * span)) {
      for (int c = 0; c < M->ncols; c++)
        memcpy(M->host() + (size_t)c * ld, tmp.data() + (size_t)c * ld, sizeof(T) * M->nrows);
      return;
    }
  }
  sd_from_device_async(M);
  sync_stream();
}

template <typename T>
void sd_from_device_async(const SdMat<T>* M) {
  if (M->nrows == 0 || M->ncols == 0) return;
  const int ld = M->root->ld;
  if (M->nrows == ld)  // whole columns: one contiguous copy
    PHX_HIP(hipMemcpyAsync(M->host(), M->dev(), sizeof(T) * ((size_t)ld * (M->ncols - 1) + M->nrows),
                           hipMemcpyDeviceToHost, rt().stream));
  else
    PHX_HIP(hipMemcpy2DAsync(M->host(), ld * sizeof(T), M->dev(), ld * sizeof(T), M->nrows * sizeof(T), M->ncols,
                             hipMemcpyDeviceToHost, rt().stream));
}

namespace api {

template <typename T>
void sdMat_get_block(const SdMat<T>* M, SdMat<T>* B, int imin, int imax, int jmin, int jmax) {
  PHX_REQUIRE(M && B, PHIST_BAD_CAST, "NULL sdMat");
  PHX_REQUIRE(imin >= 0 && imax < M->nrows && jmin >= 0 && jmax < M->ncols && imax >= imin && jmax >= jmin,
              PHIST_INVALID_INPUT, "bad block range");
  PHX_REQUIRE(B->nrows == imax - imin + 1 && B->ncols == jmax - jmin + 1, PHIST_INVALID_INPUT, "block has wrong shape");
  for (int j = jmin; j <= jmax; j++)
    for (int i = imin; i <= imax; i++) B->h(i - imin, j - jmin) = M->h(i, j);
  sd_to_device(B);
}

template <typename T>
void sdMat_set_block(SdMat<T>* M, const SdMat<T>* B, int imin, int imax, int jmin, int jmax) {
  PHX_REQUIRE(M && B, PHIST_BAD_CAST, "NULL sdMat");
  PHX_REQUIRE(imin >= 0 && imax < M->nrows && jmin >= 0 && jmax < M->ncols && imax >= imin && jmax >= jmin,
              PHIST_INVALID_INPUT, "bad block range");
  PHX_REQUIRE(B->nrows == imax - imin + 1 && B->ncols == jmax - jmin + 1, PHIST_INVALID_INPUT, "block has wrong shape");
  for (int j = jmin; j <= jmax; j++)
    for (int i = imin; i <= imax; i++) M->h(i, j) = B->h(i - imin, j - jmin);
  sd_to_device(M);
}

template <typename T>
void sdMat_set_data(SdMat<T>* M, const T* data, int64_t lda, int row_major) {
  PHX_REQUIRE(data, PHIST_INVALID_INPUT, "NULL data");
  for (int j = 0; j < M->ncols; j++)
    for (int i = 0; i < M->nrows; i++) M->h(i, j) = row_major ? data[i * lda + j] : data[(int64_t)j * lda + i];
  sd_to_device(M);
}

template <typename T>
void sdMat_get_data(const SdMat<T>* M, T* data, int64_t lda, int row_major) {
  PHX_REQUIRE(data, PHIST_INVALID_INPUT, "NULL data");
  for (int j = 0; j < M->ncols; j++)
    for (int i = 0; i < M->nrows; i++) (row_major ? data[i * lda + j] : data[(int64_t)j * lda + i]) = M->h(i, j);
}

template <typename T>
void sdMat_put_value(SdMat<T>* M, T v) {
  for (int j = 0; j < M->ncols; j++)
    for (int i = 0; i < M->nrows; i++) M->h(i, j) = v;
  sd_to_device(M);
}

template <typename T>
void sdMat_identity(SdMat<T>* M) {
  for (int j = 0; j < M->ncols; j++)
    for (int i = 0; i < M->nrows; i++) M->h(i, j) = i == j ? T(1) : T(0);
  sd_to_device(M);
}

// column-major fill from the global stream (kernels_common_impl_def.h:365-405)
template <typename T>
void sdMat_random(SdMat<T>* M) {
  const int64_t n = (int64_t)M->nrows * M->ncols;
  std::vector<double> r((size_t)std::max<int64_t>(1, n));
  kiss_draw(n, r.data());
  for (int j = 0; j < M->ncols; j++)
    for (int i = 0; i < M->nrows; i++) M->h(i, j) = (T)r[(size_t)j * M->nrows + i];
  sd_to_device(M);
}

template <typename T>
void sdMat_print(const SdMat<T>* M) {
  printf("sdMat %d x %d\n", M->nrows, M->ncols);
  for (int i = 0; i < M->nrows; i++) {
    for (int j = 0; j < M->ncols; j++) printf(" %24.16e", (double)M->h(i, j));
    printf("\n");
  }
  fflush(stdout);
}

template <typename T>
void sdMat_write_bin(const SdMat<T>* M, const char* fn) {
  FILE* f = fopen(fn, "wb");
  PHX_REQUIRE(f, PHIST_INVALID_INPUT, "cannot open %s", fn);
  int32_t hdr[2] = {M->nrows, M->ncols};
  bool ok = fwrite(hdr, sizeof hdr, 1, f) == 1;
  for (int j = 0; j < M->ncols && ok; j++)
    for (int i = 0; i < M->nrows && ok; i++) ok = fwrite(&M->h(i, j), sizeof(T), 1, f) == 1;
  fclose(f);
  PHX_REQUIRE(ok, PHIST_FUNCTIONAL_ERROR, "write to %s failed", fn);
}

template <typename T>
void sdMat_read_bin(SdMat<T>* M, const char* fn) {
  FILE* f = fopen(fn, "rb");
  PHX_REQUIRE(f, PHIST_INVALID_INPUT, "cannot open %s", fn);
  int32_t hdr[2] = {0, 0};
  bool ok = fread(hdr, sizeof hdr, 1, f) == 1 && hdr[0] == M->nrows && hdr[1] == M->ncols;
  for (int j = 0; j < M->ncols && ok; j++)
    for (int i = 0; i < M->nrows && ok; i++) ok = fread(&M->h(i, j), sizeof(T), 1, f) == 1;
  fclose(f);
  PHX_REQUIRE(ok, PHIST_INVALID_INPUT, "%s does not hold a %d x %d matrix", fn, M->nrows, M->ncols);
  sd_to_device(M);
}

// C = alpha*op(A)*op(B) + beta*C on the host copies (k x k algebra)
template <typename T>
void sdMat_gemm(T alpha, const SdMat<T>* A, bool ta, const SdMat<T>* B, bool tb, T beta, SdMat<T>* C) {
  const int m = ta ? A->ncols : A->nrows, k = ta ? A->nrows : A->ncols;
  const int kb = tb ? B->ncols : B->nrows, n = tb ? B->nrows : B->ncols;
  PHX_REQUIRE(k == kb && C->nrows == m && C->ncols == n, PHIST_INVALID_INPUT, "sdMat product shapes do not match");
  std::vector<T> out((size_t)m * n);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < m; i++) {
      T s = T(0);
      for (int l = 0; l < k; l++) s += (ta ? A->h(l, i) : A->h(i, l)) * (tb ? B->h(j, l) : B->h(l, j));
      out[(size_t)j * m + i] = s;
    }
  for (int j = 0; j < n; j++)
    for (int i = 0; i < m; i++)
      C->h(i, j) = beta == T(0) ? alpha * out[(size_t)j * m + i] : alpha * out[(size_t)j * m + i] + beta * C->h(i, j);
  sd_to_device(C);
}

template <typename T>
void sdMat_add(T alpha, const SdMat<T>* A, bool ta, T beta, SdMat<T>* B) {
  const int m = ta ? A->ncols : A->nrows, n = ta ? A->nrows : A->ncols;
  PHX_REQUIRE(B->nrows == m && B->ncols == n, PHIST_INVALID_INPUT, "sdMat shapes do not match");
  std::vector<T> a((size_t)m * n);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < m; i++) a[(size_t)j * m + i] = ta ? A->h(j, i) : A->h(i, j);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < m; i++)
      B->h(i, j) = beta == T(0) ? alpha * a[(size_t)j * m + i] : alpha * a[(size_t)j * m + i] + beta * B->h(i, j);
  sd_to_device(B);
}

}  // namespace api

#define PHX_INST(T)                                                                      \
  template T* Mvec<T>::host();                                                           \
  template void sd_to_device<T>(const SdMat<T>*);                                        \
  template void sd_from_device<T>(const SdMat<T>*);                                      \
  template void sd_from_device_async<T>(const SdMat<T>*);                                \
  namespace api {                                                                        \
  template Mvec<T>* mvec_create<T>(const Map*, int);                                     \
  template void mvec_delete<T>(Mvec<T>*);                                                \
  template void mvec_view_block<T>(Mvec<T>*, Mvec<T>**, int, int);                       \
  template void mvec_get_block<T>(const Mvec<T>*, Mvec<T>*, int, int);                   \
  template void mvec_set_block<T>(Mvec<T>*, const Mvec<T>*, int, int);                   \
  template void mvec_to_device<T>(Mvec<T>*);                                             \
  template void mvec_from_device<T>(Mvec<T>*);                                           \
  template void mvec_get_data<T>(const Mvec<T>*, T*, int64_t, int);                      \
  template void mvec_set_data<T>(Mvec<T>*, const T*, int64_t, int);                      \
  template void mvec_put_func<T>(Mvec<T>*, phist_mvec_elemFunc, void*);                  \
  template void mvec_random<T>(Mvec<T>*);                                                \
  template void mvec_print<T>(const Mvec<T>*);                                           \
  template void mvec_write_bin<T>(const Mvec<T>*, const char*);                          \
  template void mvec_read_bin<T>(Mvec<T>*, const char*);                                 \
  template SdMat<T>* sdMat_create<T>(int, int);                                          \
  template void sdMat_delete<T>(SdMat<T>*);                                              \
  template void sdMat_view_block<T>(SdMat<T>*, SdMat<T>**, int, int, int, int);          \
  template void sdMat_get_block<T>(const SdMat<T>*, SdMat<T>*, int, int, int, int);      \
  template void sdMat_set_block<T>(SdMat<T>*, const SdMat<T>*, int, int, int, int);      \
  template void sdMat_set_data<T>(SdMat<T>*, const T*, int64_t, int);                    \
  template void sdMat_get_data<T>(const SdMat<T>*, T*, int64_t, int);                    \
  template void sdMat_put_value<T>(SdMat<T>*, T);                                        \
  template void sdMat_identity<T>(SdMat<T>*);                                            \
  template void sdMat_random<T>(SdMat<T>*);                                              \
  template void sdMat_print<T>(const SdMat<T>*);                                         \
  template void sdMat_write_bin<T>(const SdMat<T>*, const char*);                        \
  template void sdMat_read_bin<T>(SdMat<T>*, const char*);                               \
  template void sdMat_gemm<T>(T, const SdMat<T>*, bool, const SdMat<T>*, bool, T, SdMat<T>*); \
  template void sdMat_add<T>(T, const SdMat<T>*, bool, T, SdMat<T>*);                    \
  }
PHX_INST(double)
PHX_INST(float)
#undef PHX_INST

}  // namespace phx
