// matfuncs.cpp -- row-function problem generators (the role of PHIST's
// drivers/matfuncs): fed to phist_[DS]sparseMat_create_fromRowFunc*.
//
//   banded random  row i holds the diagonal `diag` and nnz-1 distinct columns
//                  drawn uniformly from [max(0,i-hb), min(n-1,i+hb)] \ {i}
//                  with values U(-1,1); each row has its own SplitMix64
//                  stream seeded from (seed, i), so rows can be generated in
//                  any order or in parallel with identical results.
//   laplace3d      7-point Dirichlet Laplacian on nx*ny*nz points,
//                  6 on the diagonal, -1 for each neighbour.
#include "amd_common.hpp"

namespace {

inline uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

template <typename T>
int banded_row(phist_gidx row, phist_lidx* nnz, phist_gidx* cols, void* vals, void* arg) {
  const auto* p = static_cast<const phist_amd_banded_params*>(arg);
  T* v = static_cast<T*>(vals);
  const int64_t lo = row - p->halfband < 0 ? 0 : row - p->halfband;
  const int64_t hi = row + p->halfband > p->n - 1 ? p->n - 1 : row + p->halfband;
  const int64_t width = hi - lo + 1;  // includes the diagonal
  int want = p->nnz_per_row - 1;
  if (want > width - 1) want = (int)(width - 1);
  cols[0] = row;
  v[0] = (T)p->diag;
  int k = 1;
  uint64_t s = p->seed ^ (0xD1B54A32D192ED03ULL * (uint64_t)(row + 1));
  while (k <= want) {
    const int64_t c = lo + (int64_t)(splitmix(s) % (uint64_t)width);
    if (c == row) continue;
    bool dup = false;
    for (int j = 1; j < k; j++)
      if (cols[j] == c) {
        dup = true;
        break;
      }
    if (dup) continue;
    cols[k] = c;
    v[k] = (T)((double)(splitmix(s) >> 11) * (2.0 / 9007199254740992.0) - 1.0);
    k++;
  }
  *nnz = k;
  return 0;
}

template <typename T>
int laplace_row(phist_gidx row, phist_lidx* nnz, phist_gidx* cols, void* vals, void* arg) {
  const auto* p = static_cast<const phist_amd_laplace3d_params*>(arg);
  T* v = static_cast<T*>(vals);
  const int64_t nx = p->nx, ny = p->ny, nz = p->nz;
  const int64_t ix = row % nx, iy = (row / nx) % ny, iz = row / (nx * ny);
  int k = 0;
  auto add = [&](int64_t c, T a) {
    cols[k] = c;
    v[k] = a;
    k++;
  };
  if (iz > 0) add(row - nx * ny, T(-1));
  if (iy > 0) add(row - nx, T(-1));
  if (ix > 0) add(row - 1, T(-1));
  add(row, T(6));
  if (ix < nx - 1) add(row + 1, T(-1));
  if (iy < ny - 1) add(row + nx, T(-1));
  if (iz < nz - 1) add(row + nx * ny, T(-1));
  *nnz = k;
  return 0;
}

}  // namespace

extern "C" {

int phist_amd_banded_random_rowfunc(phist_gidx row, phist_lidx* nnz, phist_gidx* cols, void* vals, void* arg) {
  return banded_row<double>(row, nnz, cols, vals, arg);
}
int phist_amd_banded_random_rowfunc_f(phist_gidx row, phist_lidx* nnz, phist_gidx* cols, void* vals, void* arg) {
  return banded_row<float>(row, nnz, cols, vals, arg);
}
int phist_amd_laplace3d_rowfunc(phist_gidx row, phist_lidx* nnz, phist_gidx* cols, void* vals, void* arg) {
  return laplace_row<double>(row, nnz, cols, vals, arg);
}
int phist_amd_laplace3d_rowfunc_f(phist_gidx row, phist_lidx* nnz, phist_gidx* cols, void* vals, void* arg) {
  return laplace_row<float>(row, nnz, cols, vals, arg);
}
// constructor for the ...WithConstructor variants: the generators keep no
// per-thread state, the work pointer is the parameter block itself
int phist_amd_rowfunc_shared_ctor(void* arg, void** work) {
  *work = arg;
  return 0;
}

}  // extern "C"
