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
//   spin chain S_z the Heisenberg/XXZ chain of L spins restricted to nup up
//                  spins (drivers/matfuncs/SpinChainSZ.c:189-310): basis =
//                  the L-bit words with nup ones in ascending order, row r
//                  <-> r-th word; diagonal jz * sum over bonds of (+1 for
//                  parallel, -1 for antiparallel spins), omitted when 0; one
//                  entry 2*jxy per antiparallel bond at the word with that
//                  bond's two spins flipped.  Bonds (b, b+1 mod L) for b <
//                  L - obc.  Word <-> index by the combinatorial number system.
//   honeycomb      tight-binding graphene sheet of W x L sites in the
//                  brick-wall numbering of drivers/matfuncs/graphene.c:38-245
//                  (row = l*W + w, vertical bond up when (l+w) is odd):
//                  nearest neighbours t1 = -1, with long_range also second
//                  (t2) and third (t3) neighbours and 3*t3 on the diagonal;
//                  optional periodic wrap in w and l; on-site disorder
//                  gamma*(u - 1/2), u uniform from a per-row SplitMix64
//                  stream (the reference draws C rand() in call order).
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

// ---- spin chain in an S_z sector ----
struct Binom {
  uint64_t c[65][65];
  Binom() {
    for (int n = 0; n <= 64; n++)
      for (int k = 0; k <= 64; k++) c[n][k] = k == 0 ? 1 : (n == 0 ? 0 : c[n - 1][k - 1] + c[n - 1][k]);
  }
};
const Binom& binom() {
  static const Binom b;
  return b;
}

// index of a word among the L-bit words with the same number of ones,
// ascending: sum over its ones at positions p_1 < p_2 < ... of C(p_j, j)
inline int64_t word_rank(uint64_t s) {
  const Binom& b = binom();
  int64_t r = 0;
  int j = 1;
  while (s) {
    const int p = __builtin_ctzll(s);
    r += (int64_t)b.c[p][j++];
    s &= s - 1;
  }
  return r;
}

inline uint64_t word_unrank(int64_t r, int L, int k) {
  const Binom& b = binom();
  uint64_t s = 0;
  for (int p = L - 1; p >= 0 && k > 0; p--)
    if ((uint64_t)r >= b.c[p][k]) {
      s |= 1ull << p;
      r -= (int64_t)b.c[p][k];
      k--;
    }
  return s;
}

template <typename T>
int spin_sz_row(phist_gidx row, phist_lidx* nnz, phist_gidx* cols, void* vals, void* arg) {
  const auto* p = static_cast<const phist_amd_spinchain_sz_params*>(arg);
  T* v = static_cast<T*>(vals);
  const int L = p->L;
  const uint64_t s = word_unrank(row, L, p->nup);
  const int nb = L - (p->obc ? 1 : 0);
  auto bond = [L](int b) { return (1ull << b) | (1ull << ((b + 1) % L)); };
  auto anti = [&](int b) { return __builtin_popcountll(s & bond(b)) == 1; };
  int64_t hp = 0;
  if (p->jz != 0.0)
    for (int b = 0; b < nb; b++) hp += anti(b) ? -1 : 1;
  int k = 0;
  if (hp != 0) {
    cols[k] = row;
    v[k++] = (T)((double)hp * p->jz);
  }
  if (p->jxy != 0.0)
    for (int b = 0; b < nb; b++)
      if (anti(b)) {
        cols[k] = word_rank(s ^ bond(b));
        v[k++] = (T)(2.0 * p->jxy);
      }
  *nnz = k;
  return 0;
}

// ---- honeycomb lattice ----
template <typename T>
int honeycomb_row(phist_gidx row, phist_lidx* nnz, phist_gidx* cols, void* vals, void* arg) {
  const auto* p = static_cast<const phist_amd_honeycomb_params*>(arg);
  T* v = static_cast<T*>(vals);
  const int64_t W = p->W, L = p->L;
  const int64_t l = row / W, w = row % W;
  const bool up = ((l + w) & 1) != 0;
  const double t1 = -1.0, t2 = -0.1 / 2.78, t3 = -0.095 / 2.78;
  int k = 0;
  // neighbour (l + dl, w + dw) with value a, wrapped or dropped at the edges
  auto add = [&](int dl, int dw, double a) {
    int64_t ll = l + dl, ww = w + dw;
    if (ww < 0 || ww >= W) {
      if (!p->pbc_w) return;
      ww = (ww + W) % W;
    }
    if (ll < 0 || ll >= L) {
      if (!p->pbc_l) return;
      ll = (ll + L) % L;
    }
    cols[k] = ll * W + ww;
    v[k++] = (T)a;
  };
  const bool lr = p->long_range != 0;
  for (int dl = -1; dl <= 1; dl += 2) {
    // the vertical bond of this site points up (dl = +1) when up is set
    const bool bonded = (dl > 0) == up;
    if (bonded) add(dl, 0, t1);
    if (!lr) continue;
    if (!bonded) add(dl, 0, t3);
    add(dl, -1, t2);
    add(dl, +1, t2);
    if (bonded) {
      add(dl, -2, t3);
      add(dl, +2, t3);
    }
  }
  add(0, -1, t1);
  add(0, +1, t1);
  if (lr) {
    add(0, -2, t2);
    add(0, +2, t2);
  }
  uint64_t st = p->seed ^ (0xD1B54A32D192ED03ULL * (uint64_t)(row + 1));
  const double u = (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
  cols[k] = row;
  v[k++] = (T)(p->gamma * (u - 0.5) + (lr ? 3.0 * t3 : 0.0));
  *nnz = k;
  return 0;
}

}  // namespace

extern "C" {

int phist_amd_spinchain_sz_rowfunc(phist_gidx row, phist_lidx* nnz, phist_gidx* cols, void* vals, void* arg) {
  return spin_sz_row<double>(row, nnz, cols, vals, arg);
}
int phist_amd_spinchain_sz_rowfunc_f(phist_gidx row, phist_lidx* nnz, phist_gidx* cols, void* vals, void* arg) {
  return spin_sz_row<float>(row, nnz, cols, vals, arg);
}
void phist_amd_spinchain_sz_dims(const phist_amd_spinchain_sz_params* p, phist_gidx* nrows, phist_lidx* maxnne,
                                 int* iflag) {
  *iflag = 0;
  if (p->L < 2 || p->L > 62 || p->nup < 0 || p->nup > p->L) {
    *iflag = PHIST_INVALID_INPUT;
    return;
  }
  *nrows = (phist_gidx)binom().c[p->L][p->nup];
  *maxnne = p->L + 1;
}
int phist_amd_honeycomb_rowfunc(phist_gidx row, phist_lidx* nnz, phist_gidx* cols, void* vals, void* arg) {
  return honeycomb_row<double>(row, nnz, cols, vals, arg);
}
int phist_amd_honeycomb_rowfunc_f(phist_gidx row, phist_lidx* nnz, phist_gidx* cols, void* vals, void* arg) {
  return honeycomb_row<float>(row, nnz, cols, vals, arg);
}
void phist_amd_honeycomb_dims(const phist_amd_honeycomb_params* p, phist_gidx* nrows, phist_lidx* maxnne,
                              int* iflag) {
  *iflag = 0;
  if (p->W < 1 || p->L < 1) {
    *iflag = PHIST_INVALID_INPUT;
    return;
  }
  *nrows = p->W * p->L;
  *maxnne = p->long_range ? 13 : 4;
}

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
