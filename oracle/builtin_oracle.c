/*
 * builtin_oracle.c -- CPU restatement of PHIST's `builtin` kernel arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle the HIP kernel library is
 * checked against.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it, and only as the checker (or the timed CPU
 * baseline); the product library never links or calls it.
 *
 * Every function restates one reference routine (paths relative to the PHIST
 * source tree) in plain C, single-threaded, following the reference's
 * per-element floating-point order:
 *
 *   oracle_kiss_*            src/tools/phist_random.c:76-90 (initial state),
 *                            :413-432 (MWC), :479-486 (XSH), :522-526 (CNG),
 *                            :577-580 (KISSD), :107-194 (drandom_1)
 *   oracle_mvec_random       src/kernels/common/kernels_common_impl_def.h:256-361
 *   oracle_crs_sort_rows     src/kernels/builtin/crsmat_module.F90:484-516
 *                            (sort_global_cols: stable merge-rank per row)
 *   oracle_crs_from_mm       crsmat_module.F90:1674-1978 (read_mm_data)
 *   oracle_spmv              src/kernels/builtin/spmvm_kernels.F90:261-271
 *                            (+ spmvm_kernels_nt.c:103-126 when nt != 0)
 *   oracle_dot               src/kernels/builtin/dot_kernels.F90:143-156
 *   oracle_nrm2              nrm2_kernels.F90 + mvec_module.F90:1278-1280
 *   oracle_gram              gemm_kernels_sC.F90:248-265, mvec_module.F90:2738
 *   oracle_times_sdmat       gemm_kernels_sB.F90:15-43 (per-row order)
 *   oracle_times_sdmat_inplace gemm_kernels_sB.F90:595-612
 *   oracle_fused_mvsdi_mvTmv gemm_kernels_fused_sCD.F90:159-180 (W<-W*C, then V'W)
 *   oracle_fused_mvsd_mvTmv  gemm_kernels_sB_augmented.F90:149-195
 *   oracle_cholesky          src/core/phist_sdFact_kernels_def.hpp:27-130
 *   oracle_backward_subst    phist_sdFact_kernels_def.hpp:133-172
 *
 * Layouts are the builtin ones: mvecs row-major (element (i,j) at v[i*ld+j]),
 * sdMats column-major (element (i,j) at m[j*ld+i]), CRS with 1-based int64
 * row_offset and 1-based int32 col_idx.
 *
 * Parity status of each function (how it is pinned) is listed in the test
 * files that exercise it; functions without an independent pin say so there.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <stdio.h>

/* ------------------------------------------------------------------------ */
/* KISS random numbers                                                       */
/* ------------------------------------------------------------------------ */

typedef struct { uint64_t x, y, z, c; } oracle_kiss_t;

void oracle_kiss_init(oracle_kiss_t* s)
{
  s->x = 1234567890987654321ULL;
  s->y = 362436362436362436ULL;
  s->z = 1066149217761810ULL;
  s->c = 123456123456123456ULL;
}

/* one step of the three component generators, returns the raw 64-bit sum */
static inline uint64_t kiss_step(oracle_kiss_t* s)
{
  /* multiply-with-carry, multiplier 2^58+1 */
  uint64_t t = (s->x << 58) + s->c;
  s->c = s->x >> 6;
  s->x += t;
  s->c += (s->x < t);
  /* xorshift 13/17/43 */
  s->y ^= (s->y << 13);
  s->y ^= (s->y >> 17);
  s->y ^= (s->y << 43);
  /* congruential */
  s->z = 6906969069ULL * s->z + 1234567ULL;
  return s->x + s->y + s->z;
}

uint64_t oracle_kiss_raw(oracle_kiss_t* s) { return kiss_step(s); }

double oracle_kiss_double(oracle_kiss_t* s)
{
  uint64_t r = kiss_step(s);
  return 1.0 / 9223372036854775808.0 * (double)r - 1.0;
}

/* brute-force skip: the defining sequence, no jump-ahead mathematics */
void oracle_kiss_skip(oracle_kiss_t* s, uint64_t n)
{
  for (uint64_t i = 0; i < n; i++) (void)kiss_step(s);
}

void oracle_kiss_fill(oracle_kiss_t* s, int64_t n, double* out)
{
  for (int64_t i = 0; i < n; i++) out[i] = oracle_kiss_double(s);
}

/* mvec_random on a single rank: pre_skip = ilower*nvec = 0,
 * post_skip = (gnrows - iupper)*nvec with iupper = nrows-1, i.e. nvec.
 * The draws fill a row-major nrows x nvec block that is then copied into the
 * (possibly strided) mvec. */
void oracle_mvec_random(oracle_kiss_t* s, int64_t nrows, int nvec, double* v, int64_t ldv)
{
  for (int64_t i = 0; i < nrows; i++)
    for (int j = 0; j < nvec; j++)
      v[i * ldv + j] = oracle_kiss_double(s);
  oracle_kiss_skip(s, (uint64_t)nvec);
}

/* sdMat_random: column-major fill from the same global stream
 * (kernels_common_impl_def.h:365-405 via phist_Drandom_number). */
void oracle_sdmat_random(oracle_kiss_t* s, int nrows, int ncols, double* m, int ldm)
{
  for (int j = 0; j < ncols; j++)
    for (int i = 0; i < nrows; i++)
      m[(int64_t)j * ldm + i] = oracle_kiss_double(s);
}

/* ------------------------------------------------------------------------ */
/* CRS construction                                                          */
/* ------------------------------------------------------------------------ */

/* stable sort of one row's entries by column (merge sort, ties keep order) */
static void stable_sort_row(int64_t n, int64_t* col, double* val, int64_t* tc, double* tv)
{
  for (int64_t width = 1; width < n; width *= 2) {
    for (int64_t lo = 0; lo < n; lo += 2 * width) {
      int64_t mid = lo + width < n ? lo + width : n;
      int64_t hi = lo + 2 * width < n ? lo + 2 * width : n;
      int64_t a = lo, b = mid, k = lo;
      while (a < mid && b < hi) {
        if (col[b] < col[a]) { tc[k] = col[b]; tv[k] = val[b]; b++; }
        else                 { tc[k] = col[a]; tv[k] = val[a]; a++; }
        k++;
      }
      while (a < mid) { tc[k] = col[a]; tv[k] = val[a]; a++; k++; }
      while (b < hi)  { tc[k] = col[b]; tv[k] = val[b]; b++; k++; }
    }
    memcpy(col, tc, (size_t)n * sizeof(int64_t));
    memcpy(val, tv, (size_t)n * sizeof(double));
  }
}

/* Input: rows in fill order (row_ptr 0-based, gcol 0-based global columns).
 * Output: builtin CRS (row_offset 1-based int64 of size nrows+1, col_idx
 * 1-based int32, val) with each row stably sorted by column. */
void oracle_crs_sort_rows(int64_t nrows, const int64_t* row_ptr, const int64_t* gcol, const double* gval,
                          int64_t* row_offset, int32_t* col_idx, double* val)
{
  int64_t maxrow = 0;
  for (int64_t i = 0; i < nrows; i++) {
    int64_t n = row_ptr[i + 1] - row_ptr[i];
    if (n > maxrow) maxrow = n;
  }
  int64_t* c = (int64_t*)malloc((size_t)(maxrow + 1) * sizeof(int64_t));
  int64_t* tc = (int64_t*)malloc((size_t)(maxrow + 1) * sizeof(int64_t));
  double* v = (double*)malloc((size_t)(maxrow + 1) * sizeof(double));
  double* tv = (double*)malloc((size_t)(maxrow + 1) * sizeof(double));
  for (int64_t i = 0; i <= nrows; i++) row_offset[i] = row_ptr[i] + 1;
  for (int64_t i = 0; i < nrows; i++) {
    int64_t off = row_ptr[i], n = row_ptr[i + 1] - off;
    for (int64_t j = 0; j < n; j++) { c[j] = gcol[off + j] + 1; v[j] = gval[off + j]; }
    if (n > 1) stable_sort_row(n, c, v, tc, tv);
    for (int64_t j = 0; j < n; j++) { col_idx[off + j] = (int32_t)c[j]; val[off + j] = v[j]; }
  }
  free(c); free(tc); free(v); free(tv);
}

/* MatrixMarket "coordinate real general|symmetric": entries are bucketed by
 * row in file order (an off-diagonal symmetric entry (i,j) is followed by its
 * mirror (j,i)), then each row is stably sorted by column.
 * Returns 0 on success; the caller frees *row_offset, *col_idx, *val. */
int oracle_crs_from_mm(const char* path, int64_t* nrows_out, int64_t* ncols_out,
                       int64_t** row_offset_out, int32_t** col_idx_out, double** val_out)
{
  FILE* f = fopen(path, "r");
  if (!f) return -1;
  char line[512];
  if (!fgets(line, sizeof line, f)) { fclose(f); return -1; }
  int symmetric;
  if (strncmp(line, "%%MatrixMarket matrix coordinate real general", 45) == 0) symmetric = 0;
  else if (strncmp(line, "%%MatrixMarket matrix coordinate real symmetric", 47) == 0) symmetric = 1;
  else { fclose(f); return -2; }
  int64_t nr = 0, nc = 0, nl = 0;
  while (fgets(line, sizeof line, f)) {
    if (line[0] == '%') continue;
    if (sscanf(line, "%ld %ld %ld", &nr, &nc, &nl) == 3) break;
  }
  int64_t cap = symmetric ? 2 * nl : nl;
  int64_t* ri = (int64_t*)malloc((size_t)(cap + 1) * sizeof(int64_t));
  int64_t* ci = (int64_t*)malloc((size_t)(cap + 1) * sizeof(int64_t));
  double* vv = (double*)malloc((size_t)(cap + 1) * sizeof(double));
  int64_t ne = 0;
  for (int64_t k = 0; k < nl; k++) {
    int64_t i, j; double a;
    if (fscanf(f, "%ld %ld %lf", &i, &j, &a) != 3) { fclose(f); return -3; }
    ri[ne] = i - 1; ci[ne] = j - 1; vv[ne] = a; ne++;
    if (symmetric && i != j) { ri[ne] = j - 1; ci[ne] = i - 1; vv[ne] = a; ne++; }
  }
  fclose(f);
  int64_t* rp = (int64_t*)calloc((size_t)(nr + 1), sizeof(int64_t));
  for (int64_t k = 0; k < ne; k++) rp[ri[k] + 1]++;
  for (int64_t i = 0; i < nr; i++) rp[i + 1] += rp[i];
  int64_t* fill = (int64_t*)malloc((size_t)(nr + 1) * sizeof(int64_t));
  memcpy(fill, rp, (size_t)(nr + 1) * sizeof(int64_t));
  int64_t* gc = (int64_t*)malloc((size_t)(ne + 1) * sizeof(int64_t));
  double* gv = (double*)malloc((size_t)(ne + 1) * sizeof(double));
  for (int64_t k = 0; k < ne; k++) { int64_t p = fill[ri[k]]++; gc[p] = ci[k]; gv[p] = vv[k]; }
  int64_t* ro = (int64_t*)malloc((size_t)(nr + 1) * sizeof(int64_t));
  int32_t* co = (int32_t*)malloc((size_t)(ne + 1) * sizeof(int32_t));
  double* va = (double*)malloc((size_t)(ne + 1) * sizeof(double));
  oracle_crs_sort_rows(nr, rp, gc, gv, ro, co, va);
  free(ri); free(ci); free(vv); free(rp); free(fill); free(gc); free(gv);
  *nrows_out = nr; *ncols_out = nc;
  *row_offset_out = ro; *col_idx_out = co; *val_out = va;
  return 0;
}

void oracle_free(void* p) { free(p); }

/* ------------------------------------------------------------------------ */
/* sparse matrix times block vector                                          */
/* ------------------------------------------------------------------------ */

/* y(:,i) = alpha*(shift.*x(:,i) + sum_j val_j x(:,col_j)) + beta*y(:,i)
 * shifts may be NULL (all zero).  nt != 0 selects the non-temporal variant
 * used by builtin when beta == 0 and x, y are aligned: y = alpha*tmp. */
void oracle_spmv(int64_t nrows, const int64_t* row_offset, const int32_t* col_idx, const double* val,
                 int nvec, double alpha, const double* shifts, const double* x, int64_t ldx,
                 double beta, double* y, int64_t ldy, int nt)
{
  double tmp[512];
  if (nvec > 512) return;
  for (int64_t i = 0; i < nrows; i++) {
    for (int k = 0; k < nvec; k++) tmp[k] = (shifts ? shifts[k] : 0.0) * x[i * ldx + k];
    for (int64_t j = row_offset[i] - 1; j < row_offset[i + 1] - 1; j++) {
      const double a = val[j];
      const double* xc = x + (int64_t)(col_idx[j] - 1) * ldx;
      for (int k = 0; k < nvec; k++) tmp[k] = tmp[k] + a * xc[k];
    }
    if (nt) for (int k = 0; k < nvec; k++) y[i * ldy + k] = alpha * tmp[k];
    else    for (int k = 0; k < nvec; k++) y[i * ldy + k] = alpha * tmp[k] + beta * y[i * ldy + k];
  }
}

/* ------------------------------------------------------------------------ */
/* block-vector reductions                                                   */
/* ------------------------------------------------------------------------ */

void oracle_dot(int64_t nrows, int nvec, const double* v, int64_t ldv, const double* w, int64_t ldw,
                double* out)
{
  for (int k = 0; k < nvec; k++) out[k] = 0.0;
  for (int64_t i = 0; i < nrows; i++)
    for (int k = 0; k < nvec; k++) out[k] = out[k] + v[i * ldv + k] * w[i * ldw + k];
}

void oracle_nrm2(int64_t nrows, int nvec, const double* v, int64_t ldv, double* out)
{
  for (int k = 0; k < nvec; k++) out[k] = 0.0;
  for (int64_t i = 0; i < nrows; i++)
    for (int k = 0; k < nvec; k++) out[k] = out[k] + v[i * ldv + k] * v[i * ldv + k];
  for (int k = 0; k < nvec; k++) {
    double local = sqrt(out[k]);
    out[k] = sqrt(local * local);
  }
}

/* C(1:nv,1:nw) = alpha * V'W + beta * C, accumulation M(:,j) += v(:,i)*w(j,i) */
void oracle_gram(int64_t nrows, int nv, int nw, double alpha, const double* v, int64_t ldv,
                 const double* w, int64_t ldw, double beta, double* c, int ldc)
{
  double* m = (double*)calloc((size_t)nv * nw, sizeof(double));
  for (int64_t i = 0; i < nrows; i++)
    for (int j = 0; j < nw; j++) {
      const double wj = w[i * ldw + j];
      for (int r = 0; r < nv; r++) m[j * nv + r] = m[j * nv + r] + v[i * ldv + r] * wj;
    }
  for (int j = 0; j < nw; j++)
    for (int r = 0; r < nv; r++) c[j * ldc + r] = alpha * m[j * nv + r] + beta * c[j * ldc + r];
  free(m);
}

/* W(i,:) = beta*W(i,:) + sum_j alpha*C(j,:)*V(i,j)  (V nrows x k, C k x m) */
void oracle_times_sdmat(int64_t nrows, int k, int m, double alpha, const double* v, int64_t ldv,
                        const double* c, int ldc, double beta, double* w, int64_t ldw)
{
  for (int64_t i = 0; i < nrows; i++) {
    double* wi = w + i * ldw;
    if (beta == 0.0) for (int q = 0; q < m; q++) wi[q] = 0.0;
    else             for (int q = 0; q < m; q++) wi[q] = beta * wi[q];
    for (int j = 0; j < k; j++) {
      const double vij = v[i * ldv + j];
      for (int q = 0; q < m; q++) wi[q] = wi[q] + alpha * c[q * ldc + j] * vij;
    }
  }
}

/* V(i,1:m) <- V(i,1:k) * M  (M k x m, m <= k), work(q) = sum(V(i,1:k)*M(:,q)) */
void oracle_times_sdmat_inplace(int64_t nrows, int k, int m, double* v, int64_t ldv, const double* mm, int ldm)
{
  double work[512];
  if (m > 512) return;
  for (int64_t i = 0; i < nrows; i++) {
    for (int q = 0; q < m; q++) {
      double s = 0.0;
      for (int j = 0; j < k; j++) s = s + v[i * ldv + j] * mm[q * ldm + j];
      work[q] = s;
    }
    for (int q = 0; q < m; q++) v[i * ldv + q] = work[q];
  }
}

/* fused_mvsdi_mvTmv: W <- W*C (in place), D = alpha*V'W + beta*D */
void oracle_fused_mvsdi_mvTmv(int64_t nrows, int nv, int nw, double alpha, const double* v, int64_t ldv,
                              double* w, int64_t ldw, const double* c, int ldc, double beta, double* d, int ldd)
{
  oracle_times_sdmat_inplace(nrows, nw, nw, w, ldw, c, ldc);
  oracle_gram(nrows, nv, nw, alpha, v, ldv, w, ldw, beta, d, ldd);
}

/* fused_mvsd_mvTmv: W = alpha*V*C + beta*W, D = W'W */
void oracle_fused_mvsd_mvTmv(int64_t nrows, int k, int m, double alpha, const double* v, int64_t ldv,
                             const double* c, int ldc, double beta, double* w, int64_t ldw, double* d, int ldd)
{
  oracle_times_sdmat(nrows, k, m, alpha, v, ldv, c, ldc, beta, w, ldw);
  oracle_gram(nrows, m, m, 1.0, w, ldw, w, ldw, 0.0, d, ldd);
}

/* ------------------------------------------------------------------------ */
/* small dense factorisations (host side of CholQR)                          */
/* ------------------------------------------------------------------------ */

/* Rank-revealing diagonally pivoted Cholesky of the column-major n x n SPD
 * matrix a.  On exit a holds R with R(:,perm) upper triangular and
 * M = R'R; perm and rank are returned. */
int oracle_cholesky(double* a, int n, int lda, int* perm, int* rank, double rankTol)
{
  int* p = (int*)malloc((size_t)n * sizeof(int));
  double* l = (double*)calloc((size_t)n * n, sizeof(double));
  double* d = (double*)malloc((size_t)n * sizeof(double));
  for (int i = 0; i < n; i++) p[i] = i;
  double diagNorm = 0.0;
  for (int i = 0; i < n; i++) {
    d[i] = a[i * lda + i];
    if (fabs(d[i]) > diagNorm) diagNorm = fabs(d[i]);
  }
  if (diagNorm <= rankTol) diagNorm = 1.0;
  *rank = 0;
  while (*rank < n) {
    double err = 0.0;
    for (int i = *rank; i < n; i++) if (fabs(d[p[i]]) > err) err = fabs(d[p[i]]);
    if (err <= rankTol * diagNorm) break;
    int m = *rank;
    *rank = *rank + 1;
    int piv = m;
    for (int j = m + 1; j < n; j++) if (fabs(d[p[j]]) > fabs(d[p[piv]])) piv = j;
    int t = p[piv]; p[piv] = p[m]; p[m] = t;
    l[p[m] * n + m] = sqrt(d[p[m]]);
    double div = 1.0 / l[p[m] * n + m];
    for (int i = m + 1; i < n; i++) {
      double s = a[p[i] * lda + p[m]];
      for (int j = 0; j < m; j++) s -= l[p[m] * n + j] * l[p[i] * n + j];
      s *= div;
      l[p[i] * n + m] = s;
      d[p[i]] -= l[p[i] * n + m] * l[p[i] * n + m];
    }
  }
  for (int i = 0; i < n; i++) {
    perm[i] = p[i];
    for (int j = 0; j < n; j++) a[i * lda + j] = l[i * n + j];
  }
  free(p); free(l); free(d);
  return 0;
}

/* Solve R(:,perm) x = b for k right-hand sides stored column-major in x. */
void oracle_backward_subst(const double* r, int n, int ldr, const int* p, int rank,
                           double* x, int k, int ldx)
{
  double* nx = (double*)malloc((size_t)n * sizeof(double));
  for (int l = 0; l < k; l++) {
    for (int i = n - 1; i >= rank; i--) nx[i] = 0.0;
    for (int i = rank - 1; i >= 0; i--) {
      double rinv = 1.0 / r[p[i] * ldr + i];
      for (int j = i + 1; j < n; j++) x[l * ldx + i] -= r[p[j] * ldr + i] * nx[j];
      nx[i] = x[l * ldx + i] * rinv;
    }
    for (int i = 0; i < rank; i++) x[l * ldx + p[i]] = nx[i];
    for (int i = rank; i < n; i++) x[l * ldx + p[i]] = 0.0;
  }
  free(nx);
}
