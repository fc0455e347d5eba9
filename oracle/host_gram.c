/*
 * host_gram.c -- V'W of two tall-skinny row-major FP64 blocks on the host's
 * cores, the CPU reference of the Gram and CholQR benches.
 *
 * TEST INFRASTRUCTURE AND CPU REFERENCE TIMING ONLY: bench.py's cpu_baseline
 * leg and tests/ call it; the product library never links or runs it.
 *
 * Restates the reference's builtin Gram kernels dgemm_sC_k_k / dgemm_sC_k
 * (src/kernels/builtin/gemm_kernels_sC.F90:17-168): an OpenMP loop over the
 * rows, each thread adding v(:,i) * w(j,i) into its own k x m partial sums,
 * the partial sums reduced at the end (here in thread order).  That file
 * needs PHIST's CMake-generated phist_config_fortran.h and does not build
 * here.  A thread's rows are taken in chunks that stay in L2, each chunk
 * tile by tile (8 columns of V by 4 of W, the sums in registers); the loop is
 * compiled twice, for AVX2 + FMA and for the base instruction set, and the
 * first is taken where the CPU has it.
 */
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define KMAX 64
#define TA 8      /* V columns of a register tile */
#define TBW 4     /* W columns of a register tile */
#define CHUNK 2048 /* rows whose tiles are all taken before the next rows (they stay in L2) */

/* loc (k x m, column-major) += V(i0:i1, :)' W(i0:i1, :), tile by tile over row chunks; each tile's
 * TA x TBW sums live in registers across the chunk's rows (columns beyond k or m read as 0) */
#define GRAM_ROWS(NAME, ATTR)                                                                              \
  ATTR static void NAME(const double* restrict v, const double* restrict w, int k, int m, double* restrict loc, \
                        int64_t i0, int64_t i1) {                                                           \
    for (int64_t c0 = i0; c0 < i1; c0 += CHUNK) {                                                           \
      const int64_t c1 = c0 + CHUNK < i1 ? c0 + CHUNK : i1;                                                 \
      for (int b0 = 0; b0 < m; b0 += TBW)                                                                   \
        for (int a0 = 0; a0 < k; a0 += TA) {                                                                \
          double s[TBW][TA];                                                                                \
          for (int b = 0; b < TBW; b++)                                                                     \
            for (int a = 0; a < TA; a++) s[b][a] = 0.0;                                                     \
          if (a0 + TA <= k && b0 + TBW <= m) {                                                              \
            for (int64_t i = c0; i < c1; i++) {                                                             \
              const double* vi = v + i * k + a0;                                                            \
              const double* wi = w + i * m + b0;                                                            \
              for (int b = 0; b < TBW; b++)                                                                 \
                for (int a = 0; a < TA; a++) s[b][a] += vi[a] * wi[b];                                      \
            }                                                                                               \
          } else {                                                                                          \
            const int na = k - a0 < TA ? k - a0 : TA, nb = m - b0 < TBW ? m - b0 : TBW;                     \
            for (int64_t i = c0; i < c1; i++)                                                               \
              for (int b = 0; b < nb; b++)                                                                  \
                for (int a = 0; a < na; a++) s[b][a] += v[i * k + a0 + a] * w[i * m + b0 + b];               \
          }                                                                                                 \
          for (int b = 0; b < TBW && b0 + b < m; b++)                                                       \
            for (int a = 0; a < TA && a0 + a < k; a++) loc[(int64_t)(b0 + b) * k + a0 + a] += s[b][a];      \
        }                                                                                                   \
    }                                                                                                       \
  }

GRAM_ROWS(gram_rows_avx2, __attribute__((target("avx2,fma"))))
GRAM_ROWS(gram_rows_base, )

/* M (k x m, column-major, leading dimension k) = V'W for V (n x k) and W
 * (n x m), both row-major and contiguous; 0, or -1 for bad sizes */
int host_gram(const double* v, const double* w, int64_t n, int k, int m, double* M) {
  if (n < 0 || k < 1 || m < 1 || k > KMAX || m > KMAX) return -1;
  const int avx2 = __builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma");
  const int nt = omp_get_max_threads();
  double* part = calloc((size_t)nt * k * m, sizeof(double));
  if (!part) return -44;
#pragma omp parallel num_threads(nt)
  {
    const int t = omp_get_thread_num(), T = omp_get_num_threads();
    const int64_t step = (n + T - 1) / T;
    const int64_t i0 = t * step < n ? t * step : n, i1 = i0 + step < n ? i0 + step : n;
    double* loc = part + (size_t)t * k * m;
    if (avx2)
      gram_rows_avx2(v, w, k, m, loc, i0, i1);
    else
      gram_rows_base(v, w, k, m, loc, i0, i1);
  }
  memset(M, 0, sizeof(double) * k * m);
  for (int t = 0; t < nt; t++)
    for (int e = 0; e < k * m; e++) M[e] += part[(size_t)t * k * m + e];
  free(part);
  return 0;
}
