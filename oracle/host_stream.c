/*
 * host_stream.c -- STREAM triad on the host cores, reported beside the CPU
 * baselines the way the reference reports its stream benchmark next to its
 * kernel benchmarks (src/kernels/common/stream_bench.c,
 * dbench_stream_triad_run: z = x + a*y over OpenMP threads with streaming
 * stores, 3 * 8 bytes per element).  That file needs PHIST's CMake-generated
 * phist_config.h and does not build here; this is a restatement.
 *
 * TEST INFRASTRUCTURE AND CPU REFERENCE TIMING ONLY: bench.py's cpu_baseline leg calls
 * it; the product library never links or runs it.
 */
#include <emmintrin.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>

/* best and mean GB/s of `reps` triads over n doubles (n a multiple of 2),
 * after one untimed run; arrays first-touched by the threads that use them.
 * Returns 0, -1 for bad arguments, -44 if allocation fails, -2 if the result
 * differs from x + a*y */
int host_stream_triad(int64_t n, int reps, double* best_gbs, double* mean_gbs)
{
  double *x = NULL, *y = NULL, *z = NULL;
  if (n < 2 || (n & 1) || reps < 1) return -1;
  if (posix_memalign((void**)&x, 2097152, n * sizeof(double)) || posix_memalign((void**)&y, 2097152, n * sizeof(double)) ||
      posix_memalign((void**)&z, 2097152, n * sizeof(double))) {
    free(x);
    free(y);
    free(z);
    return -44;
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    x[i] = (double)i / (double)n;
    y[i] = 1.0 - (double)i / (double)n;
    z[i] = 0.0;
  }
  const __m128d a = _mm_set1_pd(0.5);
  double best = 0.0, sum = 0.0;
  for (int r = -1; r < reps; r++) {
    const double t0 = omp_get_wtime();
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i += 2) {
      const __m128d xv = _mm_load_pd(x + i), yv = _mm_load_pd(y + i);
      _mm_stream_pd(z + i, _mm_add_pd(xv, _mm_mul_pd(a, yv)));
    }
    _mm_sfence();
    const double gbs = 3.0 * sizeof(double) * (double)n / (omp_get_wtime() - t0) / 1e9;
    if (r >= 0) {
      sum += gbs;
      if (gbs > best) best = gbs;
    }
  }
  /* every element written: z = x + y/2 exactly (the operands are exact in binary at
   * these n, and x + y/2 rounds the same way in any order of evaluation) */
  int64_t bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
  for (int64_t i = 0; i < n; i++) bad += z[i] != x[i] + 0.5 * y[i];
  free(x);
  free(y);
  free(z);
  if (bad) return -2;
  *best_gbs = best;
  *mean_gbs = sum / reps;
  return 0;
}
