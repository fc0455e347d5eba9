/*
 * carp_oracle.c -- CPU restatement of PHIST's builtin Kaczmarz row update
 * (the CARP sweep kernel).
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker; the
 * product library never links or calls it.
 *
 *   oracle_kacz_rows_d / _f   src/kernels/builtin/kacz_sub_def.h:57-139, the
 *                             body of one row i of a sweep, applied to the
 *                             rows seq[0..nseq) one after another
 *                             (kacz_loop_seq_def.h:24 with a given ordering).
 *
 * The builtin sweep visits rows 1..n then n..1 (crsmat_module.F90:2686-2750)
 * or, with several threads, the colours of a distance-2 colouring; the row
 * update is the same in both.  Shift terms: (A - sigma_j I) row i, so the
 * row product starts at -sigma_r x_i (RC: -sigma_r x_r + sigma_i x_i for the
 * real part and -sigma_i x_r - sigma_r x_i for the imaginary part), the row
 * norm is sum a^2 - 2 d sigma_r + sigma_r^2 (+ sigma_i^2) with d = a_ii,
 * replaced by 1 if 0, the scaled residual is tmp * omega / norm, and the
 * projection adds tmp*sigma_r (RC: tmp_r sigma_r + tmp_i sigma_i and
 * tmp_i sigma_r - tmp_r sigma_i) to x_i before subtracting tmp * a_ij from
 * every x_j of the row.
 *
 * The reference leaves contraction to the Fortran compiler; this
 * restatement fixes it: the products accumulate with fma(a, x_j, tmp), the
 * norm with fma(a, a, nrm), the shift corrections and projections with the
 * fma forms written below.  The HIP kernel uses the same forms, so the two
 * agree bit for bit on the same row order.
 *
 * Layout: builtin CRS (1-based int64 row_offset, 1-based int32 col_idx),
 * mvecs row-major (element (i,j) at x[i*ldx+j]); si, xi and b may be NULL
 * (real shift / b = 0).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>

#define KACZ_ROWS(NAME, T, FMA)                                                                               \
  void NAME(const int64_t* row_offset, const int32_t* col_idx, const T* val, int nvec, const T* sr,         \
            const T* si, const T* omega, const T* b, int64_t ldb, T* xr, T* xi, int64_t ldx, const int32_t* seq, \
            int64_t nseq) {                                                                                  \
    for (int64_t q = 0; q < nseq; q++) {                                                                     \
      const int64_t i = seq[q];                                                                              \
      const int64_t k0 = row_offset[i] - 1, k1 = row_offset[i + 1] - 1;                                     \
      for (int c = 0; c < nvec; c++) {                                                                       \
        const T s_r = sr[c], s_i = si ? si[c] : (T)0, w = omega[c];                                          \
        const T xri = xr[i * ldx + c];                                                                       \
        T tr, ti = 0;                                                                                        \
        if (xi) {                                                                                            \
          const T xii = xi[i * ldx + c];                                                                     \
          tr = FMA(s_i, xii, -s_r * xri);                                                                    \
          ti = FMA(-s_i, xri, -s_r * xii);                                                                   \
        } else {                                                                                             \
          tr = -s_r * xri;                                                                                   \
        }                                                                                                    \
        T d = 0, nrm = 0;                                                                                    \
        for (int64_t k = k0; k < k1; k++) {                                                                  \
          const int64_t j = col_idx[k] - 1;                                                                  \
          const T a = val[k];                                                                                \
          tr = FMA(a, xr[j * ldx + c], tr);                                                                  \
          if (xi) ti = FMA(a, xi[j * ldx + c], ti);                                                          \
          nrm = FMA(a, a, nrm);                                                                              \
          if (j == i) d = a;                                                                                 \
        }                                                                                                    \
        if (b) tr = tr - b[i * ldb + c];                                                                     \
        nrm = FMA(s_r, s_r, FMA((T)(-2) * d, s_r, nrm));                                                     \
        if (xi) nrm = FMA(s_i, s_i, nrm);                                                                    \
        if (nrm == (T)0) nrm = 1;                                                                            \
        tr = (tr * w) / nrm;                                                                                 \
        if (xi) ti = (ti * w) / nrm;                                                                         \
        if (xi) {                                                                                            \
          const T xii = xi[i * ldx + c];                                                                     \
          xr[i * ldx + c] = xri + FMA(ti, s_i, tr * s_r);                                                    \
          xi[i * ldx + c] = xii + FMA(-tr, s_i, ti * s_r);                                                   \
        } else {                                                                                             \
          xr[i * ldx + c] = FMA(tr, s_r, xri);                                                               \
        }                                                                                                    \
        for (int64_t k = k0; k < k1; k++) {                                                                  \
          const int64_t j = col_idx[k] - 1;                                                                  \
          const T a = val[k];                                                                                \
          xr[j * ldx + c] = FMA(-tr, a, xr[j * ldx + c]);                                                    \
          if (xi) xi[j * ldx + c] = FMA(-ti, a, xi[j * ldx + c]);                                            \
        }                                                                                                    \
      }                                                                                                      \
    }                                                                                                        \
  }

KACZ_ROWS(oracle_kacz_rows_d, double, fma)
KACZ_ROWS(oracle_kacz_rows_f, float, fmaf)
