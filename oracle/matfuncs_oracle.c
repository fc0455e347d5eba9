/*
 * matfuncs_oracle.c -- CPU restatement of two of PHIST's test-problem
 * generators (drivers/matfuncs), for checking the library's generators.
 *
 * TEST INFRASTRUCTURE ONLY (same rules as builtin_oracle.c: loaded by tests/
 * as the checker, never by the product library).
 *
 *   oracle_spin_sz_*   drivers/matfuncs/SpinChainSZ.c: the lookup tables of
 *                      the row == -2 set-up (:279-318: cnt, luttrail, revcnt,
 *                      lutrev, lutlead), the row -> word binary search and
 *                      index map (:206-231), the diagonal (:237-255) and the
 *                      flipped-bond entries with the word -> index map
 *                      lutlead[lead] + luttrail[trail] - 1 (:262-281).
 *   oracle_honeycomb_row  drivers/matfuncs/graphene.c:38-245 (zigzag_first,
 *                      J_mode 0, symmetric flag off), as an explicit
 *                      neighbour list per (l, w) block in the order the
 *                      reference emits them; the on-site value is passed in
 *                      (the reference draws it from C rand()).
 *
 * The structures are restated in the reference's own way (tables and the
 * explicit case tree), which differs from the library's (combinatorial
 * number system, offset loop), so agreement is a real cross-check; both are
 * pinned independently by brute-force Hamiltonians in the tests.
 */
#include <stdint.h>
#include <stdlib.h>

typedef struct {
  int L, nup, obc;
  int64_t half;              /* 2^(L/2) */
  int64_t *lutlead, *luttrail, *lutrev, *revcnt, *cnt;
} oracle_spin_sz_t;

static int64_t popc(int64_t i) { int64_t c = 0; while (i) { c += i & 1; i >>= 1; } return c; }

static int64_t binomial(int64_t n, int64_t k)
{
  if (k > n || k < 0) return 0;
  if (k == n || k == 0) return 1;
  if (2 * k > n) k = n - k;
  int64_t b = 1;
  for (int64_t i = 1; i <= k; i++) b = (b * ((n - k) + i)) / i;
  return b;
}

/* SpinChainSZ.c:279-318 */
oracle_spin_sz_t* oracle_spin_sz_init(int L, int nup, int obc, int64_t* nrows)
{
  oracle_spin_sz_t* s = (oracle_spin_sz_t*)calloc(1, sizeof(*s));
  s->L = L; s->nup = nup; s->obc = obc & 1;
  s->half = (int64_t)1 << (L / 2);
  s->lutlead = (int64_t*)malloc(sizeof(int64_t) * (s->half + 1));
  s->luttrail = (int64_t*)malloc(sizeof(int64_t) * s->half);
  s->lutrev = (int64_t*)malloc(sizeof(int64_t) * (s->half + 1));
  s->cnt = (int64_t*)calloc(L + 1, sizeof(int64_t));
  s->revcnt = (int64_t*)calloc(L + 1, sizeof(int64_t));
  for (int64_t i = 0; i < s->half; i++) { int64_t t = popc(i); s->cnt[t]++; s->luttrail[i] = s->cnt[t]; }
  s->revcnt[0] = 0;
  for (int i = 1; i <= L; i++) s->revcnt[i] = s->revcnt[i - 1] + s->cnt[i - 1];
  for (int64_t i = 0; i < s->half; i++) { int64_t t = popc(i); s->lutrev[s->revcnt[t]] = i; s->revcnt[t]++; }
  s->revcnt[0] = 0;
  for (int i = 1; i <= L; i++) s->revcnt[i] = s->revcnt[i - 1] + s->cnt[i - 1];
  s->lutlead[0] = 0;
  for (int64_t i = 0; i < s->half; i++) {
    int64_t t = nup - popc(i);
    s->lutlead[i + 1] = s->lutlead[i] + ((t >= 0 && t <= nup) ? s->cnt[t] : 0);
  }
  *nrows = binomial(L, nup);
  return s;
}

void oracle_spin_sz_free(oracle_spin_sz_t* s)
{
  free(s->lutlead); free(s->luttrail); free(s->lutrev); free(s->cnt); free(s->revcnt); free(s);
}

static int64_t rot(int64_t i, int64_t sh, int64_t L)
{
  if (sh < 0) sh += L;
  return (((int64_t)1 << L) - 1) & ((i << sh) | (i >> (L - sh)));
}

/* exactly one of the two masked bits set (CHECK_DIFF_NN, :162) */
static int diff_nn(int64_t t, int64_t m) { return !((m == (m & t)) || !(m & t)); }

/* SpinChainSZ.c:206-281; jz = jxy = 1 as in the reference */
int oracle_spin_sz_row(const oracle_spin_sz_t* s, int64_t row, int* nnz, int64_t* cols, double* vals)
{
  const int64_t L = s->L;
  int64_t min_il = 0, max_il = s->half, il = 0;
  for (int64_t i = 0; i < L + 1; i++) {
    il = (min_il + max_il) / 2;
    if (row >= s->lutlead[il + 1]) min_il = il + 1;
    else if (max_il != min_il) max_il = il;
    else break;
  }
  const int64_t bp = (il << (L / 2)) + s->lutrev[row - s->lutlead[il] + s->revcnt[s->nup - popc(il)]];
  if (popc(bp) != s->nup) return -1;
  int64_t hp = 0;
  for (int64_t l = 0; l < L - s->obc; l++) hp += diff_nn(bp, rot(3, l, L)) ? -1 : 1;
  int k = 0;
  if (hp != 0) { cols[k] = row; vals[k] = (double)hp; k++; }
  for (int64_t l = 0; l < L - s->obc; l++)
    if (diff_nn(bp, rot(3, l, L))) {
      const int64_t bp2 = bp ^ rot(3, l, L);
      cols[k] = s->lutlead[bp2 >> (L / 2)] + s->luttrail[bp2 & (s->half - 1)] - 1;
      vals[k] = 2.0;
      k++;
    }
  *nnz = k;
  return 0;
}

/* graphene.c:38-245 with zigzag_first = 1, J_mode = 0 (all factors 1) */
int oracle_honeycomb_row(int64_t W, int64_t L, int pbc_w, int pbc_l, int lr, double onsite, int64_t row,
                         int* nnz, int64_t* cols, double* vals)
{
  const double t1 = -1.0, t2 = -0.1 / 2.78, t3 = -0.095 / 2.78;
  const int64_t N = W * L, w = row % W, l = row / W;
  const int up = (int)((l + w) & 1);
  int i = 0;
#define E(c, v) do { cols[i] = (c); vals[i] = (v); i++; } while (0)
  if (pbc_l && l == L - 1) {
    if (lr) {
      if (pbc_w && w > W - 2) E(row + W - N + 1 - W, t2);
      if (pbc_w && w > W - 3 && up) E(row + W - N + 2 - W, t3);
      if (w > 1 && up) E(row + W - N - 2, t3);
      if (w > 0) E(row + W - N - 1, t2);
    }
    if (up) E(row + W - N, t1);
    if (!up && lr) E(row + W - N, t3);
    if (lr) {
      if (w < W - 1) E(row + W - N + 1, t2);
      if (w < W - 2 && up) E(row + W - N + 2, t3);
      if (pbc_w && w < 2 && up) E(row + W - N - 2 + W, t3);
      if (pbc_w && w < 1) E(row + W - N - 1 + W, t2);
    }
  }
  if (l > 0) {
    if (lr) {
      if (pbc_w && w > W - 2) E(row - W + 1 - W, t2);
      if (pbc_w && w > W - 3 && !up) E(row - W + 2 - W, t3);
      if (w > 1 && !up) E(row - W - 2, t3);
      if (w > 0) E(row - W - 1, t2);
    }
    if (!up) E(row - W, t1);
    if (up && lr) E(row - W, t3);
    if (lr) {
      if (w < W - 1) E(row - W + 1, t2);
      if (w < W - 2 && !up) E(row - W + 2, t3);
      if (pbc_w && w < 2 && !up) E(row - W - 2 + W, t3);
      if (pbc_w && w < 1) E(row - W - 1 + W, t2);
    }
  }
  if (pbc_w) {
    if (w > W - 2) E(row + 1 - W, t1);
    if (w > W - 3 && lr) E(row + 2 - W, t2);
  }
  if (w > 1 && lr) E(row - 2, t2);
  if (w > 0) E(row - 1, t1);
  E(row, onsite + (lr ? 3 * t3 : 0.0));
  if (w < W - 1) E(row + 1, t1);
  if (w < W - 2 && lr) E(row + 2, t2);
  if (pbc_w) {
    if (w < 2 && lr) E(row - 2 + W, t2);
    if (w == 0) E(row - 1 + W, t1);
  }
  if (l < L - 1) {
    if (lr) {
      if (pbc_w && w > W - 2) E(row + W + 1 - W, t2);
      if (pbc_w && w > W - 3 && up) E(row + W + 2 - W, t3);
      if (w > 1 && up) E(row + W - 2, t3);
      if (w > 0) E(row + W - 1, t2);
    }
    if (up) E(row + W, t1);
    if (!up && lr) E(row + W, t3);
    if (lr) {
      if (w < W - 1) E(row + W + 1, t2);
      if (w < W - 2 && up) E(row + W + 2, t3);
      if (pbc_w && w < 2 && up) E(row + W - 2 + W, t3);
      if (pbc_w && w < 1) E(row + W - 1 + W, t2);
    }
  }
  if (pbc_l && l == 0) {
    if (lr) {
      if (pbc_w && w > W - 2) E(row - W + N + 1 - W, t2);
      if (pbc_w && w > W - 3 && !up) E(row - W + N + 2 - W, t3);
      if (w > 1 && !up) E(row - W + N - 2, t3);
      if (w > 0) E(row - W + N - 1, t2);
    }
    if (!up) E(row - W + N, t1);
    if (up && lr) E(row - W + N, t3);
    if (lr) {
      if (w < W - 1) E(row - W + N + 1, t2);
      if (w < W - 2 && !up) E(row - W + N + 2, t3);
      if (pbc_w && w < 2 && !up) E(row - W + N - 2 + W, t3);
      if (pbc_w && w < 1) E(row - W + N - 1 + W, t2);
    }
  }
#undef E
  *nnz = i;
  return 0;
}
