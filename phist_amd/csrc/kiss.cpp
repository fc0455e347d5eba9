// kiss.cpp -- PHIST's global KISS random stream with O(log n) jump-ahead,
// generated in parallel segments.
//
// The stream is Marsaglia's 64-bit KISS (multiply-with-carry with multiplier
// 2^58+1, xorshift 13/17/43, congruential 6906969069 z + 1234567) from the
// fixed initial state of src/tools/phist_random.c:76-90; a draw is
// 2^-63 * (mwc + xsh + cng) - 1 in [-1, 1) (phist_random.c:577-580).
// Reproducibility across thread counts comes from jumping each segment's
// copy of the state to its first index, as drandom_1 does
// (phist_random.c:107-194).  The jump-ahead here is derived independently:
//   * the MWC pair (x, c) is the residue y = c*2^64 + x of the
//     multiplicative congruential generator y <- a*y mod (a*2^64 - 1),
//     a = 2^58 + 1, so n steps multiply by a^n (square-and-multiply mod m);
//   * the xorshift step is a linear map over GF(2)^64, n steps apply M^n,
//     built from the precomputed squares M^(2^i);
//   * the congruential step is the affine map z <- r z + c (mod 2^64), n
//     steps compose it with itself by repeated squaring of (r, c).
#include "amd_common.hpp"
#include <omp.h>
#include <mutex>

namespace phx {
namespace {

struct Kiss {
  uint64_t x, y, z, c;
};

Kiss g_state;
std::mutex g_mutex;

inline uint64_t step(Kiss& s) {
  uint64_t t = (s.x << 58) + s.c;
  s.c = s.x >> 6;
  s.x += t;
  s.c += (s.x < t);
  s.y ^= s.y << 13;
  s.y ^= s.y >> 17;
  s.y ^= s.y << 43;
  s.z = 6906969069ULL * s.z + 1234567ULL;
  return s.x + s.y + s.z;
}

inline double to_unit(uint64_t r) { return 1.0 / 9223372036854775808.0 * (double)r - 1.0; }

// ---- MWC as a multiplicative congruential generator mod m = a*2^64-1 ----
typedef unsigned __int128 u128;
const u128 kMwcA = ((u128)1 << 58) + 1;
const u128 kMwcM = (kMwcA << 64) - 1;  // < 2^123

u128 addmod(u128 a, u128 b) {  // a, b < m < 2^123: no overflow of u128
  u128 s = a + b;
  return s >= kMwcM ? s - kMwcM : s;
}

u128 mulmod(u128 a, u128 b) {
  u128 r = 0;
  a %= kMwcM;
  b %= kMwcM;
  while (b) {
    if (b & 1) r = addmod(r, a);
    a = addmod(a, a);
    b >>= 1;
  }
  return r;
}

void mwc_jump(Kiss& s, uint64_t n) {
  u128 p = 1, base = kMwcA;
  while (n) {
    if (n & 1) p = mulmod(p, base);
    base = mulmod(base, base);
    n >>= 1;
  }
  u128 y = ((u128)s.c << 64) | s.x;
  y = mulmod(y, p);
  s.x = (uint64_t)y;
  s.c = (uint64_t)(y >> 64);
}

// ---- xorshift as a GF(2) linear map; a matrix is 64 column words ----
struct BitMat {
  uint64_t col[64];
};

uint64_t apply(const BitMat& M, uint64_t v) {
  uint64_t r = 0;
  for (int i = 0; v; i++, v >>= 1)
    if (v & 1) r ^= M.col[i];
  return r;
}

BitMat compose(const BitMat& A, const BitMat& B) {  // A*B
  BitMat C;
  for (int i = 0; i < 64; i++) C.col[i] = apply(A, B.col[i]);
  return C;
}

uint64_t xsh_once(uint64_t y) {
  y ^= y << 13;
  y ^= y >> 17;
  y ^= y << 43;
  return y;
}

BitMat g_xsh_pow[64];  // M^(2^i)
bool g_tables = false;

void build_tables() {
  BitMat M;
  for (int i = 0; i < 64; i++) M.col[i] = xsh_once(1ULL << i);
  g_xsh_pow[0] = M;
  for (int i = 1; i < 64; i++) g_xsh_pow[i] = compose(g_xsh_pow[i - 1], g_xsh_pow[i - 1]);
  g_tables = true;
}

void xsh_jump(Kiss& s, uint64_t n) {
  for (int i = 0; n; i++, n >>= 1)
    if (n & 1) s.y = apply(g_xsh_pow[i], s.y);
}

// ---- congruential generator as an affine map ----
void cng_jump(Kiss& s, uint64_t n) {
  uint64_t R = 1, Cc = 0;                      // accumulated map z -> R z + Cc
  uint64_t r = 6906969069ULL, c = 1234567ULL;  // map for 2^i steps
  while (n) {
    if (n & 1) {  // apply (r,c) after (R,Cc)
      Cc = r * Cc + c;
      R = r * R;
    }
    c = r * c + c;
    r = r * r;
    n >>= 1;
  }
  s.z = R * s.z + Cc;
}

void jump(Kiss& s, uint64_t n) {
  if (n < 4096) {
    for (uint64_t i = 0; i < n; i++) step(s);
    return;
  }
  mwc_jump(s, n);
  xsh_jump(s, n);
  cng_jump(s, n);
}

}  // namespace

void kiss_reset() {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!g_tables) build_tables();
  g_state.x = 1234567890987654321ULL;
  g_state.y = 362436362436362436ULL;
  g_state.z = 1066149217761810ULL;
  g_state.c = 123456123456123456ULL;
}

void kiss_skip(uint64_t n) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!g_tables) build_tables();
  jump(g_state, n);
}

void kiss_draw(int64_t n, double* out) {
  std::lock_guard<std::mutex> lk(g_mutex);
  if (!g_tables) build_tables();
  const Kiss start = g_state;
  const int64_t seg = 1 << 18;
  const int64_t nseg = (n + seg - 1) / seg;
#pragma omp parallel for schedule(dynamic, 1) if (nseg > 1)
  for (int64_t k = 0; k < nseg; k++) {
    Kiss s = start;
    jump(s, (uint64_t)(k * seg));
    const int64_t e = std::min(n, (k + 1) * seg);
    for (int64_t i = k * seg; i < e; i++) out[i] = to_unit(step(s));
  }
  jump(g_state, (uint64_t)n);
}

}  // namespace phx

// test hooks: raw access to the jump-ahead (declared in phist_amd_ext.h)
extern "C" void phist_amd_kiss_reset(int* iflag) {
  phx::guard(iflag, __func__, [&] {
    phx::kiss_reset();
    *iflag = 0;
  });
}
extern "C" void phist_amd_kiss_skip(uint64_t n, int* iflag) {
  phx::guard(iflag, __func__, [&] {
    phx::kiss_skip(n);
    *iflag = 0;
  });
}
extern "C" void phist_amd_kiss_draw(int64_t n, double* out, int* iflag) {
  phx::guard(iflag, __func__, [&] {
    phx::kiss_draw(n, out);
    *iflag = 0;
  });
}
