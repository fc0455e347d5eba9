// tsgemm.hip -- tall-skinny products of block vectors with small dense matrices.
//
//   gemm_gram         C = alpha*V'W + beta*C        mvecT_times_mvec
//                     (gemm_kernels_sC.F90:248-265, mvec_module.F90:2421-2743)
//   gemm_times_sd     W = alpha*V*C + beta*W        mvec_times_sdMat
//                     (gemm_kernels_sB.F90:15-43: W = beta*W, then
//                     W += (alpha*C(j,:))*V(:,j) for j = 1..k)
//   gemm_times_sd_inplace   V(:,1:m) = V*M          (gemm_kernels_sB.F90:595-612)
//   gemm_times_sd_add_times_sd  W = W*D + V*C       (gemm_kernels_sB_add_sD.F90)
//   fused_mvsdi_mvTmv W = W*C, then D = alpha*V'W + beta*D  (gemm_kernels_fused_sCD.F90)
//   fused_mvsd_mvTmv  W = alpha*V*C + beta*W, then WtW = W'W
//                     (gemm_kernels_sB_augmented.F90:149-195)
//
// Gram on the VALU: a row of the wider operand is split over LPR lanes (VPL
// columns each, 16-byte loads), the narrower operand's row is broadcast
// inside the lane group with cross-lane shuffles, every lane accumulates a
// VPL x KB register tile; the tile sums go through the fixed-order two-stage
// reduction and a finalize kernel forms alpha*S + beta*C on the device copy.
// Row-mixing products keep alpha*C (and D) in LDS; each lane produces VPL
// output columns of one row and reads the whole input row, so in-place
// updates are safe: all lanes of a row sit in one wave and every store
// depends on every load of that row.
#include "amd_device.hpp"
#include <algorithm>
#include <limits>

namespace phx {

// ---------------------------------------------------------------------------
// Gram products
// ---------------------------------------------------------------------------
template <typename T>
struct GramArgs {
  int64_t nrows, ntiles;
  const T* __restrict__ a;  // sliced operand (width na)
  int64_t lda;
  int na;
  const T* __restrict__ b;  // broadcast operand (width nb <= LPR*VPL)
  int64_t ldb;
  int nb;
  T* __restrict__ partials;  // [grid][na*nb], index c*na + r (r over a, c over b)
};

// two consecutive scalars as one 16-byte (FP64) / 8-byte (FP32) load
template <typename T>
struct alignas(2 * sizeof(T)) T2v {
  T x, y;
};

template <typename T, int VPL, int LPR, int KB>
__global__ void __launch_bounds__(kBlock) k_gram(const GramArgs<T> g) {
  constexpr int RPW = 64 / LPR, RPB = (kBlock / 64) * RPW;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int q = lane % LPR, rw = lane / LPR, c0 = q * VPL;
  const int grp = lane - q;
  T acc[VPL][KB];
#pragma unroll
  for (int e = 0; e < VPL; e++)
#pragma unroll
    for (int c = 0; c < KB; c++) acc[e][c] = T(0);

  // UNR tiles per iteration: their loads are all in flight before the first
  // shuffle; the tiles are accumulated in walk order, as one at a time would
  constexpr int UNR = 4;
  const bool self = g.a == g.b && g.lda == g.ldb && g.na == g.nb;  // the b row is the a row
  const TileWalk tw = xcd_tiles(g.ntiles);
  for (int64_t t = tw.t; t < tw.end; t += UNR * tw.stride) {
    T av[UNR][VPL], bv[UNR][VPL];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const int64_t tu = t + u * tw.stride;
      const int64_t i = tu * RPB + wv * RPW + rw;
      const bool rowok = tu < tw.end && i < g.nrows;  // all lanes stay for the shuffles
#pragma unroll
      for (int e = 0; e < VPL; e++) av[u][e] = bv[u][e] = T(0);
      if (rowok && c0 < g.na) vload<T, VPL>(g.a + i * g.lda + c0, av[u]);
      if (!self && rowok && c0 < g.nb) vload<T, VPL>(g.b + i * g.ldb + c0, bv[u]);
    }
#pragma unroll
    for (int u = 0; u < UNR; u++) {
#pragma unroll
      for (int s = 0; s < LPR; s++) {
        if (s * VPL >= KB) break;
#pragma unroll
        for (int f = 0; f < VPL; f++) {
          const int c = s * VPL + f;
          if (c >= KB) break;
          const T w = __shfl(self ? av[u][f] : bv[u][f], grp + s, 64);
#pragma unroll
          for (int e = 0; e < VPL; e++) acc[e][c] = fma(av[u][e], w, acc[e][c]);
        }
      }
    }
  }

  // reduce over the rows of the wave, then over the waves: one partial tile
  // per workgroup
  extern __shared__ char gram_lds[];
  const int nout = g.na * g.nb;
  block_tile_sum(reinterpret_cast<T*>(gram_lds), nout, g.partials + (int64_t)blockIdx.x * nout, [&](auto add) {
#pragma unroll
    for (int e = 0; e < VPL; e++)
#pragma unroll
      for (int c = 0; c < KB; c++) {
        const T s = wave_sum_strided<LPR>(acc[e][c]);
        if (lane < LPR && c0 + e < g.na && c < g.nb) add(c * g.na + c0 + e, s);
      }
  });
}

// Gram a'b for a narrow b (nb <= NB) and na <= LPR*KA without vector loads
// (rows not 16-byte aligned): a row per LPR lanes, lane q holding columns
// q*KA .. q*KA+KA-1 of a and the whole row of b, so a wave covers 64/LPR
// rows per step instead of one or two.
template <typename T, int LPR, int KA, int NB>
__global__ void __launch_bounds__(kBlock) k_gram_rows(const GramArgs<T> g) {
  constexpr int RPW = 64 / LPR, RPB = (kBlock / 64) * RPW;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int q = lane % LPR, rw = lane / LPR, a0 = q * KA;
  T acc[KA][NB];
#pragma unroll
  for (int e = 0; e < KA; e++)
#pragma unroll
    for (int c = 0; c < NB; c++) acc[e][c] = T(0);
  const TileWalk tw = xcd_tiles(g.ntiles);
  for (int64_t t = tw.t; t < tw.end; t += tw.stride) {
    const int64_t i = t * RPB + wv * RPW + rw;
    if (i >= g.nrows) continue;
    T av[KA], bv[NB];
#pragma unroll
    for (int e = 0; e < KA; e++) av[e] = a0 + e < g.na ? g.a[i * g.lda + a0 + e] : T(0);
#pragma unroll
    for (int c = 0; c < NB; c++) bv[c] = c < g.nb ? g.b[i * g.ldb + c] : T(0);
#pragma unroll
    for (int e = 0; e < KA; e++)
#pragma unroll
      for (int c = 0; c < NB; c++) acc[e][c] = fma(av[e], bv[c], acc[e][c]);
  }
  extern __shared__ char gram_lds[];
  const int nout = g.na * g.nb;
  block_tile_sum(reinterpret_cast<T*>(gram_lds), nout, g.partials + (int64_t)blockIdx.x * nout, [&](auto add) {
#pragma unroll
    for (int e = 0; e < KA; e++)
#pragma unroll
      for (int c = 0; c < NB; c++) {
        const T s = wave_sum_strided<LPR>(acc[e][c]);
        if (lane < LPR && a0 + e < g.na && c < g.nb) add(c * g.na + a0 + e, s);
      }
  });
}

// ---- MFMA Gram (both operands at least 16 columns) ----
// v_mfma_{f64,f32}_16x16x4: lane l supplies A[l%16][l/16] and B[l/16][l%16]
// (A 16x4, B 4x16) and holds four results of the 16x16 product; the result
// row of register r is l/16 + 4r for f64 and 4*(l/16) + r for f32 (measured
// on gfx950), the column is l%16.
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  __device__ static acc_t op(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ static int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mfma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  __device__ static acc_t op(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// S = a'b for na, nb <= 32 (TM = ceil(na/16), TK = ceil(nb/16)); a wave
// consumes 16 rows per tile (four 4-row MFMA steps), two tiles per
// iteration so that both tiles' loads are in flight together; with SELF
// (a == b) only the upper block tiles are computed and mirrored.  With PAIR
// (TM = 2, and TK = 2 unless SELF; even widths, 16-byte aligned rows) lane
// cc loads columns 2cc and 2cc+1 of its row as one 16-byte vector: tile 0
// holds the even columns, tile 1 the odd ones, so column (tile t, index c)
// is 2c + t instead of 16t + c.  Every output is the same sum of the same
// products in the same MFMA steps either way.
template <typename T, int TM, int TK, bool SELF, bool PAIR = false>
__global__ void __launch_bounds__(kBlock) k_gram_mfma(const GramArgs<T> g) {
  using M = Mfma<T>;
  static_assert(!PAIR || (TM == 2 && (SELF || TK == 2)), "paired loads need two column tiles per operand");
  constexpr int UNR = 2;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int kk = lane >> 4, cc = lane & 15;
  auto colof = [&](int t, int c) { return PAIR ? 2 * c + t : 16 * t + c; };
  typename M::acc_t acc[TM][TK];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TK; j++) acc[i][j] = typename M::acc_t{0, 0, 0, 0};
  const TileWalk tw = xcd_tiles(g.ntiles);
  for (int64_t t = tw.t; t < tw.end; t += UNR * tw.stride) {
    T a[UNR][4][TM], b[UNR][4][TK];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const int64_t tu = t + u * tw.stride;
      const int64_t row0 = tu * 64 + wv * 16;
#pragma unroll
      for (int st = 0; st < 4; st++) {
        const int64_t r = row0 + st * 4 + kk;
        const bool ok = tu < tw.end && r < g.nrows;
        if constexpr (PAIR) {
          T2v<T> av{T(0), T(0)};
          if (ok && 2 * cc < g.na) av = *reinterpret_cast<const T2v<T>*>(g.a + r * g.lda + 2 * cc);
          a[u][st][0] = av.x;
          a[u][st][1] = av.y;
          if constexpr (!SELF) {
            T2v<T> bv{T(0), T(0)};
            if (ok && 2 * cc < g.nb) bv = *reinterpret_cast<const T2v<T>*>(g.b + r * g.ldb + 2 * cc);
            b[u][st][0] = bv.x;
            b[u][st][1] = bv.y;
          }
        } else {
#pragma unroll
          for (int tm = 0; tm < TM; tm++) {
            const int col = tm * 16 + cc;
            a[u][st][tm] = (ok && col < g.na) ? g.a[r * g.lda + col] : T(0);
          }
          if constexpr (!SELF) {
#pragma unroll
            for (int tk = 0; tk < TK; tk++) {
              const int col = tk * 16 + cc;
              b[u][st][tk] = (ok && col < g.nb) ? g.b[r * g.ldb + col] : T(0);
            }
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < UNR; u++)
#pragma unroll
      for (int st = 0; st < 4; st++)
#pragma unroll
        for (int tm = 0; tm < TM; tm++)
#pragma unroll
          for (int tk = 0; tk < TK; tk++) {
            if (SELF && tm > tk) continue;
            acc[tm][tk] = M::op(a[u][st][tm], SELF ? a[u][st][tk] : b[u][st][tk], acc[tm][tk]);
          }
  }
  extern __shared__ char gram_lds[];
  const int nout = g.na * g.nb;
  block_tile_sum(reinterpret_cast<T*>(gram_lds), nout, g.partials + (int64_t)blockIdx.x * nout, [&](auto add) {
#pragma unroll
    for (int tm = 0; tm < TM; tm++)
#pragma unroll
      for (int tk = 0; tk < TK; tk++) {
        if (SELF && tm > tk) continue;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int i = colof(tm, M::row(lane, r)), j = colof(tk, cc);
          const T v = acc[tm][tk][r];
          if (i < g.na && j < g.nb) add(j * g.na + i, v);
          if (SELF && tm < tk && j < g.na && i < g.nb) add(i * g.na + j, v);
        }
      }
  });
}

// Self Gram V'V of a narrow block (KC = 2, 4 or 8 columns, row stride even,
// rows 2-scalar aligned) on the matrix cores with every lane loading: lane
// cc of an MFMA column group loads the column pair p = cc % P (P = KC/2) of
// a row of row set s = cc / P, so one 16x16x4 step covers 4 rows of each of
// RS = 16/P row sets (64/P rows, all 64 lanes loading 2 scalars) instead of
// 4 rows on KC of 16 lanes.  Tile t of the pair (.x / .y) holds the even /
// odd columns; of each 16x16 product only the blocks of equal row sets are
// kept (entry (s*P+p, s*P+p') = sum over the set's rows of x(2p+ta) x(2p'+tb)).
// At the end a wave adds the RS blocks of each output in set order through
// LDS, so every output is a fixed-order sum.
template <typename T, int KC>
__global__ void __launch_bounds__(kBlock) k_gram_mfma_sets(const GramArgs<T> g) {
  using M = Mfma<T>;
  constexpr int P = KC / 2, RS = 16 / P, ROWS = 64 / P;  // rows per wave and step
  constexpr int UNR = 4;                                   // steps in flight per wave
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int kk = lane >> 4, cc = lane & 15;
  const int sset = cc / P, pr = cc % P;
  typename M::acc_t acc[3];  // tile pairs (0,0), (0,1), (1,1)
#pragma unroll
  for (int t = 0; t < 3; t++) acc[t] = typename M::acc_t{0, 0, 0, 0};
  const TileWalk tw = xcd_tiles(g.ntiles);
  for (int64_t t = tw.t; t < tw.end; t += UNR * tw.stride) {
    T2v<T> v[UNR];
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      const int64_t tu = t + u * tw.stride;
      const int64_t r = tu * (4 * ROWS) + wv * ROWS + sset * 4 + kk;
      v[u] = T2v<T>{T(0), T(0)};
      if (tu < tw.end && r < g.nrows) v[u] = *reinterpret_cast<const T2v<T>*>(g.a + r * g.lda + 2 * pr);
    }
#pragma unroll
    for (int u = 0; u < UNR; u++) {
      acc[0] = M::op(v[u].x, v[u].x, acc[0]);
      acc[1] = M::op(v[u].x, v[u].y, acc[1]);
      acc[2] = M::op(v[u].y, v[u].y, acc[2]);
    }
  }
  // the wave's three 16x16 tiles through LDS, then output (i, j) = sum over
  // the row sets in order, owned by lane i*KC + j
  extern __shared__ char gram_lds[];
  T* tiles = reinterpret_cast<T*>(gram_lds) + (size_t)wv * 3 * 256;
  T* red = reinterpret_cast<T*>(gram_lds) + (size_t)(kBlock / 64) * 3 * 256;
#pragma unroll
  for (int t = 0; t < 3; t++)
#pragma unroll
    for (int r = 0; r < 4; r++) tiles[t * 256 + M::row(lane, r) * 16 + cc] = acc[t][r];
  __syncthreads();
  T val = T(0);
  const int i = lane / KC, j = lane % KC;
  if (lane < KC * KC) {
    const int ta = i & 1, pa = i >> 1, tb = j & 1, pb = j >> 1;
    // tile (ta, tb) for ta <= tb; (1, 0) is the transpose of (0, 1)
    const int tt = ta + tb, m0 = ta <= tb ? pa : pb, n0 = ta <= tb ? pb : pa;
#pragma unroll
    for (int s2 = 0; s2 < RS; s2++) val += tiles[tt * 256 + (s2 * P + m0) * 16 + (s2 * P + n0)];
  }
  block_tile_sum(red, KC * KC, g.partials + (int64_t)blockIdx.x * KC * KC, [&](auto add) {
    if (lane < KC * KC) add(j * KC + i, val);
  });
}

// C(r,c) = alpha * S + beta * C with S summed over the block partials in a
// fixed order; S is stored c*na + r, transposed when the operands were
// swapped.  One wave per output (four per workgroup, finalize_blocks(nout)
// workgroups): lane l adds the partials of blocks l, l+64, l+128, ... in
// order, then the 64 lane sums are added by a fixed shuffle tree, so every
// block's loads are in flight at once and the sum is deterministic.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_gram_finalize(const T* __restrict__ partials, int nblocks, int na, int nb,
                                                          bool transposed, T alpha, T beta, T* __restrict__ C,
                                                          int ldc) {
  const int nout = na * nb;
  const int o = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (o >= nout) return;  // whole waves leave; no workgroup barrier follows
  T s = T(0);
#pragma unroll 8
  for (int b = lane; b < nblocks; b += 64) s += partials[(int64_t)b * nout + o];
  s = wave_sum_strided<1>(s);
  if (lane == 0) {
    const int r = o % na, c = o / na;
    const int ci = transposed ? r : c, ri = transposed ? c : r;
    T* cp = C + (int64_t)ci * ldc + ri;
    *cp = beta == T(0) ? alpha * s : alpha * s + beta * *cp;
  }
}
inline int finalize_blocks(int nout) { return (nout + kBlock / 64 - 1) / (kBlock / 64); }

namespace {

inline bool al16(const void* p) { return ((uintptr_t)p & 15) == 0; }

template <typename T, int VPL, int LPR, int KB>
int gram_launch_one(GramArgs<T>& g) {
  auto kern = k_gram<T, VPL, LPR, KB>;
  constexpr int RPB = (kBlock / 64) * (64 / LPR);
  g.ntiles = (g.nrows + RPB - 1) / RPB;
  const size_t lds = sizeof(T) * g.na * g.nb;
  const int grid = persistent_blocks(kern, lds, g.ntiles, rt().gram_blocks_per_cu);
  g.partials = (T*)device_work(sizeof(T) * (size_t)grid * g.na * g.nb);
  ktimer_begin();
  hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), lds, rt().stream, g);
  PHX_HIP(hipGetLastError());
  ktimer_end();
  return grid;
}

template <typename T, int LPR, int KA, int NB>
int gram_rows_launch(GramArgs<T>& g) {
  auto kern = k_gram_rows<T, LPR, KA, NB>;
  constexpr int RPB = (kBlock / 64) * (64 / LPR);
  g.ntiles = (g.nrows + RPB - 1) / RPB;
  const size_t lds = sizeof(T) * g.na * g.nb;
  const int grid = persistent_blocks(kern, lds, g.ntiles, rt().gram_blocks_per_cu);
  g.partials = (T*)device_work(sizeof(T) * (size_t)grid * g.na * g.nb);
  ktimer_begin();
  hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), lds, rt().stream, g);
  PHX_HIP(hipGetLastError());
  ktimer_end();
  return grid;
}

// narrow b (nb <= 8), na <= 32, unaligned rows
template <typename T>
int gram_rows(GramArgs<T>& g) {
  if (g.na <= 8) return g.nb <= 4 ? gram_rows_launch<T, 2, 4, 4>(g) : gram_rows_launch<T, 2, 4, 8>(g);
  if (g.na <= 16) return g.nb <= 4 ? gram_rows_launch<T, 4, 4, 4>(g) : gram_rows_launch<T, 4, 4, 8>(g);
  return g.nb <= 4 ? gram_rows_launch<T, 4, 8, 4>(g) : gram_rows_launch<T, 4, 8, 8>(g);
}

template <typename T, int VPL, int KB>
int gram_lpr(GramArgs<T>& g, int lpr) {
  switch (lpr) {
    case 1: return gram_launch_one<T, VPL, 1, KB>(g);
    case 2: return gram_launch_one<T, VPL, 2, KB>(g);
    case 4: return gram_launch_one<T, VPL, 4, KB>(g);
    case 8: return gram_launch_one<T, VPL, 8, KB>(g);
    case 16: return gram_launch_one<T, VPL, 16, KB>(g);
    default: return gram_launch_one<T, VPL, 32, KB>(g);
  }
}

}  // namespace

template <typename T, int TM, int TK, bool SELF, bool PAIR = false>
static int gram_mfma_launch(GramArgs<T>& g) {
  auto kern = k_gram_mfma<T, TM, TK, SELF, PAIR>;
  g.ntiles = (g.nrows + 63) / 64;
  const size_t lds = sizeof(T) * g.na * g.nb;
  const int grid = persistent_blocks(kern, lds, g.ntiles, rt().gram_blocks_per_cu);
  g.partials = (T*)device_work(sizeof(T) * (size_t)grid * g.na * g.nb);
  ktimer_begin();
  hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), lds, rt().stream, g);
  PHX_HIP(hipGetLastError());
  ktimer_end();
  return grid;
}

template <typename T, int KC>
static int gram_sets_launch(GramArgs<T>& g) {
  auto kern = k_gram_mfma_sets<T, KC>;
  constexpr int ROWS = 64 / (KC / 2);
  g.ntiles = (g.nrows + 4 * ROWS - 1) / (4 * ROWS);
  const size_t lds = sizeof(T) * ((size_t)(kBlock / 64) * 3 * 256 + KC * KC);
  const int grid = persistent_blocks(kern, lds, g.ntiles, rt().gram_blocks_per_cu);
  g.partials = (T*)device_work(sizeof(T) * (size_t)grid * KC * KC);
  ktimer_begin();
  hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), lds, rt().stream, g);
  PHX_HIP(hipGetLastError());
  ktimer_end();
  return grid;
}

// self Gram of a narrow block (2, 4 or 8 columns, pairs 2-scalar aligned)
// by row sets; false when the shape does not fit
template <typename T>
static bool gram_self_sets(T alpha, const T* a, int64_t lda, int k, int64_t nrows, T beta, T* Cdev, int ldc) {
  if (!(k == 2 || k == 4 || k == 8) || lda % 2 != 0 || ((uintptr_t)a % (2 * sizeof(T))) != 0) return false;
  GramArgs<T> g;
  g.nrows = nrows;
  g.a = a;
  g.lda = lda;
  g.na = k;
  g.b = a;
  g.ldb = lda;
  g.nb = k;
  const int nparts = k == 2 ? gram_sets_launch<T, 2>(g) : k == 4 ? gram_sets_launch<T, 4>(g) : gram_sets_launch<T, 8>(g);
  hipLaunchKernelGGL(k_gram_finalize<T>, dim3(finalize_blocks(k * k)), dim3(kBlock), 0, rt().stream, g.partials,
                     nparts, k, k, false, alpha, beta, Cdev, ldc);
  PHX_HIP(hipGetLastError());
  return true;
}

// S = a'b on the matrix cores for na, nb <= 32 (zero-padded to 16-column
// tiles), written to C as C = alpha*S + beta*C; Cmirror, when given,
// receives the transposed block from the same partial sums
template <typename T>
static void gram_block_mfma(T alpha, const T* a, int64_t lda, int na, const T* b, int64_t ldb, int nb, int64_t nrows,
                            bool self, T beta, T* Cdev, int ldc, T* Cmirror = nullptr) {
  GramArgs<T> g;
  g.nrows = nrows;
  g.a = a;
  g.lda = lda;
  g.na = na;
  g.b = b;
  g.ldb = ldb;
  g.nb = nb;
  const int tm = (na + 15) / 16, tk = (nb + 15) / 16;
  // paired 16-byte loads: both operands of even width, even leading dimension, aligned
  const bool pair = tm == 2 && (self || tk == 2) && na % 2 == 0 && nb % 2 == 0 && lda % 2 == 0 &&
                    ldb % 2 == 0 && ((uintptr_t)a % (2 * sizeof(T))) == 0 && ((uintptr_t)b % (2 * sizeof(T))) == 0;
  int nparts;
  if (self && tm == 1) nparts = gram_mfma_launch<T, 1, 1, true>(g);
  else if (self && pair) nparts = gram_mfma_launch<T, 2, 2, true, true>(g);
  else if (self) nparts = gram_mfma_launch<T, 2, 2, true>(g);
  else if (tm == 1 && tk == 1) nparts = gram_mfma_launch<T, 1, 1, false>(g);
  else if (tm == 1) nparts = gram_mfma_launch<T, 1, 2, false>(g);
  else if (tk == 1) nparts = gram_mfma_launch<T, 2, 1, false>(g);
  else if (pair) nparts = gram_mfma_launch<T, 2, 2, false, true>(g);
  else nparts = gram_mfma_launch<T, 2, 2, false>(g);
  const int nout = na * nb;
  const int fblocks = finalize_blocks(nout);
  hipLaunchKernelGGL(k_gram_finalize<T>, dim3(fblocks), dim3(kBlock), 0, rt().stream, g.partials, nparts, na, nb,
                     false, alpha, beta, Cdev, ldc);
  PHX_HIP(hipGetLastError());
  if (Cmirror) {  // the transposed block of a self Gram, from the same partials
    hipLaunchKernelGGL(k_gram_finalize<T>, dim3(fblocks), dim3(kBlock), 0, rt().stream, g.partials, nparts, na, nb,
                       true, alpha, beta, Cmirror, ldc);
    PHX_HIP(hipGetLastError());
  }
}

// Gram of the column blocks a (na cols) and b (nb <= 32 cols, nb <= na):
// S = a'b, written to C as C = alpha*S + beta*C (or its transpose); the a
// side is cut into pieces one lane group covers (32 lanes of VPL columns)
template <typename T>
static void gram_block(T alpha, const T* a, int64_t lda, int na, const T* b, int64_t ldb, int nb, int64_t nrows,
                       bool transposed, T beta, T* Cdev, int ldc) {
  constexpr int V16 = 16 / sizeof(T);
  const bool vec = (na % V16 == 0) && (nb % V16 == 0) && (lda % V16 == 0) && (ldb % V16 == 0) && al16(a) && al16(b);
  const int VPL = vec ? V16 : 1;
  const int apiece = 32 * VPL;
  for (int p0 = 0; p0 < na; p0 += apiece) {
    GramArgs<T> g;
    g.nrows = nrows;
    g.a = a + p0;
    g.lda = lda;
    g.na = std::min(apiece, na - p0);
    g.b = b;
    g.ldb = ldb;
    g.nb = nb;
    // the lane group must hold every column of both operands' rows
    const int groups = (std::max(g.na, nb) + VPL - 1) / VPL;
    int lpr = 1;
    while (lpr < groups) lpr <<= 1;
    int nparts;  // partial tiles written (one per workgroup)
    if (VPL == 1 && nb <= 8 && g.na <= 32 && rt().gram_rows) {
      nparts = gram_rows<T>(g);
    } else if (nb <= 8) {
      nparts = VPL == 1 ? gram_lpr<T, 1, 8>(g, lpr) : gram_lpr<T, V16, 8>(g, lpr);
    } else {
      nparts = VPL == 1 ? gram_lpr<T, 1, 32>(g, lpr) : gram_lpr<T, V16, 32>(g, lpr);
    }
    // S rows p0.. of this piece land in C rows (or columns, transposed) p0..
    T* cpiece = transposed ? Cdev + (int64_t)p0 * ldc : Cdev + p0;
    const int nout = g.na * nb;
    const int fblocks = finalize_blocks(nout);
    hipLaunchKernelGGL(k_gram_finalize<T>, dim3(fblocks), dim3(kBlock), 0, rt().stream, g.partials, nparts, g.na, nb,
                       transposed, alpha, beta, cpiece, ldc);
    PHX_HIP(hipGetLastError());
  }
}

template <typename T>
void gemm_gram(T alpha, const Mvec<T>* V, const Mvec<T>* W, T beta, SdMat<T>* C, bool to_host) {
  const int m = V->nvec, k = W->nvec;
  PHX_REQUIRE(V->nrows == W->nrows, PHIST_INVALID_INPUT, "V and W have different lengths");
  PHX_REQUIRE(C->nrows == m && C->ncols == k, PHIST_INVALID_INPUT, "C is %dx%d, V'W is %dx%d", C->nrows, C->ncols, m,
              k);
  const int ldc = C->root->ld;
  // the sliced operand is the wider one; the broadcast side is cut in blocks
  // of at most 32 columns (each block one pass over the rows)
  const bool self = V->root == W->root && V->jofs == W->jofs && m == k;
  const int wide = std::max(m, k);
  const bool strided = V->lda != m || W->lda != k;
  if (rt().gram_mfma && self && rt().gram_sets &&
      gram_self_sets<T>(alpha, V->dev(), V->lda, m, V->nrows, beta, C->dev(), ldc)) {
    if (to_host) sd_from_device(C);
    return;
  }
  if (rt().gram_mfma && (wide >= rt().gram_mfma_wide || (strided && wide >= 4) || (self && wide >= 8))) {
    // matrix cores: 32 x 32 output blocks (the self Gram computes the upper
    // blocks and mirrors them)
    for (int i0 = 0; i0 < m; i0 += 32)
      for (int j0 = 0; j0 < k; j0 += 32) {
        const int mi = std::min(32, m - i0), kj = std::min(32, k - j0);
        T* cblk = C->dev() + (int64_t)j0 * ldc + i0;
        if (self && i0 > j0) continue;
        const bool diag_self = self && i0 == j0;
        // off-diagonal blocks of a self Gram also fill their mirror C(j0.., i0..)
        T* cmir = (self && i0 < j0) ? C->dev() + (int64_t)i0 * ldc + j0 : nullptr;
        gram_block_mfma<T>(alpha, V->dev() + i0, V->lda, mi, W->dev() + j0, W->lda, kj, V->nrows, diag_self, beta,
                           cblk, ldc, cmir);
      }
    if (to_host) sd_from_device(C);
    return;
  }
  const bool swap = k > m;
  const Mvec<T>* A = swap ? W : V;
  const Mvec<T>* B = swap ? V : W;
  const int na = A->nvec, nbt = B->nvec;
  for (int b0 = 0; b0 < nbt; b0 += 32) {
    const int nbc = std::min(32, nbt - b0);
    // output block: all A columns x B columns b0..b0+nbc
    T* cblk = swap ? C->dev() + b0 : C->dev() + (int64_t)b0 * ldc;
    gram_block<T>(alpha, A->dev(), A->lda, na, B->dev() + b0, B->lda, nbc, A->nrows, swap, beta, cblk, ldc);
  }
  if (to_host) sd_from_device(C);
}

// ---------------------------------------------------------------------------
// row-mixing products: W(i,:) = f(W(i,:)) + sum_j alpha*C(j,:) * V(i,j)
// ---------------------------------------------------------------------------
template <typename T>
struct MixArgs {
  int64_t nrows, ntiles;
  const T* v;
  int64_t ldv;
  int k;
  T* w;
  int64_t ldw;
  int m;
  const T* cdev;
  int ldc;
  const T* ddev;
  int ldd;
  T alpha, beta;
  int mode;  // 0: W = alpha*V*C + beta*W, 1: W = W*D + V*C, 2: W = V*C (W may alias V)
  int vvec;  // V and W rows may be read with 16-byte vectors
  // Gram epilogue of the matrix-core kernel: S = G'Wnew (G = gv, ngv columns)
  // or Wnew'Wnew, one partial tile per wave in partials
  const T* gv;
  int64_t ldgv;
  int ngv;
  T* partials;
};

template <typename T, int VPL, int LPR>
__global__ void __launch_bounds__(kBlock) k_rowmix(const MixArgs<T> p) {
  extern __shared__ char smem_raw[];
  T* sC = reinterpret_cast<T*>(smem_raw);  // [k][m] alpha*C(j,c) at j*m + c
  T* sD = sC + p.k * p.m;                  // [m][m] D(j,c) at j*m + c
  for (int o = threadIdx.x; o < p.k * p.m; o += kBlock) {
    const int j = o / p.m, c = o % p.m;
    sC[o] = p.alpha * p.cdev[(int64_t)c * p.ldc + j];
  }
  if (p.mode == 1)
    for (int o = threadIdx.x; o < p.m * p.m; o += kBlock) {
      const int j = o / p.m, c = o % p.m;
      sD[o] = p.ddev[(int64_t)c * p.ldd + j];
    }
  __syncthreads();
  constexpr int RPW = 64 / LPR, RPB = (kBlock / 64) * RPW;
  constexpr int V16 = 16 / sizeof(T);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int q = lane % LPR, rw = lane / LPR, c0 = q * VPL;
  if (c0 >= p.m) return;
  const TileWalk tw = xcd_tiles(p.ntiles);
  for (int64_t t = tw.t; t < tw.end; t += tw.stride) {
    const int64_t i = t * RPB + wv * RPW + rw;
    if (i >= p.nrows) continue;
    T acc[VPL];
    T* wp = p.w + i * p.ldw;
    if (p.mode == 0 && p.beta != T(0)) {
      T wo[VPL];
      vload<T, VPL>(wp + c0, wo);
#pragma unroll
      for (int e = 0; e < VPL; e++) acc[e] = p.beta * wo[e];
    } else {
#pragma unroll
      for (int e = 0; e < VPL; e++) acc[e] = T(0);
    }
    if (p.mode == 1) {  // acc = W(i,:) * D(:, c0..)
      int j = 0;
      if (p.vvec)
        for (; j + V16 <= p.m; j += V16) {
          T wr[V16];
          vload<T, V16>(wp + j, wr);
#pragma unroll
          for (int u = 0; u < V16; u++)
#pragma unroll
            for (int e = 0; e < VPL; e++) acc[e] = fma(wr[u], sD[(j + u) * p.m + c0 + e], acc[e]);
        }
      for (; j < p.m; j += 8) {
        T wr[8];
#pragma unroll
        for (int u = 0; u < 8; u++) wr[u] = j + u < p.m ? wp[j + u] : T(0);
#pragma unroll
        for (int u = 0; u < 8; u++) {
          if (j + u >= p.m) break;
#pragma unroll
          for (int e = 0; e < VPL; e++) acc[e] = fma(wr[u], sD[(j + u) * p.m + c0 + e], acc[e]);
        }
      }
    }
    const T* vp = p.v + i * p.ldv;
    int j = 0;
    if (p.vvec)
      for (; j + V16 <= p.k; j += V16) {
        T vr[V16];
        vload<T, V16>(vp + j, vr);
#pragma unroll
        for (int u = 0; u < V16; u++)
#pragma unroll
          for (int e = 0; e < VPL; e++) acc[e] = fma(sC[(j + u) * p.m + c0 + e], vr[u], acc[e]);
      }
    for (; j < p.k; j += 8) {  // scalar loads, 8 in flight before their products
      T vr[8];
#pragma unroll
      for (int u = 0; u < 8; u++) vr[u] = j + u < p.k ? vp[j + u] : T(0);
#pragma unroll
      for (int u = 0; u < 8; u++) {
        if (j + u >= p.k) break;
#pragma unroll
        for (int e = 0; e < VPL; e++) acc[e] = fma(sC[(j + u) * p.m + c0 + e], vr[u], acc[e]);
      }
    }
    vstore<T, VPL>(wp + c0, acc);
  }
}

// Row-mixing on the matrix cores (modes 0 and 2): a wave forms RT tiles of
// 16 rows x 16*TM columns per iteration as sum_s A_s * B_s over KS 16x4x16
// steps.  The k columns of V are dealt to the four lane groups g = lane/16 as
// column g*KS + s of step s, so a lane reads KS consecutive entries of its
// row and the wave reads each row exactly once; B_s = alpha*C(g*KS + s,
// 16*tm + lane%16) stays in registers for the whole launch.  All loads of a
// tile precede its stores (the stores depend on every MFMA of the tile) and
// each row belongs to one wave, so W may alias V (mode 2).
//
// GRAM = 1 adds Wnew'Wnew (upper 16x16 tiles, mirrored), GRAM = 2 adds
// G'Wnew for a second block vector G of at most 16*TV columns: the new W
// values sit in the MFMA result layout, which is the operand layout of a
// Gram step over the same four rows, so the product needs no extra traffic
// for W.
template <typename T, int KS, int TM, int RT, int GRAM = 0, int TV = 1>
__global__ void __launch_bounds__(kBlock) k_rowmix_mfma(const MixArgs<T> p) {
  using M = Mfma<T>;
  constexpr int TG = GRAM == 1 ? TM : TV;  // tiles of the Gram's left operand
  typename M::acc_t gacc[TG][TM];
  if constexpr (GRAM != 0) {
#pragma unroll
    for (int i = 0; i < TG; i++)
#pragma unroll
      for (int j = 0; j < TM; j++) gacc[i][j] = typename M::acc_t{0, 0, 0, 0};
  }
  constexpr int V16 = 16 / sizeof(T);
  constexpr int ROWS = 16 * RT;  // rows per wave and iteration
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int g = lane >> 4, cc = lane & 15;
  const int j0 = g * KS;
  T bC[KS][TM];
#pragma unroll
  for (int st = 0; st < KS; st++)
#pragma unroll
    for (int tm = 0; tm < TM; tm++) {
      const int j = j0 + st, c = tm * 16 + cc;
      bC[st][tm] = (j < p.k && c < p.m) ? p.alpha * p.cdev[(int64_t)c * p.ldc + j] : T(0);
    }
  const bool vrow = p.vvec && (KS % V16 == 0) && j0 + KS <= p.k;
  // UNR tiles per iteration where registers allow: all their loads are in
  // flight before the first MFMA (in place is still safe: the tiles' rows
  // are disjoint and each row's stores follow its own loads)
  constexpr int UNR = (GRAM == 0 && KS * TM <= 16) ? 2 : 1;
  const TileWalk tw = xcd_tiles(p.ntiles);
  for (int64_t t0 = tw.t; t0 < tw.end; t0 += UNR * tw.stride) {
  T au[UNR][RT][KS];
#pragma unroll
  for (int u = 0; u < UNR; u++) {
    const int64_t tu = t0 + u * tw.stride;
    const int64_t r0 = tu * (4 * ROWS) + wv * ROWS;
    const bool tok = tu < tw.end;
#pragma unroll
    for (int rt_ = 0; rt_ < RT; rt_++) {
      const int64_t ra = r0 + rt_ * 16 + cc;
      const T* vp = p.v + ra * p.ldv + j0;
      if (tok && ra < p.nrows && vrow) {
#pragma unroll
        for (int q = 0; q < KS; q += (KS % V16 == 0 ? V16 : 1)) {
          if constexpr (KS % V16 == 0) {
            T tmp[V16];
            vload<T, V16>(vp + q, tmp);
#pragma unroll
            for (int e = 0; e < V16; e++) au[u][rt_][q + e] = tmp[e];
          }
        }
      } else {
#pragma unroll
        for (int st = 0; st < KS; st++) au[u][rt_][st] = (tok && ra < p.nrows && j0 + st < p.k) ? vp[st] : T(0);
      }
    }
  }
#pragma unroll
  for (int u = 0; u < UNR; u++) {
    const int64_t t = t0 + u * tw.stride;
    if (t >= tw.end) break;
    const int64_t r0 = t * (4 * ROWS) + wv * ROWS;
    T (&a)[RT][KS] = au[u];
    typename M::acc_t acc[RT][TM];
#pragma unroll
    for (int rt_ = 0; rt_ < RT; rt_++)
#pragma unroll
      for (int tm = 0; tm < TM; tm++) {
        acc[rt_][tm] = typename M::acc_t{0, 0, 0, 0};
#pragma unroll
        for (int st = 0; st < KS; st++) acc[rt_][tm] = M::op(a[rt_][st], bC[st][tm], acc[rt_][tm]);
      }
#pragma unroll
    for (int rt_ = 0; rt_ < RT; rt_++) {
#pragma unroll
      for (int tm = 0; tm < TM; tm++) {
        const int c = tm * 16 + cc;
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int64_t row = r0 + rt_ * 16 + M::row(lane, r);
          if (row < p.nrows && c < p.m) {
            T* wp = p.w + row * p.ldw + c;
            T val = acc[rt_][tm][r];
            if (p.mode == 0 && p.beta != T(0)) val = fma(p.beta, *wp, val);
            *wp = val;
            acc[rt_][tm][r] = val;  // the Gram epilogue uses the stored values
          }
        }
      }
      if constexpr (GRAM == 1) {
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
          for (int tg = 0; tg < TM; tg++)
#pragma unroll
            for (int tm = tg; tm < TM; tm++) gacc[tg][tm] = M::op(acc[rt_][tg][r], acc[rt_][tm][r], gacc[tg][tm]);
      } else if constexpr (GRAM == 2) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int64_t row = r0 + rt_ * 16 + M::row(lane, r);
          T g[TV];
#pragma unroll
          for (int tv = 0; tv < TV; tv++) {
            const int c = tv * 16 + cc;
            g[tv] = (row < p.nrows && c < p.ngv) ? p.gv[row * p.ldgv + c] : T(0);
          }
#pragma unroll
          for (int tv = 0; tv < TV; tv++)
#pragma unroll
            for (int tm = 0; tm < TM; tm++) gacc[tv][tm] = M::op(g[tv], acc[rt_][tm][r], gacc[tv][tm]);
        }
      }
    }
  }
  }
  if constexpr (GRAM != 0) {
    // partial tile of this workgroup: S(i, j) at j*na + i
    extern __shared__ char gram_lds[];
    const int na = GRAM == 1 ? p.m : p.ngv, nb = p.m;
    block_tile_sum(reinterpret_cast<T*>(gram_lds), na * nb, p.partials + (int64_t)blockIdx.x * na * nb,
                   [&](auto add) {
#pragma unroll
                     for (int tg = 0; tg < TG; tg++)
#pragma unroll
                       for (int tm = 0; tm < TM; tm++) {
                         if (GRAM == 1 && tg > tm) continue;
#pragma unroll
                         for (int r = 0; r < 4; r++) {
                           const int i = tg * 16 + M::row(lane, r), j = tm * 16 + cc;
                           const T v = gacc[tg][tm][r];
                           if (i < na && j < nb) add(j * na + i, v);
                           if (GRAM == 1 && tg < tm && j < na && i < nb) add(i * na + j, v);
                         }
                       }
                   });
  }
}

namespace {

template <typename T, int VPL, int LPR>
void mix_launch_one(MixArgs<T>& p, size_t lds) {
  auto kern = k_rowmix<T, VPL, LPR>;
  constexpr int RPB = (kBlock / 64) * (64 / LPR);
  p.ntiles = (p.nrows + RPB - 1) / RPB;
  const int grid = persistent_blocks(kern, lds, p.ntiles);
  ktimer_begin();
  hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), lds, rt().stream, p);
  PHX_HIP(hipGetLastError());
  ktimer_end();
}

template <typename T, int VPL>
void mix_lpr(MixArgs<T>& p, int lpr, size_t lds) {
  switch (lpr) {
    case 1: return mix_launch_one<T, VPL, 1>(p, lds);
    case 2: return mix_launch_one<T, VPL, 2>(p, lds);
    case 4: return mix_launch_one<T, VPL, 4>(p, lds);
    case 8: return mix_launch_one<T, VPL, 8>(p, lds);
    case 16: return mix_launch_one<T, VPL, 16>(p, lds);
    case 32: return mix_launch_one<T, VPL, 32>(p, lds);
    default: return mix_launch_one<T, VPL, 64>(p, lds);
  }
}

constexpr size_t kMixLds = 32 * 1024;

// one launch; the caller guarantees the LDS footprint fits and m <= 64*VPL
template <typename T, int KS, int TM>
void mix_mfma_launch(MixArgs<T>& p) {
  constexpr int RT = KS >= 8 ? 1 : 2;
  auto kern = k_rowmix_mfma<T, KS, TM, RT>;
  p.ntiles = (p.nrows + 64 * RT - 1) / (64 * RT);
  const int grid = persistent_blocks(kern, 0, p.ntiles);
  ktimer_begin();
  hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), 0, rt().stream, p);
  PHX_HIP(hipGetLastError());
  ktimer_end();
}

template <typename T, int KS>
void mix_mfma_tm(MixArgs<T>& p) {
  switch ((p.m + 15) / 16) {
    case 1: return mix_mfma_launch<T, KS, 1>(p);
    case 2: return mix_mfma_launch<T, KS, 2>(p);
    case 3: return mix_mfma_launch<T, KS, 3>(p);
    default: return mix_mfma_launch<T, KS, 4>(p);
  }
}

template <typename T, int KS, int TM, int GRAM, int TV>
int mix_gram_launch(MixArgs<T>& p) {
  constexpr int RT = KS >= 8 ? 1 : 2;
  auto kern = k_rowmix_mfma<T, KS, TM, RT, GRAM, TV>;
  p.ntiles = (p.nrows + 64 * RT - 1) / (64 * RT);
  const int na = GRAM == 1 ? p.m : p.ngv;
  const size_t lds = sizeof(T) * na * p.m;
  const int grid = persistent_blocks(kern, lds, p.ntiles, rt().gram_blocks_per_cu);
  p.partials = (T*)device_work(sizeof(T) * (size_t)grid * na * p.m);
  ktimer_begin();
  hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), lds, rt().stream, p);
  PHX_HIP(hipGetLastError());
  ktimer_end();
  return grid;
}

template <typename T, int KS, int TM>
int mix_gram_tv(MixArgs<T>& p, bool self) {
  if (self) return mix_gram_launch<T, KS, TM, 1, 1>(p);
  if (p.ngv <= 16) return mix_gram_launch<T, KS, TM, 2, 1>(p);
  if (p.ngv <= 32) return mix_gram_launch<T, KS, TM, 2, 2>(p);
  if (p.ngv <= 48) return mix_gram_launch<T, KS, TM, 2, 3>(p);
  return mix_gram_launch<T, KS, TM, 2, 4>(p);
}

template <typename T, int KS>
int mix_gram_tm(MixArgs<T>& p, bool self) {
  if (p.m <= 16) return mix_gram_tv<T, KS, 1>(p, self);
  return mix_gram_tv<T, KS, 2>(p, self);
}

// modes 0 and 2 with 8 <= k, m <= 64 go to the matrix cores
template <typename T>
bool mix_mfma_ok(int mode, int k, int m) {
  return rt().mix_mfma && mode != 1 && k >= 8 && m >= rt().mix_mfma_min_m && k <= 64 && m <= 64;
}

template <typename T>
void mix_mfma(MixArgs<T>& p) {
  const int ks = (p.k + 3) / 4;
  if (ks <= 2) mix_mfma_tm<T, 2>(p);
  else if (ks <= 4) mix_mfma_tm<T, 4>(p);
  else if (ks <= 8) mix_mfma_tm<T, 8>(p);
  else mix_mfma_tm<T, 16>(p);
}

template <typename T>
void rowmix(int mode, T alpha, const T* v, int64_t ldv, int k, const T* cdev, int ldc, T beta, T* w, int64_t ldw,
            int m, const T* ddev, int ldd, int64_t nrows) {
  constexpr int V16 = 16 / sizeof(T);
  const bool vvec = (ldv % V16 == 0) && al16(v) && (mode != 1 || ((ldw % V16 == 0) && al16(w)));
  const bool ovec = (m % V16 == 0) && (ldw % V16 == 0) && al16(w);
  const int VPL = ovec ? V16 : 1;
  MixArgs<T> p;
  p.nrows = nrows;
  p.v = v;
  p.ldv = ldv;
  p.k = k;
  p.w = w;
  p.ldw = ldw;
  p.m = m;
  p.cdev = cdev;
  p.ldc = ldc;
  p.ddev = ddev;
  p.ldd = ldd;
  p.alpha = alpha;
  p.beta = beta;
  p.mode = mode;
  p.vvec = vvec ? 1 : 0;
  if (mix_mfma_ok<T>(mode, k, m)) return mix_mfma<T>(p);
  const size_t lds = sizeof(T) * ((size_t)k * m + (mode == 1 ? (size_t)m * m : 0));
  PHX_REQUIRE(m <= 64 * VPL && lds <= kMixLds, PHIST_FUNCTIONAL_ERROR, "row-mix launch of %d x %d exceeds one pass", k,
              m);
  const int groups = (m + VPL - 1) / VPL;
  int lpr = 1;
  while (lpr < groups) lpr <<= 1;
  if (VPL == 1)
    mix_lpr<T, 1>(p, lpr, lds);
  else
    mix_lpr<T, V16>(p, lpr, lds);
}

// one launch: at most 64 output columns (64 lane groups of >= 1 column) and
// the LDS copies of C (and D) within kMixLds
template <typename T>
bool mix_fits(int k, int m, bool with_d) {
  return m <= 64 && sizeof(T) * ((size_t)k * m + (with_d ? (size_t)m * m : 0)) <= kMixLds;
}

}  // namespace

// One sweep for W = alpha*V*C + beta*W (mode 0) or W(:,1:m) = W*C in place
// (mode 2, V == W) together with D = ga*S + gb*D, S = Wnew'Wnew (G == nullptr)
// or G'Wnew.  Returns false when the shapes are outside the fused kernel
// (8 <= k <= 64, 8 <= m <= 32, G at most 32 columns, G not overlapping W).
template <typename T>
bool mix_gram(int mode, T alpha, const Mvec<T>* V, const SdMat<T>* C, T beta, Mvec<T>* W, const Mvec<T>* G, T ga,
              T gb, SdMat<T>* D, bool to_host = true, bool narrow = false) {
  // narrow: also 1..7 output columns and a Gram side G of up to 64 columns
  const int k = V->nvec, m = C->ncols;
  if (!rt().mix_mfma || k < 8 || k > 64 || m < (narrow ? 1 : 8) || m > 32) return false;
  if (G && (G->nvec > (narrow ? 64 : 32) || G->nrows != W->nrows)) return false;
  if (G && G->root == W->root && !(G->jofs + G->nvec <= W->jofs || W->jofs + m <= G->jofs)) return false;
  constexpr int V16 = 16 / sizeof(T);
  MixArgs<T> p{};
  p.nrows = V->nrows;
  p.v = V->dev();
  p.ldv = V->lda;
  p.k = k;
  p.w = W->dev();
  p.ldw = W->lda;
  p.m = m;
  p.cdev = C->dev();
  p.ldc = C->root->ld;
  p.alpha = alpha;
  p.beta = beta;
  p.mode = mode;
  p.vvec = (p.ldv % V16 == 0) && al16(p.v) ? 1 : 0;
  p.gv = G ? G->dev() : nullptr;
  p.ldgv = G ? G->lda : 0;
  p.ngv = G ? G->nvec : 0;
  const int ks = (k + 3) / 4;
  int nparts;
  if (ks <= 2) nparts = mix_gram_tm<T, 2>(p, !G);
  else if (ks <= 4) nparts = mix_gram_tm<T, 4>(p, !G);
  else if (ks <= 8) nparts = mix_gram_tm<T, 8>(p, !G);
  else nparts = mix_gram_tm<T, 16>(p, !G);
  const int na = G ? G->nvec : m, nout = na * m;
  hipLaunchKernelGGL(k_gram_finalize<T>, dim3(finalize_blocks(nout)), dim3(kBlock), 0, rt().stream, p.partials, nparts, na,
                     m, false, ga, gb, D->dev(), D->root->ld);
  PHX_HIP(hipGetLastError());
  if (to_host) sd_from_device(D);
  return true;
}

template <typename T>
void gemm_times_sd(T alpha, const Mvec<T>* V, const SdMat<T>* C, T beta, Mvec<T>* W) {
  const int k = V->nvec, m = W->nvec;
  PHX_REQUIRE(V->nrows == W->nrows, PHIST_INVALID_INPUT, "V and W have different lengths");
  PHX_REQUIRE(C->nrows == k && C->ncols == m, PHIST_INVALID_INPUT, "C is %dx%d, expected %dx%d", C->nrows, C->ncols, k,
              m);
  if (V->root == W->root)
    PHX_REQUIRE(V->jofs + k <= W->jofs || W->jofs + m <= V->jofs, PHIST_INVALID_INPUT, "V and W overlap in memory");
  const int ldc = C->root->ld;
  // output columns in blocks of 64; k in blocks (later blocks accumulate with beta=1)
  const int mb = std::min(m, 64);
  for (int c0 = 0; c0 < m; c0 += mb) {
    const int mc = std::min(mb, m - c0);
    int kb = std::max(1, (int)(kMixLds / sizeof(T)) / std::max(1, mc));
    if (rt().mix_mfma && mc >= rt().mix_mfma_min_m && k > 64) kb = std::min(kb, 64);  // k blocks the matrix-core kernel takes
    for (int j0 = 0; j0 < k || (k == 0 && j0 == 0); j0 += kb) {
      const int kc = std::min(kb, k - j0);
      const T b = j0 == 0 ? beta : T(1);
      rowmix<T>(0, alpha, V->dev() + j0, V->lda, kc, C->dev() + (int64_t)c0 * ldc + j0, ldc, b, W->dev() + c0, W->lda,
                mc, nullptr, 0, V->nrows);
      if (k == 0) break;
    }
  }
}

template <typename T>
void gemm_times_sd_inplace(Mvec<T>* V, const SdMat<T>* M) {
  const int k = V->nvec, m = M->ncols;
  PHX_REQUIRE(M->nrows == k && m <= k, PHIST_INVALID_INPUT, "M is %dx%d, V has %d columns", M->nrows, M->ncols, k);
  if (mix_fits<T>(k, m, false)) {
    rowmix<T>(2, T(1), V->dev(), V->lda, k, M->dev(), M->root->ld, T(0), V->dev(), V->lda, m, nullptr, 0, V->nrows);
    return;
  }
  // too large for one pass: go through a copy of V
  Mvec<T> tmp;
  tmp.map = V->map;
  tmp.nrows = V->nrows;
  tmp.nvec = k;
  tmp.lda = k;
  PHX_HIP(hipMallocAsync((void**)&tmp.dev_base, sizeof(T) * std::max<int64_t>(1, V->nrows * k), rt().stream));
  blas_copy<T>(V, &tmp);
  Mvec<T> out = *V;
  out.root = V->root;
  out.nvec = m;
  out.jofs = V->jofs;
  gemm_times_sd<T>(T(1), &tmp, M, T(0), &out);
  PHX_HIP(hipFreeAsync(tmp.dev_base, rt().stream));
}

template <typename T>
void gemm_times_sd_add_times_sd(const Mvec<T>* V, const SdMat<T>* C, Mvec<T>* W, const SdMat<T>* D) {
  const int k = V->nvec, m = W->nvec;
  PHX_REQUIRE(V->nrows == W->nrows, PHIST_INVALID_INPUT, "V and W have different lengths");
  PHX_REQUIRE(C->nrows == k && C->ncols == m && D->nrows == m && D->ncols == m, PHIST_INVALID_INPUT,
              "sdMat shapes do not match V (%d cols) and W (%d cols)", k, m);
  PHX_REQUIRE(mix_fits<T>(k, m, true), PHIST_NOT_IMPLEMENTED, "W*D + V*C with %d x %d blocks exceeds one pass", k, m);
  rowmix<T>(1, T(1), V->dev(), V->lda, k, C->dev(), C->root->ld, T(0), W->dev(), W->lda, m, D->dev(), D->root->ld,
            V->nrows);
}

template <typename T>
void gemm_fused_mvsdi_mvTmv(T alpha, const Mvec<T>* V, Mvec<T>* W, const SdMat<T>* C, T beta, SdMat<T>* D) {
  const int k = W->nvec, m = C->ncols;
  PHX_REQUIRE(C->nrows == k && m <= k, PHIST_INVALID_INPUT, "C is %dx%d, W has %d columns", C->nrows, m, k);
  PHX_REQUIRE(V->nrows == W->nrows, PHIST_INVALID_INPUT, "V and W have different lengths");
  PHX_REQUIRE(D->nrows == V->nvec && D->ncols == m, PHIST_INVALID_INPUT, "D is %dx%d, expected %dx%d", D->nrows,
              D->ncols, V->nvec, m);
  // one sweep: V == W (the self product of the new block) or V apart from W
  const bool self = V->root == W->root && V->jofs == W->jofs && V->nvec == W->nvec;
  if (self ? m == k && mix_gram<T>(2, T(1), W, C, T(0), W, nullptr, alpha, beta, D)
           : mix_gram<T>(2, T(1), W, C, T(0), W, V, alpha, beta, D))
    return;
  gemm_times_sd_inplace<T>(W, C);
  Mvec<T> wv = *W;
  wv.root = W->root;
  wv.nvec = C->ncols;
  gemm_gram<T>(alpha, V, &wv, beta, D);
}

template <typename T>
void gemm_fused_mvsd_mvTmv(T alpha, const Mvec<T>* V, const SdMat<T>* C, T beta, Mvec<T>* W, SdMat<T>* WtW) {
  const int k = V->nvec, m = W->nvec;
  PHX_REQUIRE(C->nrows == k && C->ncols == m, PHIST_INVALID_INPUT, "C is %dx%d, expected %dx%d", C->nrows, C->ncols, k,
              m);
  PHX_REQUIRE(WtW->nrows == m && WtW->ncols == m, PHIST_INVALID_INPUT, "WtW must be %d x %d", m, m);
  PHX_REQUIRE(V->nrows == W->nrows, PHIST_INVALID_INPUT, "V and W have different lengths");
  if (V->root == W->root)
    PHX_REQUIRE(V->jofs + k <= W->jofs || W->jofs + m <= V->jofs, PHIST_INVALID_INPUT, "V and W overlap in memory");
  if (mix_gram<T>(0, alpha, V, C, beta, W, nullptr, T(1), T(0), WtW)) return;
  gemm_times_sd<T>(alpha, V, C, beta, W);
  gemm_gram<T>(T(1), W, W, T(0), WtW);
}

// Cholesky factor of a small Gram matrix and its inverse in LDS: column j of
// R (upper) from G(0:j, j) and the columns before it, rows of a column in
// parallel; then each thread back-substitutes one column of R^-1
template <typename T>
__global__ void __launch_bounds__(64) k_chol_inv(const T* __restrict__ G, int ldg, int k, T* __restrict__ Ri, int ldr,
                                                 T* __restrict__ flag) {
  __shared__ T R[64][65];
  __shared__ int bad;
  const int t = threadIdx.x;
  T gmax = T(0);
  for (int j = 0; j < k; j++) {
    if (t <= j) R[t][j] = G[(int64_t)j * ldg + t];
    const T gjj = G[(int64_t)j * ldg + j];
    gmax = gjj > gmax ? gjj : gmax;
  }
  if (t == 0) bad = 0;
  __syncthreads();
  const T tiny = T(64) * std::numeric_limits<T>::epsilon() * gmax;
  for (int j = 0; j < k; j++) {
    // pivot: R(j,j) = sqrt(G(j,j) - sum_p<j R(p,j)^2), then R(j, i) for i > j
    if (t == 0) {
      T d = R[j][j];
      for (int p = 0; p < j; p++) d -= R[p][j] * R[p][j];
      if (!(d > tiny)) bad = 1;
      R[j][j] = d > T(0) ? sqrt(d) : T(1);
    }
    __syncthreads();
    if (t > j && t < k) {
      T s = R[j][t];
      for (int p = 0; p < j; p++) s -= R[p][j] * R[p][t];
      R[j][t] = s / R[j][j];
    }
    __syncthreads();
  }
  const bool fail = bad != 0;
  if (t < k) {  // column t of R^-1 (R X(:, t) = e_t), back-substituted in place in Ri
    T* x = Ri + (int64_t)t * ldr;
    for (int i = k - 1; i >= 0; i--) {
      if (i > t) {
        x[i] = T(0);
        continue;
      }
      T s = (i == t) ? T(1) : T(0);
      for (int p = i + 1; p <= t; p++) s -= R[i][p] * x[p];
      x[i] = s / R[i][i];
    }
    if (fail)
      for (int i = 0; i < k; i++) x[i] = i == t ? T(1) : T(0);
  }
  if (t == 0) flag[0] = fail ? T(1) : T(0);
}

template <typename T>
void chol_inv_dev(const SdMat<T>* G, SdMat<T>* Rinv, T* d_flag) {
  const int k = G->nrows;
  PHX_REQUIRE(k == G->ncols && Rinv->nrows == k && Rinv->ncols == k && k >= 1 && k <= 64, PHIST_INVALID_INPUT,
              "device Cholesky of a %dx%d Gram matrix", G->nrows, G->ncols);
  hipLaunchKernelGGL(k_chol_inv<T>, dim3(1), dim3(64), 0, rt().stream, G->dev(), G->root->ld, k, Rinv->dev(),
                     Rinv->root->ld, d_flag);
  PHX_HIP(hipGetLastError());
}

// W -= V*C, then D = V'W with the new W, in one sweep over V where the
// shapes allow (the Gram reads the V rows the row mix has just loaded); D
// stays on the device
template <typename T>
void gemm_sub_gram_dev(const Mvec<T>* V, const SdMat<T>* C, Mvec<T>* W, SdMat<T>* D) {
  PHX_REQUIRE(V->nrows == W->nrows && C->nrows == V->nvec && C->ncols == W->nvec && D->nrows == V->nvec &&
                  D->ncols == W->nvec,
              PHIST_INVALID_INPUT, "shapes do not match");
  if (V->root == W->root)
    PHX_REQUIRE(V->jofs + V->nvec <= W->jofs || W->jofs + W->nvec <= V->jofs, PHIST_INVALID_INPUT,
                "V and W overlap in memory");
  if (mix_gram<T>(0, T(-1), V, C, T(1), W, V, T(1), T(0), D, false, true)) return;
  gemm_times_sd<T>(T(-1), V, C, T(1), W);
  gemm_gram<T>(T(1), V, W, T(0), D, false);
}

#define PHX_INST(T)                                                                                  \
  template void chol_inv_dev<T>(const SdMat<T>*, SdMat<T>*, T*);                                     \
  template void gemm_sub_gram_dev<T>(const Mvec<T>*, const SdMat<T>*, Mvec<T>*, SdMat<T>*);          \
  template void gemm_gram<T>(T, const Mvec<T>*, const Mvec<T>*, T, SdMat<T>*, bool);                       \
  template void gemm_times_sd<T>(T, const Mvec<T>*, const SdMat<T>*, T, Mvec<T>*);                   \
  template void gemm_times_sd_inplace<T>(Mvec<T>*, const SdMat<T>*);                                 \
  template void gemm_times_sd_add_times_sd<T>(const Mvec<T>*, const SdMat<T>*, Mvec<T>*, const SdMat<T>*); \
  template void gemm_fused_mvsdi_mvTmv<T>(T, const Mvec<T>*, Mvec<T>*, const SdMat<T>*, T, SdMat<T>*); \
  template void gemm_fused_mvsd_mvTmv<T>(T, const Mvec<T>*, const SdMat<T>*, T, Mvec<T>*, SdMat<T>*);
PHX_INST(double)
PHX_INST(float)
#undef PHX_INST

}  // namespace phx
