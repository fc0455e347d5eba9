// spmv.hip -- SELL-C-sigma sparse matrix times row-major block vector.
//
// Computes, for every matrix row i and block column c (reference order of
// spmvm_kernels.F90:261-271):
//     tmp = shift_c * x(i,c);  tmp += val_j * x(col_j, c)  (row entries in
//     ascending column order);  y(i,c) = alpha*tmp + beta*y(i,c)
// with beta == 0 meaning y is not read (the non-temporal builtin path,
// spmvm_kernels_nt.c:103-126), plus optional epilogues fused into the same
// pass: column dots <y_new, y_new> and <x, y_new> (fused_spmv_mvdot,
// phist_kernels_fused_decl.h:56) and U = gamma*y_new + delta*U
// (fused_spmv_mvdot_mvadd, :77), or the Gram matrices W'W and V'W of the
// new block (fused_spmv_mvTmv, :68) on the matrix cores.
//
// Mapping onto CDNA4: a row is owned by LPR lanes of one wave64, each lane
// holding VPL consecutive block columns (16-byte vectors where the layout
// allows), so a wave covers 64/LPR consecutive SELL rows and the x gather of
// one nonzero is one contiguous LPR*VPL*sizeof(T) segment.  val/col are read
// coalesced from the column-major chunk layout with non-temporal loads, so
// the L2 keeps the x rows being gathered.  The grid is persistent and
// XCD-aware (amd_device.hpp).  Dots are reduced per wave with shuffles, per
// workgroup through LDS, and across workgroups by a second fixed-order pass.
#include "amd_device.hpp"
#include <algorithm>

namespace phx {

template <typename T>
struct SpmvArgs {
  const int64_t* __restrict__ chunk_ptr;
  const int32_t* __restrict__ chunk_w;
  const int32_t* __restrict__ perm;
  const int32_t* __restrict__ col;
  const T* __restrict__ val;
  int clog2;
  int64_t nslots;
  int64_t ntiles;
  const T* __restrict__ x;
  int64_t ldx;
  T* __restrict__ y;
  int64_t ldy;
  T* __restrict__ u;
  int64_t ldu;
  int ncols;  // block columns handled by this launch
  int nact;   // lanes of columns >= nact skip their loads and stores (same kernel, grid and sums)
  T alpha, beta, gamma, delta;
  int flags;
  int blocks_per_cu;  // host side: persistent grid size per CU (0 = as many as fit)
  T* __restrict__ partials;  // [gridDim.x][2][ncols]
  const T* __restrict__ shift_dev;  // shifts on the device (this launch's columns), else shift[]
  const double* __restrict__ xscale;  // per-column factors of x (XSC kernels), stride xscale_stride
  int64_t xscale_stride;
  // staged kernels: the tile's column intervals and each slot's window row
  // (SellStage), 16-byte pieces per window row (st_ppr) and log2 of the
  // pieces a window row occupies in LDS (st_pshift, a power of two >= st_ppr)
  const int32_t* __restrict__ st_ivptr;
  const int32_t* __restrict__ st_iv;
  const uint16_t* __restrict__ st_lcol;
  int st_ppr, st_pshift;
  size_t st_lds_bytes;  // host side: dynamic LDS of the staged window
  T shift[64 * 4];
};

// Workgroups per CU by row length and block width, from a measured sweep
// (1..8 per CU on MI355X; FP64, nvec 1..16; banded 32-nnz rows, XXZ chain,
// 3D Laplacian, honeycomb sheets): long rows gathered from a wide band run
// best with few workgroups (the gathered x rows stay in L2), short rows need
// more waves in flight to cover the gather latency.
inline int spmv_table_blocks_per_cu(double nnz_per_row, int nvec) {
  if (nnz_per_row >= 24.0) return nvec >= 4 ? 2 : 3;
  if (nnz_per_row >= 16.0) return 4;
  return 6;
}

enum : int {
  F_SHIFT = 1, F_BETA = 2, F_DWW = 4, F_DVW = 8, F_AXPY = 16, F_GWW = 32, F_GVW = 64, F_NOSTORE = 128,
  F_NTY = 256  // y written with non-temporal stores
};

// v_mfma_{f64,f32}_16x16x4: lane l supplies A[l%16][l/16] and B[l/16][l%16]
// and holds four results of the 16x16 product in column l%16; the row of
// result register r is l/16 + 4r for f64 and 4*(l/16) + r for f32 (the
// layout the Gram kernels of tsgemm.hip use).
template <typename T>
struct SpmvMfma;
template <>
struct SpmvMfma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  __device__ static acc_t op(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
  __device__ static int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct SpmvMfma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  __device__ static acc_t op(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
  __device__ static int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// orders the LDS accesses of one wave (its lanes exchange data through LDS;
// a wave's LDS instructions complete in order)
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Broadcast of lane `src` (0 <= src < LPR) of every LPR-lane group: DPP
// quad permutes for groups of 2 and 4 lanes (VALU, no LDS traffic), a
// cross-lane shuffle for wider groups.
template <int LPR>
__device__ inline int group_bcast(int v, int src, int lane) {
  if constexpr (LPR == 1) {
    return v;
  } else if constexpr (LPR == 2) {
    return src == 0 ? __builtin_amdgcn_mov_dpp(v, 0xA0, 0xF, 0xF, false)   // quad_perm [0,0,2,2]
                    : __builtin_amdgcn_mov_dpp(v, 0xF5, 0xF, 0xF, false);  // quad_perm [1,1,3,3]
  } else if constexpr (LPR == 4) {
    switch (src) {
      case 0: return __builtin_amdgcn_mov_dpp(v, 0x00, 0xF, 0xF, false);
      case 1: return __builtin_amdgcn_mov_dpp(v, 0x55, 0xF, 0xF, false);
      case 2: return __builtin_amdgcn_mov_dpp(v, 0xAA, 0xF, 0xF, false);
      default: return __builtin_amdgcn_mov_dpp(v, 0xFF, 0xF, 0xF, false);
    }
  } else {
    return __shfl(v, (lane & ~(LPR - 1)) + src, 64);
  }
}
template <int LPR, typename T>
__device__ inline T group_bcast_t(T v, int src, int lane) {
  if constexpr (sizeof(T) == 4) {
    return __int_as_float(group_bcast<LPR>(__float_as_int(v), src, lane));
  } else {
    int2 p = *reinterpret_cast<int2*>(&v);
    p.x = group_bcast<LPR>(p.x, src, lane);
    p.y = group_bcast<LPR>(p.y, src, lane);
    return *reinterpret_cast<T*>(&p);
  }
}

// One row per LPR lanes; the lanes of a row load the row's val/col
// cooperatively (lane q takes slot j = jb*LPR + q, coalesced over the rows
// of the chunk) and share them by group broadcast, so every index is read
// from memory once.  A batch covers SL entries of each row (the depth: 8, or
// 32 for long rows): its KB = SL/LPR val/col loads per lane, then all its x
// gathers, go out before the first FMA of the batch, so a row of 32 entries
// costs one round trip for its indices and one for its gathers.
//
// STG: before a tile's rows, the workgroup copies the x rows the tile
// gathers from (its column intervals, each taken by one wave with 16-byte
// loads) into LDS, and the gathers read that window at each slot's window
// row (st_lcol, 16 bits, read instead of the 32-bit column).  The values
// gathered and their order are the same, so are the results.
//
// GR > 0 (fused_spmv_mvTmv, at most 32 columns, GR = 16-column tiles): after
// each tile the wave writes its rows of the new W (and of V) to LDS and
// reads them back in the matrix-core operand layout, four rows per MFMA
// step, accumulating W'W (upper tiles, mirrored at the end) and V'W in
// registers; the wave tiles are added in LDS in wave order and the
// workgroup tiles by a second fixed-order pass, so the sums are
// deterministic.  W itself is not stored with F_NOSTORE (W = NULL).
template <typename T, int VPL, int LPR, bool XSC, int GR, int SL, bool STG = false>
__global__ void __launch_bounds__(kBlock) k_sell_spmv(const SpmvArgs<T> a) {
  static_assert(!STG || (VPL * sizeof(T) >= 16 && !XSC && GR == 0), "staged windows hold whole 16-byte vectors");
  constexpr int RPW = 64 / LPR;
  constexpr int RPB = (kBlock / 64) * RPW;
  constexpr int KB = SL > LPR ? SL / LPR : 1;  // val/col loads per lane and batch
  constexpr int BATCH = KB * LPR;              // slots per batch and row
  // gathers issued before their FMAs: up to 32 at depth 32 with at most 16
  // gathered bytes per lane, otherwise 16 (register budget)
  constexpr int SUBMAX = (SL >= 32 && VPL * (int)sizeof(T) <= 16) ? 32 : 16;
  constexpr int SUB = BATCH < SUBMAX ? BATCH : SUBMAX;  // gathers issued before their FMAs
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int q = lane % LPR;
  const int rw = lane / LPR;
  const int c0 = q * VPL;
  const bool colok = c0 < a.nact;
  const int C = 1 << a.clog2;
  const int flags = a.flags;

  T sh[VPL], xsc[VPL];
#pragma unroll
  for (int e = 0; e < VPL; e++) {
    sh[e] = (colok && (flags & F_SHIFT)) ? (a.shift_dev ? a.shift_dev[c0 + e] : a.shift[c0 + e]) : T(0);
    if constexpr (XSC) xsc[e] = colok ? (T)a.xscale[(int64_t)(c0 + e) * a.xscale_stride] : T(0);
  }
  T dww[VPL], dvw[VPL];
#pragma unroll
  for (int e = 0; e < VPL; e++) dww[e] = dvw[e] = T(0);
  using GM = SpmvMfma<T>;
  constexpr int TG = GR > 0 ? GR : 1;  // 16-column tiles of the Gram matrices
  constexpr int GW = LPR * VPL;        // row stride of the exchange tiles
  __shared__ T gex[GR > 0 ? kBlock / 64 : 1][2][GR > 0 ? 64 * VPL : 1];
  typename GM::acc_t gww[TG][TG], gvw[TG][TG];
#pragma unroll
  for (int i = 0; i < TG; i++)
#pragma unroll
    for (int j = 0; j < TG; j++) gww[i][j] = gvw[i][j] = typename GM::acc_t{0, 0, 0, 0};

  extern __shared__ __align__(16) unsigned char spmv_window[];
  const T* win_x = reinterpret_cast<const T*>(spmv_window) + c0;  // this lane's columns of window row 0
  const int win_ld = (16 / (int)sizeof(T)) << a.st_pshift;        // scalars per window row

  const TileWalk tw = xcd_tiles(a.ntiles);
  for (int64_t t = tw.t; t < tw.end; t += tw.stride) {
    if constexpr (STG) {
      // the tile's window: wave w copies intervals w, w+4, ...; piece p of an
      // interval is piece p & (2^pshift - 1) of its row p >> pshift
      __syncthreads();  // every wave is done with the previous tile's window
      const int k0 = a.st_ivptr[t], k1 = a.st_ivptr[t + 1];
      const int ps = a.st_pshift, ppr = a.st_ppr, pmask = (1 << ps) - 1;
      typedef unsigned int piece_t __attribute__((ext_vector_type(4)));
      piece_t* win = reinterpret_cast<piece_t*>(spmv_window);
      for (int k = k0 + wv; k < k1; k += kBlock / 64) {
        const int32_t src = a.st_iv[3 * k], off = a.st_iv[3 * k + 1], len = a.st_iv[3 * k + 2];
        const int n = len << ps;
        const unsigned char* xs = reinterpret_cast<const unsigned char*>(a.x + (int64_t)src * a.ldx);
        const int64_t ldb = a.ldx * (int64_t)sizeof(T);
        for (int p0 = 0; p0 < n; p0 += 64 * 8) {
          piece_t v[8];
#pragma unroll
          for (int u = 0; u < 8; u++) {
            const int p = p0 + u * 64 + lane;
            v[u] = piece_t{0, 0, 0, 0};
            if (p < n && (p & pmask) < ppr)
              v[u] = *reinterpret_cast<const piece_t*>(xs + (int64_t)(p >> ps) * ldb + 16 * (p & pmask));
          }
#pragma unroll
          for (int u = 0; u < 8; u++) {
            const int p = p0 + u * 64 + lane;
            if (p < n && (p & pmask) < ppr) win[(off << ps) + p] = v[u];
          }
        }
      }
      __syncthreads();
    }
    // all lanes stay in the loop body (group broadcasts need every lane)
    const int64_t s = t * RPB + wv * RPW + rw;
    const bool sok = s < a.nslots;
    const int32_t row = sok ? a.perm[s] : -1;
    const bool live = row >= 0;
    const int64_t ch = s >> a.clog2;
    const int r = (int)(s & (C - 1));
    const int w = sok ? a.chunk_w[ch] : 0;
    const int64_t base = sok ? a.chunk_ptr[ch] + r : 0;
    // loop bound: the widest chunk among the wave's rows
    int wmax = w;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) wmax = max(wmax, __shfl_xor(wmax, off, 64));
    const T* __restrict__ xb = a.x + c0;

    T acc[VPL], xo[VPL];
#pragma unroll
    for (int e = 0; e < VPL; e++) xo[e] = T(0);
    if (live && colok && (flags & (F_SHIFT | F_DVW | F_GVW))) vload<T, VPL>(xb + (int64_t)row * a.ldx, xo);
#pragma unroll
    for (int e = 0; e < VPL; e++) acc[e] = (flags & F_SHIFT) ? sh[e] * xo[e] : T(0);

    // x row cc of this lane's columns: from the window (STG: cc is a window row) or from x
    auto gather = [&](int32_t cc, T (&out)[VPL]) {
      if constexpr (STG)
        vload<T, VPL>(win_x + cc * win_ld, out);
      else
        vload<T, VPL>(xb + (int64_t)cc * a.ldx, out);
    };
    // the column (STG: window row) of a slot
    auto load_col = [&](int64_t slot) -> int32_t {
      if constexpr (STG)
        return (int32_t)a.st_lcol[slot];
      else
        return __builtin_nontemporal_load(a.col + slot);
    };
    // the batch's x gathers, then its FMAs in entry order: slot j0+k of this
    // row sits in lane k % LPR of the group, register k / LPR
    auto spmv_batch = [&](int j0, const T (&v)[KB], const int32_t (&cj)[KB]) {
#pragma unroll
      for (int k0 = 0; k0 < BATCH; k0 += SUB) {
        T vv[SUB];
        T xg[SUB][VPL];
        bool ok[SUB];
#pragma unroll
        for (int kk = 0; kk < SUB; kk++) {
          const int k = k0 + kk;
          const int32_t cc = group_bcast<LPR>(cj[k / LPR], k % LPR, lane);
          vv[kk] = group_bcast_t<LPR>(v[k / LPR], k % LPR, lane);
          ok[kk] = colok && (j0 + k) < w;
          if (ok[kk]) gather(cc, xg[kk]);
        }
#pragma unroll
        for (int kk = 0; kk < SUB; kk++)
          if (ok[kk]) {
#pragma unroll
            for (int e = 0; e < VPL; e++) acc[e] = fma(vv[kk], xg[kk][e], acc[e]);
          }
      }
    };
    // val/col of the batch starting at entry j0 (0 / column 0 beyond the row)
    auto load_batch = [&](int j0, T (&v)[KB], int32_t (&cj)[KB]) {
#pragma unroll
      for (int u = 0; u < KB; u++) {
        const int j = j0 + u * LPR + q;
        if (j < w) {
          const int64_t slot = base + (int64_t)j * C;
          v[u] = __builtin_nontemporal_load(a.val + slot);
          cj[u] = load_col(slot);
        } else {
          v[u] = T(0);
          cj[u] = 0;
        }
      }
    };
    // depth >= 16: the next batch's val/col are requested before this
    // batch's gathers, so the index loads overlap the gathers; depth 8
    // loads each batch's indices, then gathers
    if constexpr (SL >= 16) {
      T v[KB];
      int32_t cj[KB];
      if (wmax > 0) load_batch(0, v, cj);
      for (int j0 = 0; j0 < wmax; j0 += BATCH) {
        T vn[KB];
        int32_t cn[KB];
        const bool more = j0 + BATCH < wmax;  // uniform over the wave
        if (more) load_batch(j0 + BATCH, vn, cn);
        spmv_batch(j0, v, cj);
        if (more) {
#pragma unroll
          for (int u = 0; u < KB; u++) {
            v[u] = vn[u];
            cj[u] = cn[u];
          }
        }
      }
    } else {
      for (int j0 = 0; j0 < wmax; j0 += BATCH) {
        T v[KB];
        int32_t cj[KB];
#pragma unroll
        for (int u = 0; u < KB; u++) {
          const int j = j0 + u * LPR + q;
          if (j < w) {
            const int64_t slot = base + (int64_t)j * C;
            v[u] = __builtin_nontemporal_load(a.val + slot);
            cj[u] = load_col(slot);
          } else {
            v[u] = T(0);
            cj[u] = 0;
          }
        }
#pragma unroll
        for (int k0 = 0; k0 < BATCH; k0 += SUB) {
          T vv[SUB];
          T xg[SUB][VPL];
          bool ok[SUB];
#pragma unroll
          for (int kk = 0; kk < SUB; kk++) {
            const int k = k0 + kk;
            const int32_t cc = group_bcast<LPR>(cj[k / LPR], k % LPR, lane);
            vv[kk] = group_bcast_t<LPR>(v[k / LPR], k % LPR, lane);
            ok[kk] = colok && (j0 + k) < w;
            if (ok[kk]) gather(cc, xg[kk]);
          }
#pragma unroll
          for (int kk = 0; kk < SUB; kk++)
            if (ok[kk]) {
#pragma unroll
              for (int e = 0; e < VPL; e++) acc[e] = fma(vv[kk], xg[kk][e], acc[e]);
            }
        }
      }
    }
    T yn[VPL];
#pragma unroll
    for (int e = 0; e < VPL; e++) yn[e] = T(0);
    if (live && colok) {
      if constexpr (XSC) {  // (A + s)(x .* f) = ((A + s) x) .* f, one product per result
#pragma unroll
        for (int e = 0; e < VPL; e++) {
          acc[e] *= xsc[e];
          xo[e] *= xsc[e];
        }
      }
      T* yp = a.y + (int64_t)row * a.ldy + c0;
      if (flags & F_BETA) {
        T yo[VPL];
        vload<T, VPL>(yp, yo);
#pragma unroll
        for (int e = 0; e < VPL; e++) yn[e] = a.alpha * acc[e] + a.beta * yo[e];
      } else {
#pragma unroll
        for (int e = 0; e < VPL; e++) yn[e] = a.alpha * acc[e];
      }
      if (!(flags & F_NOSTORE)) {
        if (flags & F_NTY)
          vstore_nt<T, VPL>(yp, yn);
        else
          vstore<T, VPL>(yp, yn);
      }
      if (flags & F_DWW) {
#pragma unroll
        for (int e = 0; e < VPL; e++) dww[e] = fma(yn[e], yn[e], dww[e]);
      }
      if (flags & F_DVW) {
#pragma unroll
        for (int e = 0; e < VPL; e++) dvw[e] = fma(xo[e], yn[e], dvw[e]);
      }
      if (flags & F_AXPY) {
        T* up = a.u + (int64_t)row * a.ldu + c0;
        T uo[VPL], un[VPL];
        vload<T, VPL>(up, uo);
#pragma unroll
        for (int e = 0; e < VPL; e++) un[e] = a.gamma * yn[e] + a.delta * uo[e];
        vstore<T, VPL>(up, un);
      }
    }
    if constexpr (GR > 0) {
      // rows of the wave (dead rows and inactive columns are zero) through LDS into the MFMA layout
      T* wt = gex[wv][0];
      T* vt = gex[wv][1];
#pragma unroll
      for (int e = 0; e < VPL; e++) {
        wt[rw * GW + c0 + e] = yn[e];
        vt[rw * GW + c0 + e] = (live && colok) ? xo[e] : T(0);
      }
      wave_lds_sync();
      constexpr int KS = (RPW + 3) / 4;
#pragma unroll
      for (int ks = 0; ks < KS; ks++) {
        const int r = 4 * ks + (lane >> 4);
        T aw[TG], av[TG];
#pragma unroll
        for (int tm = 0; tm < TG; tm++) {
          const int col = tm * 16 + (lane & 15);
          const bool ok = r < RPW && col < GW;
          aw[tm] = ok ? wt[r * GW + col] : T(0);
          av[tm] = ok ? vt[r * GW + col] : T(0);
        }
#pragma unroll
        for (int tm = 0; tm < TG; tm++)
#pragma unroll
          for (int tk = 0; tk < TG; tk++) {
            if ((flags & F_GWW) && tm <= tk) gww[tm][tk] = GM::op(aw[tm], aw[tk], gww[tm][tk]);
            if (flags & F_GVW) gvw[tm][tk] = GM::op(av[tm], aw[tk], gvw[tm][tk]);
          }
      }
      wave_lds_sync();  // the next tile overwrites the exchange tiles
    }
  }

  if constexpr (GR > 0) {
    // S = [W'W, V'W] of this workgroup: the four waves' tiles added in LDS in
    // wave order, then stored as partials[blockIdx.x][n*n] and
    // partials[gridDim.x + blockIdx.x][n*n] (column-major n x n)
    __shared__ T gsum[2][(16 * TG) * (16 * TG)];
    const int n = a.ncols;
#pragma unroll 1
    for (int w = 0; w < kBlock / 64; w++) {
      if (wv == w) {
#pragma unroll
        for (int tm = 0; tm < TG; tm++)
#pragma unroll
          for (int tk = 0; tk < TG; tk++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
              const int i = tm * 16 + GM::row(lane, r), j = tk * 16 + (lane & 15);
              if (i >= n || j >= n) continue;
              if ((flags & F_GWW) && tm <= tk) {
                const T v = gww[tm][tk][r];
                gsum[0][j * n + i] = w == 0 ? v : gsum[0][j * n + i] + v;
                if (tm < tk) gsum[0][i * n + j] = w == 0 ? v : gsum[0][i * n + j] + v;
              }
              if (flags & F_GVW) {
                const T v = gvw[tm][tk][r];
                gsum[1][j * n + i] = w == 0 ? v : gsum[1][j * n + i] + v;
              }
            }
      }
      __syncthreads();
    }
    const int nn = n * n;
    for (int o = threadIdx.x; o < nn; o += kBlock) {
      a.partials[(int64_t)blockIdx.x * nn + o] = gsum[0][o];
      a.partials[((int64_t)gridDim.x + blockIdx.x) * nn + o] = gsum[1][o];
    }
  }

  if (flags & (F_DWW | F_DVW)) {
    __shared__ T red[2][kBlock / 64][64 * VPL];
#pragma unroll
    for (int e = 0; e < VPL; e++) {
      dww[e] = wave_sum_strided<LPR>(dww[e]);
      dvw[e] = wave_sum_strided<LPR>(dvw[e]);
    }
    if (lane < LPR) {
#pragma unroll
      for (int e = 0; e < VPL; e++) {
        red[0][wv][lane * VPL + e] = dww[e];
        red[1][wv][lane * VPL + e] = dvw[e];
      }
    }
    __syncthreads();
    const int width = a.ncols;
    for (int i = threadIdx.x; i < 2 * width; i += kBlock) {
      const int which = i / width, c = i % width;
      T s = T(0);
#pragma unroll
      for (int k = 0; k < kBlock / 64; k++) s += red[which][k][c];
      a.partials[(int64_t)blockIdx.x * 2 * width + i] = s;
    }
  }
}

namespace {

template <typename T, int VPL, int LPR, bool XSC, int GR = 0, int SL = 8, bool STG = false>
int launch_one(SpmvArgs<T>& a) {
  auto kern = k_sell_spmv<T, VPL, LPR, XSC, GR, SL, STG>;
  constexpr int RPB = (kBlock / 64) * (64 / LPR);
  a.ntiles = (a.nslots + RPB - 1) / RPB;
  const size_t lds = STG ? a.st_lds_bytes : 0;
  const int grid = persistent_blocks(kern, lds, a.ntiles, a.blocks_per_cu);
  if (GR > 0)
    a.partials = (T*)device_work(sizeof(T) * (size_t)2 * grid * a.ncols * a.ncols);
  else if (a.flags & (F_DWW | F_DVW))
    a.partials = (T*)device_work(sizeof(T) * ((size_t)grid * 2 * a.ncols + 2 * a.ncols));
  ktimer_begin();
  hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), lds, rt().stream, a);
  PHX_HIP(hipGetLastError());
  ktimer_end();
  return grid;
}

template <typename T, int VPL, bool XSC>
int launch_lpr(SpmvArgs<T>& a, int lpr) {
  switch (lpr) {
    case 1: return launch_one<T, VPL, 1, XSC>(a);
    case 2: return launch_one<T, VPL, 2, XSC>(a);
    case 4: return launch_one<T, VPL, 4, XSC>(a);
    case 8: return launch_one<T, VPL, 8, XSC>(a);
    case 16: return launch_one<T, VPL, 16, XSC>(a);
    case 32: return launch_one<T, VPL, 32, XSC>(a);
    default: return launch_one<T, VPL, 64, XSC>(a);
  }
}
// depth 16 or 32 (rows of many entries; no per-column x factors): row groups of at most 8 lanes
template <typename T, int VPL, int SL>
int launch_deep(SpmvArgs<T>& a, int lpr) {
  switch (lpr) {
    case 1: return launch_one<T, VPL, 1, false, 0, SL>(a);
    case 2: return launch_one<T, VPL, 2, false, 0, SL>(a);
    case 4: return launch_one<T, VPL, 4, false, 0, SL>(a);
    default: return launch_one<T, VPL, 8, false, 0, SL>(a);
  }
}

// the deeper batches hold their gathered values in registers: depth 16 and
// 32 take at most two 16-byte vectors per lane (wider lanes would need more
// registers than the file holds without spilling)
template <typename T, int VPL>
int launch_xsc(SpmvArgs<T>& a, int lpr, int depth) {
  if (!a.xscale && lpr <= 8) {
    if constexpr (VPL * sizeof(T) <= 32)
      if (depth >= 32) return launch_deep<T, VPL, 32>(a, lpr);
    if constexpr (VPL * sizeof(T) <= 32)
      if (depth >= 16) return launch_deep<T, VPL, 16>(a, lpr);
  }
  return a.xscale ? launch_lpr<T, VPL, true>(a, lpr) : launch_lpr<T, VPL, false>(a, lpr);
}

// staged windows (row groups of at most 8 lanes of 16-byte vectors): depth 8, or 16 for lanes of at most 32 bytes
template <typename T, int VPL, int SL>
int launch_stage_sl(SpmvArgs<T>& a, int lpr) {
  switch (lpr) {
    case 1: return launch_one<T, VPL, 1, false, 0, SL, true>(a);
    case 2: return launch_one<T, VPL, 2, false, 0, SL, true>(a);
    case 4: return launch_one<T, VPL, 4, false, 0, SL, true>(a);
    default: return launch_one<T, VPL, 8, false, 0, SL, true>(a);
  }
}
template <typename T, int VPL>
int launch_stage(SpmvArgs<T>& a, int lpr, int depth) {
  if constexpr (VPL * sizeof(T) <= 32)
    if (depth >= 16) return launch_stage_sl<T, VPL, 16>(a, lpr);
  return launch_stage_sl<T, VPL, 8>(a, lpr);
}

// the Gram-epilogue variants: at most 32 columns (lane groups of at most 32 lanes)
template <typename T, int VPL, int GR>
int launch_gram_lpr(SpmvArgs<T>& a, int lpr) {
  switch (lpr) {
    case 1: return launch_one<T, VPL, 1, false, GR>(a);
    case 2: return launch_one<T, VPL, 2, false, GR>(a);
    case 4: return launch_one<T, VPL, 4, false, GR>(a);
    case 8: return launch_one<T, VPL, 8, false, GR>(a);
    case 16: return launch_one<T, VPL, 16, false, GR>(a);
    default: return launch_one<T, VPL, 32, false, GR>(a);
  }
}
template <typename T, int VPL>
int launch_gram(SpmvArgs<T>& a, int lpr) {
  return a.ncols <= 16 ? launch_gram_lpr<T, VPL, 1>(a, lpr) : launch_gram_lpr<T, VPL, 2>(a, lpr);
}

// C(0:n, 0:n) (column stride ldc) = sum over the nblocks partial n x n tiles, in a fixed order
template <typename T>
__global__ void __launch_bounds__(kBlock) k_spmv_gram_out(const T* __restrict__ partials, int nblocks, int n,
                                                          T* __restrict__ C, int ldc) {
  const T s = block_sum16(partials, nblocks, n * n);
  const int o = blockIdx.x * 16 + threadIdx.x;
  if (threadIdx.x < 16 && o < n * n) C[(int64_t)(o / n) * ldc + o % n] = s;
}

inline bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

}  // namespace

template <typename T>
const SellStage* spmv_stage_for(const SparseMat<T>* A, int nc, int vpl, int* pshift, size_t* lds_bytes) {
  if (vpl * (int)sizeof(T) < 16 || (nc * (int)sizeof(T)) % 16 != 0) return nullptr;
  const int groups = (nc + vpl - 1) / vpl;
  int lpr = 1;
  while (lpr < groups) lpr <<= 1;
  if (lpr > 8) return nullptr;
  const SellStage* S = sell_stage(A, (kBlock / 64) * (64 / lpr));
  if (!S) return nullptr;
  const int ppr = nc * (int)sizeof(T) / 16;
  int ps = 0;
  while ((1 << ps) < ppr) ps++;
  const size_t bytes = (size_t)S->max_rows * ((size_t)16 << ps);
  // the kernel's own LDS: the dot partials of its four waves (2 x 4 x 64 lane groups' columns), and a margin
  const int64_t static_lds = (int64_t)2 * (kBlock / 64) * 64 * vpl * (int64_t)sizeof(T) + 256;
  if (bytes > (size_t)kStageLdsBytes || (int64_t)bytes + static_lds > rt().lds_per_block) return nullptr;
  if (pshift) *pshift = ps;
  if (lds_bytes) *lds_bytes = std::max<size_t>(16, bytes);
  return S;
}
template const SellStage* spmv_stage_for<double>(const SparseMat<double>*, int, int, int*, size_t*);
template const SellStage* spmv_stage_for<float>(const SparseMat<float>*, int, int, int*, size_t*);

template <typename T>
void spmv_launch(const SparseMat<T>* A, T alpha, const T* shifts_host, const Mvec<T>* x, T beta, Mvec<T>* y,
                 const SpmvRequest& req, T* dot_ww_host, T* dot_vw_host, T gamma, T delta, Mvec<T>* U) {
  const int nvec = x->nvec;
  const bool gram = req.gram_ww_dev || req.gram_vw_dev;
  PHX_REQUIRE(y || (req.no_store && gram && !req.has_beta && !req.axpy_u && !req.dot_ww && !req.dot_vw),
              PHIST_INVALID_INPUT, "no output block");
  PHX_REQUIRE(!gram || (nvec <= 32 && A->nrows == A->ncols && !req.dot_ww && !req.dot_vw && !req.axpy_u &&
                        !req.xscale && req.ncols_active < 0),
              PHIST_INVALID_INPUT, "Gram epilogue: at most 32 columns of a square matrix, no other epilogue");
  if (y) {
    PHX_REQUIRE(y->nvec == nvec, PHIST_FUNCTIONAL_ERROR, "x has %d columns, y has %d", nvec, y->nvec);
    PHX_REQUIRE(y->nrows == A->nrows, PHIST_INVALID_INPUT, "shape mismatch: A has %lld rows, y has %lld",
                (long long)A->nrows, (long long)y->nrows);
  }
  PHX_REQUIRE(x->nrows == A->ncols, PHIST_INVALID_INPUT, "shape mismatch: A has %lld columns, x has %lld rows",
              (long long)A->ncols, (long long)x->nrows);
  PHX_REQUIRE(!req.axpy_u || (U && U->nvec == nvec && U->nrows == y->nrows), PHIST_INVALID_INPUT, "bad U block");
  PHX_REQUIRE(!(req.has_shift || req.dot_vw) || A->nrows == A->ncols, PHIST_INVALID_INPUT,
              "shifts and <x,y> dots need a square matrix (A is %lld x %lld)", (long long)A->nrows,
              (long long)A->ncols);
  if (y && x->root == y->root)
    PHX_REQUIRE(x->jofs + nvec <= y->jofs || y->jofs + nvec <= x->jofs, PHIST_INVALID_INPUT,
                "x and y overlap in memory");

  constexpr int V16 = 16 / sizeof(T);
  bool vec_ok = (nvec % V16 == 0) && (x->lda % V16 == 0) && aligned16(x->dev());
  if (y) vec_ok = vec_ok && (y->lda % V16 == 0) && aligned16(y->dev());
  if (req.axpy_u) vec_ok = vec_ok && (U->lda % V16 == 0) && aligned16(U->dev());
  int VPL = vec_ok ? V16 : 1;
  const auto tuned = A->tuned.find(nvec);
  if (!gram) {  // the gathered bytes per lane and nonzero (same sums whichever is taken)
    int lb = rt().spmv_lane_bytes;
    if (req.lane_bytes >= 0)
      lb = req.lane_bytes;
    else if (lb == 0 && tuned != A->tuned.end())
      lb = tuned->second.lane_bytes;
    const int w = lb / (int)sizeof(T);
    if (lb > 0 && lb <= (int)sizeof(T)) {
      VPL = 1;
    } else if (vec_ok && (w == 2 * V16 || w == 4 * V16)) {
      bool ok = nvec % w == 0 && x->lda % w == 0;
      if (y) ok = ok && y->lda % w == 0;
      if (req.axpy_u) ok = ok && U->lda % w == 0;
      if (ok) VPL = w;
    }
  }
  // block columns per launch: 64 lane groups, or fewer as set or tuned (each
  // further launch reads the matrix again; each row's sum is the same)
  int maxcols = 64 * VPL;
  int colblock = req.col_block;
  if (colblock < 0) colblock = rt().spmv_col_block;
  if (colblock < 0) colblock = tuned != A->tuned.end() ? tuned->second.col_block : 0;
  if (!gram && colblock > 0 && colblock % VPL == 0) maxcols = std::min(maxcols, colblock);

  SpmvArgs<T> a;
  a.chunk_ptr = A->d_chunk_ptr;
  a.chunk_w = A->d_chunk_w;
  a.perm = A->d_perm;
  a.col = A->d_col;
  a.val = A->d_val;
  int clog2 = 0;
  while ((1 << clog2) < A->C) clog2++;
  a.clog2 = clog2;
  a.nslots = A->nchunks * A->C;
  a.ldx = x->lda;
  a.ldy = y ? y->lda : 0;
  a.ldu = U ? U->lda : 0;
  a.alpha = alpha;
  a.beta = beta;
  a.gamma = gamma;
  a.delta = delta;
  a.flags = (req.has_shift ? F_SHIFT : 0) | (req.has_beta ? F_BETA : 0) | (req.dot_ww ? F_DWW : 0) |
            (req.dot_vw ? F_DVW : 0) | (req.axpy_u ? F_AXPY : 0) | (req.gram_ww_dev ? F_GWW : 0) |
            (req.gram_vw_dev ? F_GVW : 0) | ((req.no_store || !y) ? F_NOSTORE : 0);
  if (y && (rt().spmv_nt_y == 1 || (rt().spmv_nt_y < 0 && (int64_t)sizeof(T) * y->nrows * nvec >= Runtime::kNtYBytes &&
                                     (int)sizeof(T) * nvec < 128)))
    a.flags |= F_NTY;
  if (req.blocks_per_cu >= 0)
    a.blocks_per_cu = req.blocks_per_cu;
  else if (rt().spmv_blocks_per_cu >= 0)
    a.blocks_per_cu = rt().spmv_blocks_per_cu;
  else if (tuned != A->tuned.end())
    a.blocks_per_cu = tuned->second.blocks_per_cu;
  else
    a.blocks_per_cu = spmv_table_blocks_per_cu(A->nrows ? (double)A->nnz / (double)A->nrows : 0.0, nvec);
  // entries per row whose loads go out together (8 or 32)
  int depth;
  if (req.depth >= 0)
    depth = req.depth;
  else if (rt().spmv_depth >= 0)
    depth = rt().spmv_depth;
  else if (tuned != A->tuned.end())
    depth = tuned->second.depth;
  else
    depth = 8;
  const bool dots = req.dot_ww || req.dot_vw;
  // x rows of each tile staged in LDS (where the matrix's tiles allow it)
  int stage = req.stage;
  if (stage < 0) stage = rt().spmv_stage;
  if (stage < 0) stage = tuned != A->tuned.end() ? tuned->second.stage : 0;
  a.st_ivptr = nullptr;
  a.st_iv = nullptr;
  a.st_lcol = nullptr;
  a.st_ppr = a.st_pshift = 0;
  a.st_lds_bytes = 0;

  for (int cb = 0; cb < nvec; cb += maxcols) {
    const int nc = std::min(maxcols, nvec - cb);
    a.ncols = nc;
    a.nact = req.ncols_active >= 0 ? std::max(0, std::min(nc, req.ncols_active - cb)) : nc;
    a.x = x->dev() + cb;
    a.y = y ? y->dev() + cb : nullptr;
    a.u = U ? U->dev() + cb : nullptr;
    a.shift_dev = (req.has_shift && req.shifts_dev) ? static_cast<const T*>(req.shifts_dev) + cb : nullptr;
    for (int c = 0; c < nc; c++) a.shift[c] = (req.has_shift && !a.shift_dev) ? shifts_host[cb + c] : T(0);
    const int groups = (nc + VPL - 1) / VPL;
    int lpr = 1;
    while (lpr < groups) lpr <<= 1;
    a.xscale = req.xscale ? req.xscale + (int64_t)cb * req.xscale_stride : nullptr;
    a.xscale_stride = req.xscale_stride;
    if (gram) {  // one column block (nvec <= 32)
      const int grid = VPL == 1 ? launch_gram<T, 1>(a, lpr) : launch_gram<T, V16>(a, lpr);
      const int nn = nc * nc;
      if (req.gram_ww_dev) {
        hipLaunchKernelGGL(k_spmv_gram_out<T>, dim3((nn + 15) / 16), dim3(kBlock), 0, rt().stream, a.partials, grid,
                           nc, static_cast<T*>(req.gram_ww_dev), req.gram_ww_ld);
        PHX_HIP(hipGetLastError());
      }
      if (req.gram_vw_dev) {
        hipLaunchKernelGGL(k_spmv_gram_out<T>, dim3((nn + 15) / 16), dim3(kBlock), 0, rt().stream,
                           a.partials + (size_t)grid * nn, grid, nc, static_cast<T*>(req.gram_vw_dev),
                           req.gram_vw_ld);
        PHX_HIP(hipGetLastError());
      }
      continue;
    }
    const SellStage* S = stage > 0 && !a.xscale ? spmv_stage_for(A, nc, VPL, &a.st_pshift, &a.st_lds_bytes) : nullptr;
    int grid;
    if (S) {
      rt().spmv_staged_launches++;
      a.st_ivptr = S->d_ivptr;
      a.st_iv = S->d_iv;
      a.st_lcol = S->d_lcol;
      a.st_ppr = nc * (int)sizeof(T) / 16;
      grid = VPL == V16 ? launch_stage<T, V16>(a, lpr, depth)
                        : (VPL == 2 * V16 ? launch_stage<T, 2 * V16>(a, lpr, depth) : launch_stage<T, 4 * V16>(a, lpr, depth));
    } else {
      grid = VPL == 1     ? launch_xsc<T, 1>(a, lpr, depth)
             : VPL == V16 ? launch_xsc<T, V16>(a, lpr, depth)
             : VPL == 2 * V16 ? launch_xsc<T, 2 * V16>(a, lpr, depth)
                              : launch_xsc<T, 4 * V16>(a, lpr, depth);
    }
    if (dots && req.dots_dev && cb == 0 && nc == nvec) {
      hipLaunchKernelGGL(k_reduce_partials<T>, dim3((2 * nc + 15) / 16), dim3(kBlock), 0, rt().stream, a.partials,
                         grid, 2 * nc, static_cast<T*>(req.dots_dev));
      PHX_HIP(hipGetLastError());
    } else if (dots && (req.dots_dev || req.dot_ww_dev || req.dot_vw_dev)) {
      // fixed-order sums into the tail of the partials buffer (sized for them), then to the two outputs
      // (dots_dev: ww at [0, nvec), vw at [nvec, 2 nvec), this block's columns at offset cb)
      T* sums = a.partials + (size_t)grid * 2 * nc;
      hipLaunchKernelGGL(k_reduce_partials<T>, dim3((2 * nc + 15) / 16), dim3(kBlock), 0, rt().stream, a.partials,
                         grid, 2 * nc, sums);
      PHX_HIP(hipGetLastError());
      T* ww_out = req.dots_dev ? static_cast<T*>(req.dots_dev) : static_cast<T*>(req.dot_ww_dev);
      T* vw_out = req.dots_dev ? static_cast<T*>(req.dots_dev) + nvec : static_cast<T*>(req.dot_vw_dev);
      if (ww_out)
        PHX_HIP(hipMemcpyAsync(ww_out + cb, sums, sizeof(T) * nc, hipMemcpyDeviceToDevice, rt().stream));
      if (vw_out)
        PHX_HIP(hipMemcpyAsync(vw_out + cb, sums + nc, sizeof(T) * nc, hipMemcpyDeviceToDevice, rt().stream));
    } else if (dots) {
      std::vector<T> sums(2 * nc);
      reduce_partials_to_host(a.partials, grid, 2 * nc, sums.data());
      for (int c = 0; c < nc; c++) {
        if (dot_ww_host) dot_ww_host[cb + c] = sums[c];
        if (dot_vw_host) dot_vw_host[cb + c] = sums[nc + c];
      }
    }
  }
}

template void spmv_launch<double>(const SparseMat<double>*, double, const double*, const Mvec<double>*, double,
                                  Mvec<double>*, const SpmvRequest&, double*, double*, double, double, Mvec<double>*);
template void spmv_launch<float>(const SparseMat<float>*, float, const float*, const Mvec<float>*, float, Mvec<float>*,
                                 const SpmvRequest&, float*, float*, float, float, Mvec<float>*);

}  // namespace phx
