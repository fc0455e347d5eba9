// blas1.hip -- BLAS-1 on row-major block vectors.
//
//   k_axpby      Y(:,c) = a_c * X(:,c) + b_c * Y(:,c)   covers mvec_add_mvec,
//                mvec_vadd_mvec, mvec_scale, mvec_vscale, mvec_put_value and
//                block copies (mvec_module.F90:1288-1527, axpy_kernels.F90);
//                X is not read when every a_c is 0, Y not read when every
//                b_c is 0 (builtin's dset/dscal/dcopy special cases).
//   k_elemwise   W = alpha * V .* W  (mvec_times_mvec_elemwise)
//   k_coldot     column dots <V(:,c), W(:,c)> (dot_kernels.F90:143-156),
//                norms through V == W (nrm2_kernels.F90)
//
// Same mapping as the SpMV: a row is owned by LPR lanes holding VPL
// consecutive columns each, workgroups walk row tiles on a persistent
// XCD-aware grid, reductions are fixed-order two-stage.  The streaming
// kernels take kRowUnroll rows per lane group and tile, whose loads are all
// in flight before the first store.
#include "amd_device.hpp"
#include <algorithm>

namespace phx {

template <typename T>
struct RowArgs {
  int64_t nrows;
  int64_t ntiles;
  int ncols;
  const T* __restrict__ x;
  int64_t ldx;
  T* __restrict__ y;
  int64_t ldy;
  T alpha;
  int flags;
  T* __restrict__ partials;
  T a[64 * 4];
  T b[64 * 4];
};

enum : int { B_READX = 1, B_READY = 2 };

// rows per lane group and tile of the streaming kernels (k_axpby, k_elemwise)
constexpr int kRowUnroll = 4;

template <typename T, int VPL, int LPR>
__global__ void __launch_bounds__(kBlock) k_axpby(const RowArgs<T> p) {
  constexpr int RPW = 64 / LPR, RPB = (kBlock / 64) * RPW;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int q = lane % LPR, rw = lane / LPR, c0 = q * VPL;
  if (c0 >= p.ncols) return;
  T ca[VPL], cb[VPL];
#pragma unroll
  for (int e = 0; e < VPL; e++) {
    ca[e] = p.a[c0 + e];
    cb[e] = p.b[c0 + e];
  }
  // a tile is kRowUnroll sub-tiles of RPB rows: all their loads go out before the first store
  const TileWalk tw = xcd_tiles(p.ntiles);
  for (int64_t t = tw.t; t < tw.end; t += tw.stride) {
    int64_t iu[kRowUnroll];
    T xv[kRowUnroll][VPL], yv[kRowUnroll][VPL];
#pragma unroll
    for (int u = 0; u < kRowUnroll; u++) {
      iu[u] = (t * kRowUnroll + u) * RPB + wv * RPW + rw;
      if (iu[u] < p.nrows) {
        if (p.flags & B_READX) vload<T, VPL>(p.x + iu[u] * p.ldx + c0, xv[u]);
        if (p.flags & B_READY) vload<T, VPL>(p.y + iu[u] * p.ldy + c0, yv[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < kRowUnroll; u++) {
      if (iu[u] >= p.nrows) continue;
      T out[VPL];
#pragma unroll
      for (int e = 0; e < VPL; e++) out[e] = T(0);
      if (p.flags & B_READX) {
#pragma unroll
        for (int e = 0; e < VPL; e++) out[e] = ca[e] * xv[u][e];
      }
      if (p.flags & B_READY) {
#pragma unroll
        for (int e = 0; e < VPL; e++) out[e] = (p.flags & B_READX) ? out[e] + cb[e] * yv[u][e] : cb[e] * yv[u][e];
      }
      vstore<T, VPL>(p.y + iu[u] * p.ldy + c0, out);
    }
  }
}

template <typename T, int VPL, int LPR>
__global__ void __launch_bounds__(kBlock) k_elemwise(const RowArgs<T> p) {
  constexpr int RPW = 64 / LPR, RPB = (kBlock / 64) * RPW;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int q = lane % LPR, rw = lane / LPR, c0 = q * VPL;
  if (c0 >= p.ncols) return;
  const TileWalk tw = xcd_tiles(p.ntiles);
  for (int64_t t = tw.t; t < tw.end; t += tw.stride) {
    int64_t iu[kRowUnroll];
    T xv[kRowUnroll][VPL], yv[kRowUnroll][VPL];
#pragma unroll
    for (int u = 0; u < kRowUnroll; u++) {
      iu[u] = (t * kRowUnroll + u) * RPB + wv * RPW + rw;
      if (iu[u] < p.nrows) {
        vload<T, VPL>(p.x + iu[u] * p.ldx + c0, xv[u]);
        vload<T, VPL>(p.y + iu[u] * p.ldy + c0, yv[u]);
      }
    }
#pragma unroll
    for (int u = 0; u < kRowUnroll; u++) {
      if (iu[u] >= p.nrows) continue;
#pragma unroll
      for (int e = 0; e < VPL; e++) yv[u][e] = p.alpha * xv[u][e] * yv[u][e];
      vstore<T, VPL>(p.y + iu[u] * p.ldy + c0, yv[u]);
    }
  }
}

template <typename T, int VPL, int LPR>
__global__ void __launch_bounds__(kBlock) k_coldot(const RowArgs<T> p) {
  constexpr int RPW = 64 / LPR, RPB = (kBlock / 64) * RPW;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int q = lane % LPR, rw = lane / LPR, c0 = q * VPL;
  const bool colok = c0 < p.ncols;
  T acc[VPL];
#pragma unroll
  for (int e = 0; e < VPL; e++) acc[e] = T(0);
  const TileWalk tw = xcd_tiles(p.ntiles);
  for (int64_t t = tw.t; t < tw.end; t += tw.stride) {
    const int64_t i = t * RPB + wv * RPW + rw;
    if (i >= p.nrows || !colok) continue;
    T xv[VPL], yv[VPL];
    vload<T, VPL>(p.x + i * p.ldx + c0, xv);
    vload<T, VPL>(p.y + i * p.ldy + c0, yv);
#pragma unroll
    for (int e = 0; e < VPL; e++) acc[e] = fma(xv[e], yv[e], acc[e]);
  }
  __shared__ T red[kBlock / 64][64 * VPL];
#pragma unroll
  for (int e = 0; e < VPL; e++) acc[e] = wave_sum_strided<LPR>(acc[e]);
  if (lane < LPR) {
#pragma unroll
    for (int e = 0; e < VPL; e++) red[wv][lane * VPL + e] = acc[e];
  }
  __syncthreads();
  for (int c = threadIdx.x; c < p.ncols; c += kBlock) {
    T s = T(0);
#pragma unroll
    for (int k = 0; k < kBlock / 64; k++) s += red[k][c];
    p.partials[(int64_t)blockIdx.x * p.ncols + c] = s;
  }
}

namespace {

inline bool al16(const void* ptr) { return ((uintptr_t)ptr & 15) == 0; }

enum Kind { K_AXPBY, K_ELEM, K_DOT };

template <typename T, int VPL, int LPR>
int launch_kind(Kind kind, RowArgs<T>& p) {
  constexpr int RPB = (kBlock / 64) * (64 / LPR);
  const int64_t rows_per_tile = kind == K_DOT ? RPB : (int64_t)RPB * kRowUnroll;
  p.ntiles = (p.nrows + rows_per_tile - 1) / rows_per_tile;
  int grid;
  ktimer_begin();
  switch (kind) {
    case K_AXPBY:
      grid = persistent_blocks(k_axpby<T, VPL, LPR>, 0, p.ntiles);
      hipLaunchKernelGGL((k_axpby<T, VPL, LPR>), dim3(grid), dim3(kBlock), 0, rt().stream, p);
      break;
    case K_ELEM:
      grid = persistent_blocks(k_elemwise<T, VPL, LPR>, 0, p.ntiles);
      hipLaunchKernelGGL((k_elemwise<T, VPL, LPR>), dim3(grid), dim3(kBlock), 0, rt().stream, p);
      break;
    default:
      grid = persistent_blocks(k_coldot<T, VPL, LPR>, 0, p.ntiles);
      p.partials = (T*)device_work(sizeof(T) * ((size_t)grid * p.ncols + p.ncols));
      hipLaunchKernelGGL((k_coldot<T, VPL, LPR>), dim3(grid), dim3(kBlock), 0, rt().stream, p);
      break;
  }
  PHX_HIP(hipGetLastError());
  ktimer_end();
  return grid;
}

template <typename T, int VPL>
int launch_vpl(Kind kind, RowArgs<T>& p, int lpr) {
  switch (lpr) {
    case 1: return launch_kind<T, VPL, 1>(kind, p);
    case 2: return launch_kind<T, VPL, 2>(kind, p);
    case 4: return launch_kind<T, VPL, 4>(kind, p);
    case 8: return launch_kind<T, VPL, 8>(kind, p);
    case 16: return launch_kind<T, VPL, 16>(kind, p);
    case 32: return launch_kind<T, VPL, 32>(kind, p);
    default: return launch_kind<T, VPL, 64>(kind, p);
  }
}

// runs kind over the columns of Y (and X) in blocks of at most 64*VPL;
// per-column coefficients come from fa/fb, dots are written to out_host (or,
// when out_dev is set, left on the device in out_dev)
template <typename T, class FA, class FB>
void run_rows(Kind kind, const T* xdev, int64_t ldx, T* ydev, int64_t ldy, int64_t nrows, int ncols, T alpha, int flags,
              FA fa, FB fb, T* out_host, T* out_dev = nullptr) {
  PHX_REQUIRE(!(kind == K_DOT && out_dev && nrows == 0), PHIST_INVALID_INPUT, "device dots of an empty block");
  if (nrows == 0 || ncols == 0) {
    if (kind == K_DOT)
      for (int c = 0; c < ncols; c++) out_host[c] = T(0);
    return;
  }
  constexpr int V16 = 16 / sizeof(T);
  bool vec = (ncols % V16 == 0) && (ldy % V16 == 0) && al16(ydev);
  if (xdev) vec = vec && (ldx % V16 == 0) && al16(xdev);
  const int VPL = vec ? V16 : 1;
  const int maxc = 64 * VPL;
  RowArgs<T> p;
  p.nrows = nrows;
  p.ldx = ldx;
  p.ldy = ldy;
  p.alpha = alpha;
  p.flags = flags;
  for (int cb = 0; cb < ncols; cb += maxc) {
    const int nc = std::min(maxc, ncols - cb);
    p.ncols = nc;
    p.x = xdev ? xdev + cb : nullptr;
    p.y = ydev + cb;
    for (int c = 0; c < nc; c++) {
      p.a[c] = fa(cb + c);
      p.b[c] = fb(cb + c);
    }
    const int groups = (nc + VPL - 1) / VPL;
    int lpr = 1;
    while (lpr < groups) lpr <<= 1;
    const int grid = VPL == 1 ? launch_vpl<T, 1>(kind, p, lpr) : launch_vpl<T, V16>(kind, p, lpr);
    if (kind == K_DOT && out_dev) {
      hipLaunchKernelGGL(k_reduce_partials<T>, dim3((nc + 15) / 16), dim3(kBlock), 0, rt().stream, p.partials, grid, nc,
                         out_dev + cb);
      PHX_HIP(hipGetLastError());
    } else if (kind == K_DOT) {
      reduce_partials_to_host(p.partials, grid, nc, out_host + cb);
    }
  }
}

template <typename T>
T zero_fn(int) {
  return T(0);
}

}  // namespace

template <typename T>
void blas_put_value(Mvec<T>* V, T value) {
  if (value == T(0)) {  // neither operand read: writes zeros
    run_rows<T>(K_AXPBY, nullptr, 0, V->dev(), V->lda, V->nrows, V->nvec, T(0), 0, zero_fn<T>, zero_fn<T>, nullptr);
    return;
  }
  // X is a single row of `value`s broadcast to every row (ldx = 0)
  std::vector<T> rowv(V->nvec, value);
  T* d_row = (T*)device_work(sizeof(T) * V->nvec);
  const size_t bytes = sizeof(T) * V->nvec;
  const bool staged = stage_upload_2d(d_row, bytes, rowv.data(), bytes, bytes, 1);
  if (!staged) PHX_HIP(hipMemcpyAsync(d_row, rowv.data(), bytes, hipMemcpyHostToDevice, rt().stream));
  run_rows<T>(
      K_AXPBY, d_row, 0, V->dev(), V->lda, V->nrows, V->nvec, T(0), B_READX, [](int) { return T(1); }, zero_fn<T>,
      nullptr);
  if (!staged) sync_stream();  // rowv (the copy source) is released on return
}

template <typename T>
void blas_scale_cols(Mvec<T>* V, const T* s, bool uniform) {
  run_rows<T>(
      K_AXPBY, nullptr, 0, V->dev(), V->lda, V->nrows, V->nvec, T(0), B_READY, zero_fn<T>,
      [&](int c) { return uniform ? s[0] : s[c]; }, nullptr);
}

template <typename T>
void blas_axpby(const T* alpha, bool alpha_uniform, const Mvec<T>* X, T beta, Mvec<T>* Y) {
  PHX_REQUIRE(X->nvec == Y->nvec && X->nrows == Y->nrows, PHIST_INVALID_INPUT, "block shapes differ (%d vs %d columns)",
              X->nvec, Y->nvec);
  bool any_a = false;
  for (int c = 0; c < X->nvec; c++) any_a |= (alpha_uniform ? alpha[0] : alpha[c]) != T(0);
  int flags = (any_a ? B_READX : 0) | (beta != T(0) ? B_READY : 0);
  run_rows<T>(
      K_AXPBY, X->dev(), X->lda, Y->dev(), Y->lda, Y->nrows, Y->nvec, T(0), flags,
      [&](int c) { return alpha_uniform ? alpha[0] : alpha[c]; }, [&](int) { return beta; }, nullptr);
}

template <typename T>
void blas_copy(const Mvec<T>* X, Mvec<T>* Y) {
  PHX_REQUIRE(X->nvec == Y->nvec && X->nrows == Y->nrows, PHIST_INVALID_INPUT, "block shapes differ");
  run_rows<T>(
      K_AXPBY, X->dev(), X->lda, Y->dev(), Y->lda, Y->nrows, Y->nvec, T(0), B_READX, [](int) { return T(1); },
      zero_fn<T>, nullptr);
}

template <typename T>
void blas_elemwise(T alpha, const Mvec<T>* V, Mvec<T>* W) {
  PHX_REQUIRE(V->nvec == W->nvec && V->nrows == W->nrows, PHIST_INVALID_INPUT, "block shapes differ");
  run_rows<T>(K_ELEM, V->dev(), V->lda, W->dev(), W->lda, W->nrows, W->nvec, alpha, 0, zero_fn<T>, zero_fn<T>,
              nullptr);
}

template <typename T>
void blas_dot(const Mvec<T>* V, const Mvec<T>* W, T* out_host) {
  PHX_REQUIRE(V->nvec == W->nvec && V->nrows == W->nrows, PHIST_INVALID_INPUT, "block shapes differ (%d vs %d columns)",
              V->nvec, W->nvec);
  run_rows<T>(K_DOT, V->dev(), V->lda, const_cast<T*>(W->dev()), W->lda, W->nrows, W->nvec, T(0), 0, zero_fn<T>,
              zero_fn<T>, out_host);
}

// V(:, c) *= (T)(1 / sqrt(n2[c])) with the squared norms on the device (0
// when n2[c] is not positive): a grid-stride loop over the elements
template <typename T>
__global__ void __launch_bounds__(kBlock) k_scale_inv_norm(T* __restrict__ v, int64_t ld, int64_t nrows, int ncols,
                                                           const T* __restrict__ n2) {
  const int64_t n = nrows * ncols;
  for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < n; e += (int64_t)gridDim.x * kBlock) {
    const int64_t i = e / ncols;
    const int c = (int)(e - i * ncols);
    const double q = (double)n2[c];
    const T s = q > 0.0 ? (T)(1.0 / sqrt(q)) : T(0);
    v[i * ld + c] *= s;
  }
}

template <typename T>
void blas_scale_inv_norm_dev(Mvec<T>* V, const T* n2_dev) {
  const int64_t n = V->nrows * V->nvec;
  if (n == 0) return;
  const int grid = (int)std::min<int64_t>((n + kBlock - 1) / kBlock, (int64_t)rt().num_cu * 8);
  hipLaunchKernelGGL(k_scale_inv_norm<T>, dim3(std::max(grid, 1)), dim3(kBlock), 0, rt().stream, V->dev(), V->lda,
                     V->nrows, V->nvec, n2_dev);
  PHX_HIP(hipGetLastError());
}

template <typename T>
void blas_dot_dev(const Mvec<T>* V, const Mvec<T>* W, T* out_dev) {
  PHX_REQUIRE(V->nvec == W->nvec && V->nrows == W->nrows && V->nrows > 0, PHIST_INVALID_INPUT,
              "block shapes differ or are empty");
  run_rows<T>(K_DOT, V->dev(), V->lda, const_cast<T*>(W->dev()), W->lda, W->nrows, W->nvec, T(0), 0, zero_fn<T>,
              zero_fn<T>, nullptr, out_dev);
}

#define PHX_INST(T)                                                           \
  template void blas_dot_dev<T>(const Mvec<T>*, const Mvec<T>*, T*);          \
  template void blas_scale_inv_norm_dev<T>(Mvec<T>*, const T*);               \
  template void blas_put_value<T>(Mvec<T>*, T);                               \
  template void blas_scale_cols<T>(Mvec<T>*, const T*, bool);                 \
  template void blas_axpby<T>(const T*, bool, const Mvec<T>*, T, Mvec<T>*);   \
  template void blas_copy<T>(const Mvec<T>*, Mvec<T>*);                       \
  template void blas_elemwise<T>(T, const Mvec<T>*, Mvec<T>*);                \
  template void blas_dot<T>(const Mvec<T>*, const Mvec<T>*, T*);
PHX_INST(double)
PHX_INST(float)
#undef PHX_INST

}  // namespace phx
