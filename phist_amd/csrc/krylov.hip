// krylov.hip -- fused vector updates of the blocked MINRES iteration and
// precision conversion of block vectors.
//
// MINRES (Paige & Saunders) on k independent shifted systems at once, one
// column per system, the per-column scalars kept on the host:
//
//   k_lanczos    w(:,c) -= a_c * v(:,c) + b_c * vold(:,c), and the column
//                norms of the new w (fixed-order two-stage reduction)
//   k_minres     d_new = (v - delta_c*d1 - eps_c*d2) / gamma_c  (into d2),
//                x += tau_c * d_new,  vold = w / beta_c  (the next Lanczos
//                vector, written over the previous one)
//   k_convert    Y = X with a change of precision (FP64 <-> FP32), used for
//                the FP32 inner solves of the mixed-precision eigensolver
//
// Same mapping as the BLAS-1 kernels: LPR lanes per row with VPL consecutive
// columns each, persistent XCD-aware grid.
#include "amd_device.hpp"
#include <algorithm>

namespace phx {

constexpr int kKrylovMaxCols = 64;  // systems per block (coefficients travel as kernel arguments)

template <typename T>
struct KrylovArgs {
  int64_t nrows, ntiles;
  int ncols;
  T* v;
  T* vold;
  T* w;
  T* d1;
  T* d2;
  T* x;
  int64_t ld;  // common row stride of every block (elements)
  int flags;
  T* partials;
  T c0[kKrylovMaxCols], c1[kKrylovMaxCols], c2[kKrylovMaxCols], c3[kKrylovMaxCols], c4[kKrylovMaxCols];
};

enum : int { KF_VOLD = 1, KF_D2 = 2 };

template <typename T, int VPL, int LPR>
__global__ void __launch_bounds__(kBlock) k_lanczos(const KrylovArgs<T> p) {
  constexpr int RPW = 64 / LPR, RPB = (kBlock / 64) * RPW;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int q = lane % LPR, rw = lane / LPR, c0 = q * VPL;
  const bool colok = c0 < p.ncols;
  T a[VPL], b[VPL], acc[VPL];
#pragma unroll
  for (int e = 0; e < VPL; e++) {
    a[e] = c0 + e < p.ncols ? p.c0[c0 + e] : T(0);
    b[e] = c0 + e < p.ncols ? p.c1[c0 + e] : T(0);
    acc[e] = T(0);
  }
  const TileWalk tw = xcd_tiles(p.ntiles);
  for (int64_t t = tw.t; t < tw.end; t += tw.stride) {
    const int64_t i = t * RPB + wv * RPW + rw;
    if (i >= p.nrows || !colok) continue;
    const int64_t o = i * p.ld + c0;
    T vv[VPL], ww[VPL];
    vload<T, VPL>(p.v + o, vv);
    vload<T, VPL>(p.w + o, ww);
#pragma unroll
    for (int e = 0; e < VPL; e++) ww[e] = fma(-a[e], vv[e], ww[e]);
    if (p.flags & KF_VOLD) {
      T vo[VPL];
      vload<T, VPL>(p.vold + o, vo);
#pragma unroll
      for (int e = 0; e < VPL; e++) ww[e] = fma(-b[e], vo[e], ww[e]);
    }
    vstore<T, VPL>(p.w + o, ww);
#pragma unroll
    for (int e = 0; e < VPL; e++) acc[e] = fma(ww[e], ww[e], acc[e]);
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

template <typename T, int VPL, int LPR>
__global__ void __launch_bounds__(kBlock) k_minres(const KrylovArgs<T> p) {
  constexpr int RPW = 64 / LPR, RPB = (kBlock / 64) * RPW;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int q = lane % LPR, rw = lane / LPR, c0 = q * VPL;
  if (c0 >= p.ncols) return;
  T dl[VPL], ep[VPL], ig[VPL], ta[VPL], ib[VPL];
#pragma unroll
  for (int e = 0; e < VPL; e++) {
    const bool ok = c0 + e < p.ncols;
    dl[e] = ok ? p.c0[c0 + e] : T(0);
    ep[e] = ok ? p.c1[c0 + e] : T(0);
    ig[e] = ok ? p.c2[c0 + e] : T(0);
    ta[e] = ok ? p.c3[c0 + e] : T(0);
    ib[e] = ok ? p.c4[c0 + e] : T(0);
  }
  const TileWalk tw = xcd_tiles(p.ntiles);
  for (int64_t t = tw.t; t < tw.end; t += tw.stride) {
    const int64_t i = t * RPB + wv * RPW + rw;
    if (i >= p.nrows) continue;
    const int64_t o = i * p.ld + c0;
    T vv[VPL], d1[VPL], dn[VPL], xx[VPL], ww[VPL];
    vload<T, VPL>(p.v + o, vv);
    vload<T, VPL>(p.d1 + o, d1);
#pragma unroll
    for (int e = 0; e < VPL; e++) dn[e] = fma(-dl[e], d1[e], vv[e]);
    if (p.flags & KF_D2) {
      T d2[VPL];
      vload<T, VPL>(p.d2 + o, d2);
#pragma unroll
      for (int e = 0; e < VPL; e++) dn[e] = fma(-ep[e], d2[e], dn[e]);
    }
#pragma unroll
    for (int e = 0; e < VPL; e++) dn[e] *= ig[e];
    vstore<T, VPL>(p.d2 + o, dn);
    vload<T, VPL>(p.x + o, xx);
#pragma unroll
    for (int e = 0; e < VPL; e++) xx[e] = fma(ta[e], dn[e], xx[e]);
    vstore<T, VPL>(p.x + o, xx);
    vload<T, VPL>(p.w + o, ww);
#pragma unroll
    for (int e = 0; e < VPL; e++) ww[e] *= ib[e];
    vstore<T, VPL>(p.vold + o, ww);
  }
}

template <typename TS, typename TD>
__global__ void __launch_bounds__(kBlock) k_convert(const TS* __restrict__ x, int64_t ldx, TD* __restrict__ y,
                                                    int64_t ldy, int64_t nrows, int ncols) {
  const int64_t n = nrows * ncols;
  for (int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x; e < n; e += (int64_t)gridDim.x * kBlock) {
    const int64_t i = e / ncols;
    const int c = (int)(e - i * ncols);
    y[i * ldy + c] = (TD)x[i * ldx + c];
  }
}

namespace {

inline bool al16(const void* ptr) { return ((uintptr_t)ptr & 15) == 0; }

template <typename T, int VPL, int LPR>
int launch_lpr(bool minres, KrylovArgs<T>& p) {
  constexpr int RPB = (kBlock / 64) * (64 / LPR);
  p.ntiles = (p.nrows + RPB - 1) / RPB;
  int grid;
  if (minres) {
    grid = persistent_blocks(k_minres<T, VPL, LPR>, 0, p.ntiles);
    hipLaunchKernelGGL((k_minres<T, VPL, LPR>), dim3(grid), dim3(kBlock), 0, rt().stream, p);
  } else {
    grid = persistent_blocks(k_lanczos<T, VPL, LPR>, 0, p.ntiles);
    p.partials = (T*)device_work(sizeof(T) * ((size_t)grid * p.ncols + p.ncols));
    hipLaunchKernelGGL((k_lanczos<T, VPL, LPR>), dim3(grid), dim3(kBlock), 0, rt().stream, p);
  }
  PHX_HIP(hipGetLastError());
  return grid;
}

template <typename T, int VPL>
int launch_vpl(bool minres, KrylovArgs<T>& p) {
  const int groups = (p.ncols + VPL - 1) / VPL;
  int lpr = 1;
  while (lpr < groups) lpr <<= 1;
  switch (lpr) {
    case 1: return launch_lpr<T, VPL, 1>(minres, p);
    case 2: return launch_lpr<T, VPL, 2>(minres, p);
    case 4: return launch_lpr<T, VPL, 4>(minres, p);
    case 8: return launch_lpr<T, VPL, 8>(minres, p);
    case 16: return launch_lpr<T, VPL, 16>(minres, p);
    case 32: return launch_lpr<T, VPL, 32>(minres, p);
    default: return launch_lpr<T, VPL, 64>(minres, p);
  }
}

template <typename T>
int launch(bool minres, KrylovArgs<T>& p, std::initializer_list<const T*> ptrs) {
  constexpr int V16 = 16 / sizeof(T);
  bool vec = (p.ncols % V16 == 0) && (p.ld % V16 == 0);
  for (const T* q : ptrs) vec = vec && al16(q);
  PHX_REQUIRE(p.ncols <= kKrylovMaxCols, PHIST_INVALID_INPUT, "Krylov block of %d columns too wide", p.ncols);
  return vec ? launch_vpl<T, V16>(minres, p) : launch_vpl<T, 1>(minres, p);
}

}  // namespace

// w -= a.*v + b.*vold (vold skipped when b == 0 everywhere); norms of the new
// w columns returned in nrm2_host (squared norms summed in a fixed order,
// then square-rooted).  All blocks share the row stride ld.
template <typename T>
void krylov_lanczos(int64_t nrows, int ncols, int64_t ld, const T* v, const T* vold, T* w, const T* a, const T* b,
                    T* nrm_host) {
  if (nrows == 0 || ncols == 0) {
    for (int c = 0; c < ncols; c++) nrm_host[c] = T(0);
    return;
  }
  KrylovArgs<T> p;
  p.nrows = nrows;
  p.ncols = ncols;
  p.ld = ld;
  p.v = const_cast<T*>(v);
  p.vold = const_cast<T*>(vold);
  p.w = w;
  p.flags = 0;
  PHX_REQUIRE(ncols <= kKrylovMaxCols, PHIST_INVALID_INPUT, "Krylov block of %d columns too wide", ncols);
  for (int c = 0; c < ncols; c++) {
    p.c0[c] = a[c];
    p.c1[c] = b[c];
    if (b[c] != T(0)) p.flags |= KF_VOLD;
  }
  const int grid = launch(false, p, {v, w, (p.flags & KF_VOLD) ? vold : w});
  std::vector<T> sq(ncols);
  reduce_partials_to_host(p.partials, grid, ncols, sq.data());
  for (int c = 0; c < ncols; c++) nrm_host[c] = std::sqrt(sq[c]);
}

// d2 = (v - delta.*d1 - eps.*d2) .* inv_gamma;  x += tau.*d2;  vold = w .* inv_beta
template <typename T>
void krylov_minres_update(int64_t nrows, int ncols, int64_t ld, const T* v, T* vold, const T* w, const T* d1, T* d2,
                          T* x, const T* delta, const T* eps, const T* inv_gamma, const T* tau, const T* inv_beta) {
  if (nrows == 0 || ncols == 0) return;
  KrylovArgs<T> p;
  p.nrows = nrows;
  p.ncols = ncols;
  p.ld = ld;
  p.v = const_cast<T*>(v);
  p.vold = vold;
  p.w = const_cast<T*>(w);
  p.d1 = const_cast<T*>(d1);
  p.d2 = d2;
  p.x = x;
  p.flags = 0;
  PHX_REQUIRE(ncols <= kKrylovMaxCols, PHIST_INVALID_INPUT, "Krylov block of %d columns too wide", ncols);
  for (int c = 0; c < ncols; c++) {
    p.c0[c] = delta[c];
    p.c1[c] = eps[c];
    p.c2[c] = inv_gamma[c];
    p.c3[c] = tau[c];
    p.c4[c] = inv_beta[c];
    if (eps[c] != T(0)) p.flags |= KF_D2;
  }
  launch(true, p, {v, vold, w, d1, d2, x});
}

// Y = X converted to TD (same shape; any row strides)
template <typename TS, typename TD>
void convert_block(const TS* x, int64_t ldx, TD* y, int64_t ldy, int64_t nrows, int ncols) {
  if (nrows == 0 || ncols == 0) return;
  const int64_t n = nrows * ncols;
  const int grid = (int)std::min<int64_t>((n + kBlock - 1) / kBlock, (int64_t)rt().num_cu * 8);
  hipLaunchKernelGGL((k_convert<TS, TD>), dim3(std::max(grid, 1)), dim3(kBlock), 0, rt().stream, x, ldx, y, ldy, nrows,
                     ncols);
  PHX_HIP(hipGetLastError());
}

template void krylov_lanczos<double>(int64_t, int, int64_t, const double*, const double*, double*, const double*,
                                     const double*, double*);
template void krylov_lanczos<float>(int64_t, int, int64_t, const float*, const float*, float*, const float*,
                                    const float*, float*);
template void krylov_minres_update<double>(int64_t, int, int64_t, const double*, double*, const double*,
                                           const double*, double*, double*, const double*, const double*,
                                           const double*, const double*, const double*);
template void krylov_minres_update<float>(int64_t, int, int64_t, const float*, float*, const float*, const float*,
                                          float*, float*, const float*, const float*, const float*, const float*,
                                          const float*);
template void convert_block<double, float>(const double*, int64_t, float*, int64_t, int64_t, int);
template void convert_block<float, double>(const float*, int64_t, double*, int64_t, int64_t, int);

}  // namespace phx
