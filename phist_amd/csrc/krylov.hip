// krylov.hip -- fused vector updates of the blocked MINRES iteration and
// precision conversion of block vectors.
//
// MINRES (Paige & Saunders) on k independent shifted systems at once, one
// column per system, the per-column scalars on the host (kernel arguments) or,
// for krylov_minres_step_dev, on the device (column states, k_minres_rot):
//
//   k_lanczos    w(:,c) -= a_c * v(:,c) + b_c * vold(:,c), and the column
//                norms of the new w (fixed-order two-stage reduction); on the
//                device path a, b come from the column states and the SpMV dots
//   k_minres     d_new = (v - delta_c*d1 - eps_c*d2) / gamma_c  (into d2),
//                x += tau_c * d_new,  and on the host-scalar path
//                vold = w / beta_c (the next Lanczos vector, written over the
//                previous one; the device path keeps w unnormalised instead)
//   k_minres_rot one workgroup: the norms from the k_lanczos partials, then
//                the rotations of every column
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
  // coefficients: from cdev (device, coefficient q of column c at cdev[q*ncols + c]) when it is set,
  // else from the kernel arguments c0..c4
  const T* cdev;
  // optional projection fused into k_lanczos: w -= Q*C (Q nrows x nq, row stride ldq; C nq x ncols,
  // column-major with leading dimension ldcm, on the device)
  const T* q;
  int64_t ldq;
  int nq;
  const T* cmat;
  int ldcm;
  // optional per-column factors of v and vold (v(i,c) * scv[c * sc_stride]):
  // the device MINRES keeps its Lanczos vectors unnormalised
  const double* scv = nullptr;
  const double* scvold = nullptr;
  int64_t sc_stride = 0;
  // optional: the Lanczos coefficients a, b taken from the MINRES column
  // states and the fused SpMV dots (dots[ncols + c] = <v, Av>) instead of cdev
  const MinresState* st = nullptr;
  const T* dots = nullptr;
  // lanes of columns >= nact (when >= 0) skip their rows (same launch and sums for the others)
  int nact = -1;
  T c0[kKrylovMaxCols], c1[kKrylovMaxCols], c2[kKrylovMaxCols], c3[kKrylovMaxCols], c4[kKrylovMaxCols];
};

template <typename T>
__device__ inline T kcoef(const KrylovArgs<T>& p, int q, int c) {
  if (c >= p.ncols) return T(0);
  if (p.st && q < 2) {  // a_c = <v, Av>_c, b_c = beta_c after the first iteration; 0 when frozen
    const MinresState& s = p.st[c];
    if (q == 0) return s.active ? p.dots[p.ncols + c] : T(0);
    return (s.active && s.iters > 0) ? (T)s.beta : T(0);
  }
  if (p.cdev) return p.cdev[q * p.ncols + c];
  switch (q) {
    case 0: return p.c0[c];
    case 1: return p.c1[c];
    case 2: return p.c2[c];
    case 3: return p.c3[c];
    default: return p.c4[c];
  }
}

enum : int { KF_VOLD = 1, KF_D2 = 2, KF_NOVOLD = 4 };  // KF_NOVOLD: k_minres writes no vold

template <typename T, int VPL, int LPR>
__global__ void __launch_bounds__(kBlock) k_lanczos(const KrylovArgs<T> p) {
  constexpr int RPW = 64 / LPR, RPB = (kBlock / 64) * RPW;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int q = lane % LPR, rw = lane / LPR, c0 = q * VPL;
  const bool colok = c0 < (p.nact >= 0 ? p.nact : p.ncols);
  // projection coefficients C(j, c0+e) in LDS, row-interleaved: sc[j*VPL*LPR + c]
  extern __shared__ char lanczos_lds[];
  T* sc = reinterpret_cast<T*>(lanczos_lds);
  const int nq = p.nq;
  if (nq > 0) {
    const int width = VPL * LPR;
    for (int e = threadIdx.x; e < nq * width; e += kBlock) {
      const int j = e / width, c = e - j * width;
      sc[e] = c < p.ncols ? p.cmat[(int64_t)c * p.ldcm + j] : T(0);
    }
    __syncthreads();
  }
  T a[VPL], b[VPL], acc[VPL], sv[VPL], svo[VPL];
  const bool scaled = p.scv != nullptr;
#pragma unroll
  for (int e = 0; e < VPL; e++) {
    a[e] = kcoef(p, 0, c0 + e);
    b[e] = kcoef(p, 1, c0 + e);
    acc[e] = T(0);
    sv[e] = (scaled && colok) ? (T)p.scv[(int64_t)(c0 + e) * p.sc_stride] : T(1);
    svo[e] = (scaled && colok) ? (T)p.scvold[(int64_t)(c0 + e) * p.sc_stride] : T(1);
  }
  const TileWalk tw = xcd_tiles(p.ntiles);
  for (int64_t t = tw.t; t < tw.end; t += tw.stride) {
    const int64_t i = t * RPB + wv * RPW + rw;
    if (i >= p.nrows || !colok) continue;
    const int64_t o = i * p.ld + c0;
    T vv[VPL], ww[VPL];
    vload<T, VPL>(p.v + o, vv);
    vload<T, VPL>(p.w + o, ww);
    if (scaled) {
#pragma unroll
      for (int e = 0; e < VPL; e++) vv[e] *= sv[e];
    }
    if (nq > 0) {  // w(i,:) -= Q(i,:) * C, the Q row read 8 entries at a time (loads in flight together)
      const T* qi = p.q + i * p.ldq;
      for (int j0 = 0; j0 < nq; j0 += 8) {
        T qv[8];
#pragma unroll
        for (int u = 0; u < 8; u++) qv[u] = j0 + u < nq ? qi[j0 + u] : T(0);
#pragma unroll
        for (int u = 0; u < 8; u++) {
          if (j0 + u >= nq) break;
          const T* cj = sc + (j0 + u) * (VPL * LPR) + c0;
#pragma unroll
          for (int e = 0; e < VPL; e++) ww[e] = fma(-qv[u], cj[e], ww[e]);
        }
      }
    }
#pragma unroll
    for (int e = 0; e < VPL; e++) ww[e] = fma(-a[e], vv[e], ww[e]);
    if (p.flags & KF_VOLD) {
      T vo[VPL];
      vload<T, VPL>(p.vold + o, vo);
      if (scaled) {
#pragma unroll
        for (int e = 0; e < VPL; e++) vo[e] *= svo[e];
      }
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
  if (c0 >= (p.nact >= 0 ? p.nact : p.ncols)) return;
  T dl[VPL], ep[VPL], ig[VPL], ta[VPL], ib[VPL], sv[VPL];
  const bool scaled = p.scv != nullptr;
#pragma unroll
  for (int e = 0; e < VPL; e++) {
    dl[e] = kcoef(p, 0, c0 + e);
    ep[e] = kcoef(p, 1, c0 + e);
    ig[e] = kcoef(p, 2, c0 + e);
    ta[e] = kcoef(p, 3, c0 + e);
    ib[e] = kcoef(p, 4, c0 + e);
    sv[e] = scaled ? (T)p.scv[(int64_t)(c0 + e) * p.sc_stride] : T(1);
  }
  const TileWalk tw = xcd_tiles(p.ntiles);
  for (int64_t t = tw.t; t < tw.end; t += tw.stride) {
    const int64_t i = t * RPB + wv * RPW + rw;
    if (i >= p.nrows) continue;
    const int64_t o = i * p.ld + c0;
    T vv[VPL], d1[VPL], dn[VPL], xx[VPL], ww[VPL];
    vload<T, VPL>(p.v + o, vv);
    vload<T, VPL>(p.d1 + o, d1);
    if (scaled) {
#pragma unroll
      for (int e = 0; e < VPL; e++) vv[e] *= sv[e];
    }
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
    if (!(p.flags & KF_NOVOLD)) {
      vload<T, VPL>(p.w + o, ww);
#pragma unroll
      for (int e = 0; e < VPL; e++) ww[e] *= ib[e];
      vstore<T, VPL>(p.vold + o, ww);
    }
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

// One step of the Paige & Saunders recurrences per active column, in FP64:
// column `iters` of the Lanczos tridiagonal (beta_k, alpha_k, beta_{k+1})
// meets the two previous rotations, a new rotation annihilates beta_{k+1};
// writes the k_minres coefficients (delta, eps, 1/gamma, tau, 1/beta_{k+1})
// and freezes columns that converged, broke down or hit their limit.
// With phase >= 0 the factor of the next (unnormalised) Lanczos vector,
// 1/beta_{k+1} or 0 for a column that was frozen, goes to sc[(phase + 1) % 3].
// The squared norms of the new Lanczos vectors are first summed from the
// k_lanczos block partials (in the order of k_reduce_partials).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_minres_rot(const T* __restrict__ dots, const T* __restrict__ partials,
                                                       int nparts, MinresState* __restrict__ st, int k,
                                                       T* __restrict__ coef, int phase) {
  __shared__ T nrm2[kKrylovMaxCols];
  for (int g = 0; g * 16 < k; g++) {
    const T s = block_sum16(partials, nparts, k, g);
    if (threadIdx.x < 16 && g * 16 + (int)threadIdx.x < k) nrm2[g * 16 + threadIdx.x] = s;
  }
  __syncthreads();
  const int c = threadIdx.x;
  if (c >= k) return;
  MinresState q = st[c];
  T dl = T(0), ep = T(0), ig = T(0), ta = T(0), ib = T(0);
  double scn = 0.0;
  if (q.active) {
    const T bnext = sqrt(nrm2[c]);
    const double bk = q.iters > 0 ? q.beta : 0.0, ak = (double)dots[k + c], bn = (double)bnext;
    const double eps_k = q.s2 * bk, dbar = q.c2 * bk;
    const double delta = q.c1 * dbar + q.s1 * ak;
    const double gbar = -q.s1 * dbar + q.c1 * ak;
    const double gamma = hypot(gbar, bn);
    if (!(gamma > 0.0) || !isfinite(gamma)) {
      q.active = 0;
    } else {
      const double cs = gbar / gamma, sn = bn / gamma;
      const double tau = cs * q.phibar;
      q.phibar = -sn * q.phibar;
      dl = (T)delta;
      ep = (T)eps_k;
      ig = (T)(1.0 / gamma);
      ta = (T)tau;
      ib = bn > 0.0 ? (T)(1.0 / bn) : T(0);
      scn = bn > 0.0 ? 1.0 / bn : 0.0;
      q.c2 = q.c1;
      q.s2 = q.s1;
      q.c1 = cs;
      q.s1 = sn;
      q.beta = bn;
      q.iters++;
      q.relres = fabs(q.phibar) / q.beta1;
      if (q.relres <= q.tol || bn == 0.0 || q.iters >= q.max_iter) q.active = 0;
    }
  }
  st[c] = q;
  if (phase >= 0) st[c].sc[(phase + 1) % 3] = scn;  // (a run-time index into the local copy would put it in scratch)
  coef[c] = dl;
  coef[k + c] = ep;
  coef[2 * k + c] = ig;
  coef[3 * k + c] = ta;
  coef[4 * k + c] = ib;
}

// Column states j0 .. j0+n-1 of freshly loaded systems from the squared norms
// of their right-hand sides (nrm2[0..n)): beta1 = phibar = ||b||, active when
// ||b|| > 0 and iterations are allowed, and the factor 1/||b|| of the
// unnormalised first Lanczos vector in sc[ph].
template <typename T>
__global__ void k_minres_init(MinresState* __restrict__ st, const T* __restrict__ nrm2, int j0, int n, int ph) {
  const int c = threadIdx.x;
  if (c >= n) return;
  MinresState q = st[j0 + c];
  const T nt = sqrt(nrm2[c]);
  const double nn = (double)nt;
  q.beta1 = nn;
  q.phibar = nn;
  q.relres = nn > 0.0 ? 1.0 : 0.0;
  q.active = (nn > 0.0 && q.max_iter > 0) ? 1 : 0;
  st[j0 + c] = q;
  st[j0 + c].sc[ph] = nn > 0.0 ? 1.0 / nn : 0.0;
}

template <typename T>
void krylov_minres_init_dev(MinresState* st, const T* nrm2, int j0, int n, int ph) {
  PHX_REQUIRE(n >= 1 && n <= kKrylovMaxCols && ph >= 0 && ph < 3, PHIST_INVALID_INPUT, "MINRES init of %d columns", n);
  hipLaunchKernelGGL(k_minres_init<T>, dim3(1), dim3(64), 0, rt().stream, st, nrm2, j0, n, ph);
  PHX_HIP(hipGetLastError());
}
template void krylov_minres_init_dev<double>(MinresState*, const double*, int, int, int);
template void krylov_minres_init_dev<float>(MinresState*, const float*, int, int, int);

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
    const size_t lds = p.nq > 0 ? sizeof(T) * (size_t)p.nq * VPL * LPR : 0;
    grid = persistent_blocks(k_lanczos<T, VPL, LPR>, lds, p.ntiles);
    p.partials = (T*)device_work(sizeof(T) * ((size_t)grid * p.ncols + p.ncols));
    hipLaunchKernelGGL((k_lanczos<T, VPL, LPR>), dim3(grid), dim3(kBlock), lds, rt().stream, p);
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
  p.cdev = nullptr;
  p.q = nullptr;
  p.ldq = 0;
  p.nq = 0;
  p.cmat = nullptr;
  p.ldcm = 0;
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
  p.cdev = nullptr;
  p.q = nullptr;
  p.ldq = 0;
  p.nq = 0;
  p.cmat = nullptr;
  p.ldcm = 0;
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

// Device-resident variant of one MINRES iteration's vector work and scalars
// (no host synchronisation): given the fused SpMV dots (dots[k..2k) = <v, Av>)
// and the column states st, it forms w -= a.*v + b.*vold with the new norms,
// advances the recurrences on the device and updates d2, x and vold.
// work: device scratch of 7*k values of T (a, b, then the five k_minres
// coefficients) plus k squared norms.  With phase >= 0 the Lanczos vectors
// are unnormalised: v and vold are read times their factors in the column
// states (sc[phase % 3], sc[(phase + 2) % 3]), w itself becomes the next v
// (factor sc[(phase + 1) % 3]) and no vold is written.
template <typename T>
void krylov_minres_step_dev(int64_t nrows, int k, int64_t ld, const T* v, T* vold, T* w, const T* d1, T* d2, T* x,
                            const T* dots, MinresState* st, T* work, const T* Q, int64_t ldq, int nq, const T* C,
                            int ldc, int phase, int nact) {
  PHX_REQUIRE(k >= 1 && k <= kKrylovMaxCols, PHIST_INVALID_INPUT, "Krylov block of %d columns", k);
  if (nrows == 0) return;  // every column starts frozen (zero right-hand side)
  hipStream_t s = rt().stream;
  T* coef = work + 2 * k;
  KrylovArgs<T> p;
  p.nrows = nrows;
  p.ncols = k;
  p.ld = ld;
  p.v = const_cast<T*>(v);
  p.vold = vold;
  p.w = w;
  p.d1 = const_cast<T*>(d1);
  p.d2 = d2;
  p.x = x;
  p.flags = KF_VOLD;
  p.cdev = nullptr;
  p.st = st;  // a and b from the column states and the SpMV dots (no separate launch)
  p.dots = dots;
  p.nact = nact >= 0 ? std::min(nact, k) : -1;
  p.q = Q;
  p.ldq = ldq;
  p.nq = Q ? nq : 0;
  p.cmat = C;
  p.ldcm = ldc;
  if (phase >= 0) {
    p.scv = &st[0].sc[phase % 3];
    p.scvold = &st[0].sc[(phase + 2) % 3];
    p.sc_stride = sizeof(MinresState) / sizeof(double);
  }
  const int grid = launch(false, p, {v, w, vold});
  p.q = nullptr;
  p.nq = 0;
  p.st = nullptr;
  hipLaunchKernelGGL(k_minres_rot<T>, dim3(1), dim3(kBlock), 0, s, dots, p.partials, grid, st, k, coef, phase);
  PHX_HIP(hipGetLastError());
  p.flags = KF_D2 | (phase >= 0 ? KF_NOVOLD : 0);
  p.cdev = coef;
  p.scvold = nullptr;
  launch(true, p, {v, vold, w, d1, d2, x});
}

// ---------------------------------------------------------------------------
// Classical Gram-Schmidt of the blocked GMRES: kk systems, each with its own
// basis, laid side by side (column l*kk + a of V is basis vector l of system
// a).  System a only ever meets its own columns, so a sweep reads every basis
// entry once and does one multiply-add with it (no Gram of the whole block):
//   mode 0: D(l,a) = <V(:,l*kk+a), w(:,a)>
//   mode 1: w(:,a) -= sum_l V(:,l*kk+a) C(l,a), then D(l,a) of the new w
//   mode 2: as mode 1, then N(a) = ||new w(:,a)||^2
//   mode 3: as mode 1 without outputs (x += V y at the end of a cycle)
// with C(l,a) = coef[l*kk+a] where act[a] != 0, else 0 (act == nullptr: all).
// Thread (rl, a) of a workgroup owns system a of row t*RP + rl (RP = 256/KP,
// KP the next power of two >= kk) and keeps the row's nl basis entries in
// registers (buffer loads with the vector's offset as scalar operand); the outputs are reduced per workgroup (lanes of a wave that
// share a, then the four waves in order) and over workgroups by
// k_reduce_partials, so every sum has a fixed order.
template <typename T>
struct CgsArgs {
  int64_t nrows, ntiles;
  const T* v;
  int64_t ldv;
  T* w;
  int64_t ldw;
  int kk, kp, nl;
  const T* coef;
  const T* act;
  T* partials;
};

template <typename T, int LMAX, int MODE>
__global__ void __launch_bounds__(kBlock) k_gmres_cgs(const CgsArgs<T> p) {
  extern __shared__ unsigned char cgs_lds[];
  T* sh = reinterpret_cast<T*>(cgs_lds);
  const int kp = p.kp, rp = kBlock / kp;
  const int a = threadIdx.x % kp, rl = threadIdx.x / kp;
  const bool live = a < p.kk;
  // coefficients in LDS after the reduction area (registers hold the row's basis entries and the sums)
  // (zero for l >= nl, so the loads of those slots, which repeat vector nl-1, add nothing)
  T* cf = sh + (MODE == 3 ? 0 : (MODE == 2 ? p.kk : p.nl * p.kk));
  if (MODE > 0) {
    for (int e = threadIdx.x; e < LMAX * p.kk; e += kBlock) {
      const int c = e % p.kk;
      cf[e] = (e < p.nl * p.kk && (!p.act || p.act[c] != T(0))) ? p.coef[e] : T(0);
    }
    __syncthreads();
  }
  T acc[LMAX];
  T nacc = T(0);
#pragma unroll
  for (int l = 0; l < LMAX; l++) acc[l] = T(0);
  const TileWalk tw = xcd_tiles(p.ntiles);
  for (int64_t t = tw.t; t < tw.end; t += tw.stride) {
    const int64_t r = t * rp + rl;
    if (!live || r >= p.nrows) continue;
    // buffer loads: the tile's rows as the resource (uniform base), the lane's offset in one register
    // and the vector's offset l*kk (clamped to nl-1: every slot loads, no branch) as the scalar offset
    const int64_t rows = p.nrows - t * rp < rp ? p.nrows - t * rp : rp;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<T*>(p.v + t * rp * p.ldv), (short)0, (int)(rows * p.ldv * (int64_t)sizeof(T)), 0x00020000);
    const int vo = (int)((rl * p.ldv + a) * sizeof(T));
    T vv[LMAX];
#pragma unroll
    for (int l = 0; l < LMAX; l++) {
      const int so = (l < p.nl ? l : p.nl - 1) * p.kk * (int)sizeof(T);
      if constexpr (sizeof(T) == 8)
        vv[l] = __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(rs, vo, so, 0));
      else
        vv[l] = __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(rs, vo, so, 0));
    }
    T* wr = p.w + r * p.ldw + a;
    T s = *wr;
    // keep every load of the row in flight together (the scheduler would otherwise pair each load with
    // its use to save registers, one memory latency per vector)
    __builtin_amdgcn_sched_barrier(0);
    if (MODE > 0) {
#pragma unroll
      for (int l = 0; l < LMAX; l++) s -= vv[l] * cf[l * p.kk + a];
      *wr = s;
    }
    if (MODE == 0 || MODE == 1) {
#pragma unroll
      for (int l = 0; l < LMAX; l++) acc[l] += vv[l] * s;
    } else if (MODE == 2) {
      nacc += s * s;
    }
  }
  if (MODE == 3) return;
  // workgroup sums: lanes of a wave with the same a, then the waves in order
  const int nout = MODE == 2 ? p.kk : p.nl * p.kk;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  constexpr int nouts = MODE == 2 ? 1 : LMAX;
  if (MODE == 2) acc[0] = nacc;
  // (acc is indexed by compile-time l only, so it stays in registers)
  for (int off = kp; off < 64; off <<= 1) {
#pragma unroll
    for (int l = 0; l < nouts; l++) acc[l] += __shfl_xor(acc[l], off, 64);
  }
#pragma unroll 1
  for (int w = 0; w < kBlock / 64; w++) {
    if (wv == w && lane < kp && live) {
#pragma unroll
      for (int l = 0; l < nouts; l++) {
        if (MODE == 2 || l < p.nl) {
          const int o = MODE == 2 ? a : l * p.kk + a;
          sh[o] = w == 0 ? acc[l] : sh[o] + acc[l];
        }
      }
    }
    __syncthreads();
  }
  for (int o = threadIdx.x; o < nout; o += kBlock) p.partials[(int64_t)blockIdx.x * nout + o] = sh[o];
}

template <typename T, int LMAX, int MODE>
static void cgs_launch(CgsArgs<T> p, int grid, size_t lds) {
  hipLaunchKernelGGL((k_gmres_cgs<T, LMAX, MODE>), dim3(grid), dim3(kBlock), lds, rt().stream, p);
  PHX_HIP(hipGetLastError());
}

template <typename T, int MODE>
static void cgs_lmax(CgsArgs<T> p, int grid, size_t lds) {
  switch ((p.nl + 3) / 4) {  // register slots: nl rounded up to a multiple of 4
    case 1: cgs_launch<T, 4, MODE>(p, grid, lds); break;
    case 2: cgs_launch<T, 8, MODE>(p, grid, lds); break;
    case 3: cgs_launch<T, 12, MODE>(p, grid, lds); break;
    case 4: cgs_launch<T, 16, MODE>(p, grid, lds); break;
    case 5: cgs_launch<T, 20, MODE>(p, grid, lds); break;
    case 6: cgs_launch<T, 24, MODE>(p, grid, lds); break;
    case 7: cgs_launch<T, 28, MODE>(p, grid, lds); break;
    default: cgs_launch<T, 32, MODE>(p, grid, lds); break;
  }
}

// Workgroups of a gmres_cgs_dev sweep (the partial sums need grid * nout
// elements of work).
int gmres_cgs_grid(int64_t nrows, int kk) {
  int kp = 1;
  while (kp < kk) kp <<= 1;
  const int64_t ntiles = (nrows + kBlock / kp - 1) / (kBlock / kp);
  return persistent_blocks(k_gmres_cgs<double, 32, 1>, 2 * kGmresCgsMaxNl * 64 * sizeof(double), ntiles, 4);
}

template <typename T>
void gmres_cgs_dev(int mode, int64_t nrows, const T* v, int64_t ldv, T* w, int64_t ldw, int kk, int nl,
                   const T* coef, const T* act, T* work, int64_t work_elems, T* out) {
  PHX_REQUIRE(kk >= 1 && kk <= kGmresCgsMaxKk && nl >= 1 && nl <= kGmresCgsMaxNl && mode >= 0 && mode <= 3,
              PHIST_INVALID_INPUT, "Gram-Schmidt sweep over %d systems x %d vectors (mode %d)", kk, nl, mode);
  PHX_REQUIRE((int64_t)kBlock * ldv * (int64_t)sizeof(T) < (int64_t(1) << 31), PHIST_INVALID_INPUT,
              "basis row stride %ld too large", (long)ldv);  // a tile's rows fit one buffer resource
  CgsArgs<T> p;
  p.nrows = nrows;
  p.v = v;
  p.ldv = ldv;
  p.w = w;
  p.ldw = ldw;
  p.kk = kk;
  p.kp = 1;
  while (p.kp < kk) p.kp <<= 1;
  p.nl = nl;
  p.coef = coef;
  p.act = act;
  p.partials = work;
  p.ntiles = (nrows + kBlock / p.kp - 1) / (kBlock / p.kp);
  const int nout = mode == 2 ? kk : nl * kk;
  const int grid = gmres_cgs_grid(nrows, kk);
  PHX_REQUIRE(mode == 3 || (int64_t)grid * nout + nout <= work_elems, PHIST_FUNCTIONAL_ERROR,
              "Gram-Schmidt work of %ld elements is too small", (long)work_elems);
  const int lmax = (nl + 3) / 4 * 4;
  const size_t lds = sizeof(T) * ((mode == 3 ? 0 : nout) + (mode > 0 ? (size_t)lmax * kk : 0));
  switch (mode) {
    case 0: cgs_lmax<T, 0>(p, grid, lds); break;
    case 1: cgs_lmax<T, 1>(p, grid, lds); break;
    case 2: cgs_lmax<T, 2>(p, grid, lds); break;
    default: cgs_lmax<T, 3>(p, grid, lds); return;
  }
  hipLaunchKernelGGL(k_reduce_partials<T>, dim3((nout + 15) / 16), dim3(kBlock), 0, rt().stream, work, grid, nout,
                     out);
  PHX_HIP(hipGetLastError());
}

// One Arnoldi step's Givens rotations of the blocked GMRES on the device, one
// thread per system a (the host loop's arithmetic, in double): from the
// sweeps' sums outs = D1 | D2 | N (D1, D2: (i+1)*kk coefficients of the two
// passes, N: the squared norms of w) it forms Hessenberg column i, applies
// the system's earlier rotations and a new one, updates g, and decides
// convergence (|g_{i+1}| < 1000 eps or |g_{i+1}| / ||r0|| < tol, or a zero
// norm; the iteration limit; a non-finite norm fails the system); act[a]
// (the sweeps' mask) and inv[a] (1 / ||w||, 0 for stopped systems) are
// written for the next kernels.
template <typename T>
__global__ void __launch_bounds__(64) k_gmres_givens(GmresDevState st, int kk, int m, int i, const T* __restrict__ outs,
                                                     T* act, T* inv, double eps1000, int max_iter) {
  const int a = threadIdx.x;
  if (a >= kk) return;
  const int nc = (i + 1) * kk;
  const double hn = sqrt((double)outs[2 * nc + a]);
  const bool live = act[a] != T(0);
  inv[a] = (live && hn > 0.0 && isfinite(hn)) ? T(1.0 / hn) : T(0);
  if (!live) return;
  if (!isfinite(hn)) {
    act[a] = T(0);
    st.status[a] = 3;
    return;
  }
  const int m1 = m + 1;
  double* h = st.H + ((size_t)a * m + i) * m1;  // column i of system a, rows 0..m
  double* cs = st.cs + (size_t)a * m;
  double* sn = st.sn + (size_t)a * m;
  double* g = st.g + (size_t)a * m1;
  for (int l = 0; l <= i; l++) h[l] = (double)outs[l * kk + a] + (double)outs[nc + l * kk + a];
  h[i + 1] = hn;
  for (int l = 0; l < i; l++) {
    const double c = cs[l], sv = sn[l];
    const double p0 = h[l], p1 = h[l + 1];
    h[l] = c * p0 + sv * p1;
    h[l + 1] = -sv * p0 + c * p1;
  }
  const double p0 = h[i], p1 = h[i + 1];
  const double r = hypot(p0, p1);
  const double c = r > 0.0 ? p0 / r : 1.0, sv = r > 0.0 ? p1 / r : 0.0;
  cs[i] = c;
  sn[i] = sv;
  h[i] = r;
  h[i + 1] = 0.0;
  g[i + 1] = -sv * g[i];
  g[i] = c * g[i];
  st.steps[a] = i + 1;
  const double nr = fabs(g[i + 1]);
  const double rel = nr / st.normR0[a];
  st.relres[a] = rel;
  if (nr < eps1000 || rel < st.tol[a] || hn == 0.0) {
    act[a] = T(0);
    st.status[a] = 0;
  } else if (st.iters0[a] + i + 1 >= max_iter) {
    act[a] = T(0);
    st.status[a] = 3;
  }
}

// v = w * inv (column a by inv[a]; 0 where ||w(:,a)||^2 = nrm2[a] is not
// finite, so a broken column cannot reach other systems as 0 * NaN) written
// to both vcur and the basis block vb.  A lane owns VS consecutive systems of
// a row (16-byte vectors where kk and the row strides allow, else one), kp
// lanes per row (a power of two), and four rows per tile whose loads all go
// out before the stores.
template <typename T, int VS>
__global__ void __launch_bounds__(kBlock) k_gmres_normalize(int64_t nrows, int kk, int kshift, const T* __restrict__ w,
                                                            int64_t ldw, const T* __restrict__ inv,
                                                            const T* __restrict__ nrm2, T* __restrict__ vcur,
                                                            int64_t ldc, T* __restrict__ vb, int64_t ldv) {
  constexpr int U = 4;
  const int q = threadIdx.x & ((1 << kshift) - 1), a0 = q * VS;
  if (a0 >= kk) return;
  T f[VS];
#pragma unroll
  for (int e = 0; e < VS; e++) {
    const bool ok = isfinite((double)nrm2[a0 + e]);
    f[e] = ok ? inv[a0 + e] : T(0);
  }
  bool zero[VS];
#pragma unroll
  for (int e = 0; e < VS; e++) zero[e] = !isfinite((double)nrm2[a0 + e]);
  const int64_t rpb = kBlock >> kshift;
  const int64_t rl = threadIdx.x >> kshift;
  for (int64_t r0 = (int64_t)blockIdx.x * rpb * U; r0 < nrows; r0 += (int64_t)gridDim.x * rpb * U) {
    T x[U][VS];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t r = r0 + u * rpb + rl;
      if (r < nrows) vload<T, VS>(w + r * ldw + a0, x[u]);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t r = r0 + u * rpb + rl;
      if (r >= nrows) continue;
      T v[VS];
#pragma unroll
      for (int e = 0; e < VS; e++) v[e] = zero[e] ? T(0) : f[e] * x[u][e];
      vstore<T, VS>(vcur + r * ldc + a0, v);
      vstore<T, VS>(vb + r * ldv + a0, v);
    }
  }
}

template <typename T>
void gmres_givens_dev(const GmresDevState& st, int kk, int m, int i, const T* outs, T* act, T* inv, double eps1000,
                      int max_iter) {
  PHX_REQUIRE(kk >= 1 && kk <= 64, PHIST_INVALID_INPUT, "%d systems (1..64)", kk);
  hipLaunchKernelGGL(k_gmres_givens<T>, dim3(1), dim3(64), 0, rt().stream, st, kk, m, i, outs, act, inv, eps1000,
                     max_iter);
  PHX_HIP(hipGetLastError());
}

template <typename T>
void gmres_normalize_dev(int64_t nrows, int kk, const T* w, int64_t ldw, const T* inv, const T* nrm2, T* vcur,
                         int64_t ldc, T* vb, int64_t ldv) {
  if (nrows == 0) return;
  PHX_REQUIRE(kk >= 1 && kk <= 64, PHIST_INVALID_INPUT, "%d systems (1..64)", kk);
  constexpr int V = 16 / sizeof(T);
  auto al = [](const void* ptr, int64_t ld) { return ((uintptr_t)ptr % 16) == 0 && (ld * (int64_t)sizeof(T)) % 16 == 0; };
  const bool vec = kk % V == 0 && al(w, ldw) && al(vcur, ldc) && al(vb, ldv);
  const int lanes = vec ? kk / V : kk;
  int ks = 0;
  while ((1 << ks) < lanes) ks++;
  const int64_t rows_per_tile = (int64_t)(kBlock >> ks) * 4;
  const int grid =
      (int)std::max<int64_t>(1, std::min<int64_t>((nrows + rows_per_tile - 1) / rows_per_tile, (int64_t)rt().num_cu * 8));
  if (vec)
    hipLaunchKernelGGL((k_gmres_normalize<T, V>), dim3(grid), dim3(kBlock), 0, rt().stream, nrows, kk, ks, w, ldw, inv,
                       nrm2, vcur, ldc, vb, ldv);
  else
    hipLaunchKernelGGL((k_gmres_normalize<T, 1>), dim3(grid), dim3(kBlock), 0, rt().stream, nrows, kk, ks, w, ldw, inv,
                       nrm2, vcur, ldc, vb, ldv);
  PHX_HIP(hipGetLastError());
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
template void krylov_minres_step_dev<double>(int64_t, int, int64_t, const double*, double*, double*, const double*,
                                             double*, double*, const double*, MinresState*, double*, const double*,
                                             int64_t, int, const double*, int, int, int);
template void krylov_minres_step_dev<float>(int64_t, int, int64_t, const float*, float*, float*, const float*, float*,
                                            float*, const float*, MinresState*, float*, const float*, int64_t, int,
                                            const float*, int, int, int);
template void gmres_cgs_dev<double>(int, int64_t, const double*, int64_t, double*, int64_t, int, int,
                                    const double*, const double*, double*, int64_t, double*);
template void gmres_cgs_dev<float>(int, int64_t, const float*, int64_t, float*, int64_t, int, int, const float*,
                                   const float*, float*, int64_t, float*);
template void gmres_givens_dev<double>(const GmresDevState&, int, int, int, const double*, double*, double*, double,
                                       int);
template void gmres_givens_dev<float>(const GmresDevState&, int, int, int, const float*, float*, float*, double, int);
template void gmres_normalize_dev<double>(int64_t, int, const double*, int64_t, const double*, const double*, double*,
                                          int64_t, double*, int64_t);
template void gmres_normalize_dev<float>(int64_t, int, const float*, int64_t, const float*, const float*, float*,
                                         int64_t, float*, int64_t);
template void convert_block<double, float>(const double*, int64_t, float*, int64_t, int64_t, int);
template void convert_block<float, double>(const float*, int64_t, double*, int64_t, int64_t, int);

}  // namespace phx
