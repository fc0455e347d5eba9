// jada.cpp -- device-resident eigensolver and shifted linear solvers.
//
//   minres_block     k shifted symmetric systems P(A + s_j I)P x_j = b_j at
//                    once (P = I - QQ' or I), MINRES with Paige & Saunders'
//                    short recurrences: one shifted SpMV with the fused
//                    <v, Av> dots, the projection (Gram + row mix), one fused
//                    Lanczos kernel (update + norms) and one fused solution
//                    update per iteration; the 2x2 rotations of the k
//                    columns run on the host.  Algorithmically the systems
//                    phist_blockedMINRESstates_iterate solves
//                    (src/krylov/phist_blockedminres_def.hpp:201-440), with
//                    the solution formed by the three-term direction
//                    recurrence instead of a stored basis.
//   subspacejada     block Jacobi-Davidson QR for exterior eigenpairs of a
//                    symmetric operator (src/jada/phist_subspacejada_def.hpp:
//                    Arnoldi start :302-345, Ritz extraction :455-512,
//                    residuals :585-604, locking :713-771, selection and
//                    inner tolerances :774-853, correction solve and basis
//                    extension :861-904); the correction equations
//                    (I - QQ')(A - sigma_j I) t_j = -r_j (phist_jadaOp_def.hpp)
//                    go to minres_block, in FP64 or on an FP32 copy of A.
//   linearOp_wrap_sparseMat, jadaOpts_setDefaults: the reference's operator
//                    wrapper (src/core/phist_operator_def.hpp:84-106) and
//                    option defaults (src/jada/phist_jadaOpts.cpp:19-58).
#include "objects.hpp"
#include "phist_amd_jada.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

namespace phx {

namespace dense {
void sym_eig(double* a, int n, int lda, double* w);
}

template <typename T>
void krylov_lanczos(int64_t nrows, int ncols, int64_t ld, const T* v, const T* vold, T* w, const T* a, const T* b,
                    T* nrm_host);
template <typename T>
void krylov_minres_update(int64_t nrows, int ncols, int64_t ld, const T* v, T* vold, const T* w, const T* d1, T* d2,
                          T* x, const T* delta, const T* eps, const T* inv_gamma, const T* tau, const T* inv_beta);
template <typename TS, typename TD>
void convert_block(const TS* x, int64_t ldx, TD* y, int64_t ldy, int64_t nrows, int ncols);

namespace {

bool g_inner_fp32 = false;
bool g_inner_fp32_init = false;

// PHIST_AMD_JADA_VERBOSE=1 prints one line per outer iteration to stderr
bool jada_verbose() {
  static const bool v = [] {
    const char* s = getenv("PHIST_AMD_JADA_VERBOSE");
    return s && atoi(s) != 0;
  }();
  return v;
}

bool inner_fp32() {
  if (!g_inner_fp32_init) {
    if (const char* s = getenv("PHIST_AMD_JADA_INNER_FP32")) g_inner_fp32 = atoi(s) != 0;
    g_inner_fp32_init = true;
  }
  return g_inner_fp32;
}

// owned block vector, freed at scope exit
template <typename T>
struct MvBox {
  Mvec<T>* p = nullptr;
  MvBox() = default;
  MvBox(const Map* map, int nvec) : p(api::mvec_create<T>(map, nvec)) {}
  ~MvBox() {
    try {
      if (p) api::mvec_delete(p);
    } catch (...) {
    }
  }
  MvBox(const MvBox&) = delete;
  MvBox& operator=(const MvBox&) = delete;
  Mvec<T>* operator->() const { return p; }
  Mvec<T>* get() const { return p; }
};

// column view [j0, j1] of a block vector, deleted at scope exit
template <typename T>
struct View {
  Mvec<T>* p = nullptr;
  View(const Mvec<T>* V, int j0, int j1) { api::mvec_view_block(const_cast<Mvec<T>*>(V), &p, j0, j1); }
  ~View() {
    try {
      if (p) api::mvec_delete(p);
    } catch (...) {
    }
  }
  View(const View&) = delete;
  View& operator=(const View&) = delete;
  Mvec<T>* operator->() const { return p; }
  Mvec<T>* get() const { return p; }
};

template <typename T>
struct SdBox {
  SdMat<T>* p = nullptr;
  SdBox(int m, int n) : p(api::sdMat_create<T>(std::max(1, m), std::max(1, n))) {}
  ~SdBox() {
    try {
      if (p) api::sdMat_delete(p);
    } catch (...) {
    }
  }
  SdBox(const SdBox&) = delete;
  SdBox& operator=(const SdBox&) = delete;
  SdMat<T>* operator->() const { return p; }
  SdMat<T>* get() const { return p; }
};

template <typename T>
struct SdView {
  SdMat<T>* p = nullptr;
  SdView(SdMat<T>* M, int i0, int i1, int j0, int j1) { api::sdMat_view_block(M, &p, i0, i1, j0, j1); }
  ~SdView() {
    try {
      if (p) api::sdMat_delete(p);
    } catch (...) {
    }
  }
  SdView(const SdView&) = delete;
  SdView& operator=(const SdView&) = delete;
  SdMat<T>* operator->() const { return p; }
  SdMat<T>* get() const { return p; }
};

// ---------------------------------------------------------------------------
// operators: either a sparse matrix of this library (device fast path) or a
// user linearOp called through its function pointers
// ---------------------------------------------------------------------------
template <typename T> struct LinOpT;
template <> struct LinOpT<double> {
  using type = phist_DlinearOp;
  using mvec = phist_Dmvec;
  using sd = phist_DsdMat;
  static void apply(double alpha, const void* A, const phist_Dmvec* X, double beta, phist_Dmvec* Y, int* iflag) {
    phist_DsparseMat_times_mvec(alpha, (phist_Dconst_sparseMat_ptr)A, X, beta, Y, iflag);
  }
  static void applyT(double, const void*, const phist_Dmvec*, double, phist_Dmvec*, int* iflag) {
    *iflag = PHIST_NOT_IMPLEMENTED;
  }
  static void apply_shifted(double alpha, const void* A, const double* sigma, const phist_Dmvec* X, double beta,
                            phist_Dmvec* Y, int* iflag) {
    phist_DsparseMat_times_mvec_vadd_mvec(alpha, (phist_Dconst_sparseMat_ptr)A, sigma, X, beta, Y, iflag);
  }
  static void fused(double alpha, const void* A, const phist_Dmvec* V, double beta, phist_Dmvec* W, phist_DsdMat* WtW,
                    phist_DsdMat* VtW, int* iflag) {
    phist_Dfused_spmv_mvTmv(alpha, (phist_Dconst_sparseMat_ptr)A, V, beta, W, WtW, VtW, iflag);
  }
  static void destroy(phist_DlinearOp*, int* iflag) { *iflag = 0; }
};
template <> struct LinOpT<float> {
  using type = phist_SlinearOp;
  using mvec = phist_Smvec;
  using sd = phist_SsdMat;
  static void apply(float alpha, const void* A, const phist_Smvec* X, float beta, phist_Smvec* Y, int* iflag) {
    phist_SsparseMat_times_mvec(alpha, (phist_Sconst_sparseMat_ptr)A, X, beta, Y, iflag);
  }
  static void applyT(float, const void*, const phist_Smvec*, float, phist_Smvec*, int* iflag) {
    *iflag = PHIST_NOT_IMPLEMENTED;
  }
  static void apply_shifted(float alpha, const void* A, const float* sigma, const phist_Smvec* X, float beta,
                            phist_Smvec* Y, int* iflag) {
    phist_SsparseMat_times_mvec_vadd_mvec(alpha, (phist_Sconst_sparseMat_ptr)A, sigma, X, beta, Y, iflag);
  }
  static void fused(float alpha, const void* A, const phist_Smvec* V, float beta, phist_Smvec* W, phist_SsdMat* WtW,
                    phist_SsdMat* VtW, int* iflag) {
    phist_Sfused_spmv_mvTmv(alpha, (phist_Sconst_sparseMat_ptr)A, V, beta, W, WtW, VtW, iflag);
  }
  static void destroy(phist_SlinearOp*, int* iflag) { *iflag = 0; }
};

// the sparse matrix behind an operator made by linearOp_wrap_sparseMat
template <typename T>
const SparseMat<T>* own_matrix(const typename LinOpT<T>::type* op) {
  if (op && op->apply == &LinOpT<T>::apply && op->apply_shifted == &LinOpT<T>::apply_shifted)
    return reinterpret_cast<const SparseMat<T>*>(op->A);
  return nullptr;
}

// Y = alpha*(A + diag(s))*X + beta*Y with optional column dots <X, Y>
template <typename T>
struct Operator {
  const SparseMat<T>* A = nullptr;
  const typename LinOpT<T>::type* user = nullptr;
  void apply_shifted(T alpha, const T* s, const Mvec<T>* X, T beta, Mvec<T>* Y, T* xy_dots) const {
    if (A) {
      SpmvRequest req;
      req.has_shift = s != nullptr;
      req.has_beta = beta != T(0);
      req.dot_vw = xy_dots != nullptr;
      spmv_launch<T>(A, alpha, s, X, beta, Y, req, nullptr, xy_dots);
      return;
    }
    int iflag = 0;
    auto* x = reinterpret_cast<const typename LinOpT<T>::mvec*>(X);
    auto* y = reinterpret_cast<typename LinOpT<T>::mvec*>(Y);
    if (s) {
      PHX_REQUIRE(user->apply_shifted != nullptr, PHIST_NOT_IMPLEMENTED, "operator has no apply_shifted");
      user->apply_shifted(alpha, user->A, s, x, beta, y, &iflag);
    } else {
      user->apply(alpha, user->A, x, beta, y, &iflag);
    }
    PHX_REQUIRE(iflag == 0, iflag, "user operator returned %d", iflag);
    if (xy_dots) blas_dot<T>(X, Y, xy_dots);
  }
};

// ---------------------------------------------------------------------------
// blocked MINRES on P(A + s_j I)P x_j = b_j, x = 0 on entry
// ---------------------------------------------------------------------------
template <typename T>
struct MinresOut {
  std::vector<int> iters;
  std::vector<double> relres;
  int unconverged = 0;
};

template <typename T>
MinresOut<T> minres_block(const Operator<T>& op, const std::vector<T>& s, const Mvec<T>* Q, const Mvec<T>* B,
                           Mvec<T>* X, const std::vector<double>& tol, int maxIter) {
  const int k = B->nvec;
  const Map* map = B->map;
  const int64_t n = B->nrows;
  MinresOut<T> out;
  out.iters.assign(k, 0);
  out.relres.assign(k, 1.0);
  // five k-column work blocks with one row stride: v, vold, w, d1, d2 and x
  MvBox<T> work(map, 6 * k);
  View<T> v0(work.get(), 0, k - 1), v1(work.get(), k, 2 * k - 1), wv(work.get(), 2 * k, 3 * k - 1),
      da(work.get(), 3 * k, 4 * k - 1), db(work.get(), 4 * k, 5 * k - 1), xv(work.get(), 5 * k, 6 * k - 1);
  const int64_t ld = work->lda;
  Mvec<T>* v = v0.get();
  Mvec<T>* vold = v1.get();
  Mvec<T>* w = wv.get();
  Mvec<T>* d1 = da.get();
  Mvec<T>* d2 = db.get();
  blas_put_value<T>(work.get(), T(0));

  std::vector<T> beta1(k), sc(k);
  blas_dot<T>(B, B, beta1.data());
  for (int j = 0; j < k; j++) {
    beta1[j] = std::sqrt(beta1[j]);
    sc[j] = beta1[j] > T(0) ? T(1) / beta1[j] : T(0);
  }
  blas_copy<T>(B, v);  // v = b / beta1 (B may have another row stride)
  blas_scale_cols<T>(v, sc.data(), false);
  // per-column recurrence state
  std::vector<double> beta(k, 0.0), c1(k, 1.0), s1(k, 0.0), c2(k, 1.0), s2(k, 0.0), phibar(k);
  std::vector<char> active(k, 1);
  for (int j = 0; j < k; j++) {
    phibar[j] = (double)beta1[j];
    if (beta1[j] == T(0)) {
      active[j] = 0;
      out.relres[j] = 0.0;
    }
  }
  SdBox<T> C(Q ? Q->nvec : 1, k);
  std::vector<T> alpha(k), bnext(k), la(k), lb(k), dl(k), ep(k), ig(k), ta(k), ib(k);
  for (int it = 1; it <= maxIter; it++) {
    bool any = false;
    for (int j = 0; j < k; j++) any = any || active[j];
    if (!any) break;
    // w = P (A + s) v, alpha = <v, w>
    op.apply_shifted(T(1), s.data(), v, T(0), w, alpha.data());
    if (Q && Q->nvec > 0) {
      gemm_gram<T>(T(1), Q, w, T(0), C.get());
      gemm_times_sd<T>(T(-1), Q, C.get(), T(1), w);
    }
    for (int j = 0; j < k; j++) {
      la[j] = active[j] ? alpha[j] : T(0);
      lb[j] = active[j] && it > 1 ? (T)beta[j] : T(0);
    }
    krylov_lanczos<T>(n, k, ld, v->dev(), vold->dev(), w->dev(), la.data(), lb.data(), bnext.data());
    for (int j = 0; j < k; j++) {
      dl[j] = ep[j] = ig[j] = ta[j] = ib[j] = T(0);
      if (!active[j]) continue;
      // column `it` of the Lanczos tridiagonal: beta_k above, alpha_k on,
      // beta_{k+1} below the diagonal; two old rotations, one new one
      const double bk = it > 1 ? beta[j] : 0.0, ak = (double)alpha[j], bn = (double)bnext[j];
      const double eps_k = s2[j] * bk, dbar = c2[j] * bk;
      const double delta = c1[j] * dbar + s1[j] * ak;
      const double gbar = -s1[j] * dbar + c1[j] * ak;
      const double gamma = std::hypot(gbar, bn);
      if (!(gamma > 0.0) || !std::isfinite(gamma)) {  // singular or broken down: stop this system
        active[j] = 0;
        continue;
      }
      const double c = gbar / gamma, sn = bn / gamma;
      const double tau = c * phibar[j];
      phibar[j] = -sn * phibar[j];
      dl[j] = (T)delta;
      ep[j] = (T)eps_k;
      ig[j] = (T)(1.0 / gamma);
      ta[j] = (T)tau;
      ib[j] = bn > 0.0 ? (T)(1.0 / bn) : T(0);
      c2[j] = c1[j];
      s2[j] = s1[j];
      c1[j] = c;
      s1[j] = sn;
      beta[j] = bn;
      out.iters[j] = it;
      out.relres[j] = std::fabs(phibar[j]) / (double)beta1[j];
      if (out.relres[j] <= tol[j] || bn == 0.0) active[j] = 0;
    }
    krylov_minres_update<T>(n, k, ld, v->dev(), vold->dev(), w->dev(), d1->dev(), d2->dev(), xv->dev(), dl.data(),
                            ep.data(), ig.data(), ta.data(), ib.data());
    std::swap(v, vold);  // the next Lanczos vector was written over the old one
    std::swap(d1, d2);   // the new direction was written over d_{k-2}
  }
  for (int j = 0; j < k; j++)
    if (out.relres[j] > tol[j]) out.unconverged++;
  blas_copy<T>(xv.get(), X);
  return out;
}

// correction equations in the precision TI: t_j solves
// (I - QQ')(A + s_j I) t_j = rhs_j, converted from/to T
template <typename T, typename TI>
MinresOut<TI> minres_lowered(const Operator<TI>& op, const std::vector<T>& s, const Mvec<T>* Q, const Mvec<T>* rhs,
                             Mvec<T>* t, const std::vector<double>& tol, int maxIter) {
  const Map* map = rhs->map;
  const int k = rhs->nvec;
  MvBox<TI> b(map, k), x(map, k);
  MvBox<TI> q;
  if (Q && Q->nvec > 0) {
    q.p = api::mvec_create<TI>(map, Q->nvec);
    convert_block<T, TI>(Q->dev(), Q->lda, q->dev(), q->lda, Q->nrows, Q->nvec);
  }
  convert_block<T, TI>(rhs->dev(), rhs->lda, b->dev(), b->lda, rhs->nrows, k);
  std::vector<TI> si(s.begin(), s.end());
  MinresOut<TI> r = minres_block<TI>(op, si, q.get(), b.get(), x.get(), tol, maxIter);
  convert_block<TI, T>(x->dev(), x->lda, t->dev(), t->lda, t->nrows, k);
  return r;
}

// ---------------------------------------------------------------------------
// symmetric block Jacobi-Davidson (standard Ritz values, B = I)
// ---------------------------------------------------------------------------
// V(:, c0:c0+ncols) <- V(:, c0:c1) * Q_H(:, 0:ncols), same for AV, and
// H <- Q_H' H Q_H on that block (transform_searchSpace,
// src/core/phist_transform_searchspace_def.hpp:11-45); qh is the m x m
// column-major Ritz basis of the block, ncols = mcols - c0
template <typename T>
void transform(Mvec<T>* V_, Mvec<T>* AV_, SdMat<T>* H_, int c0, int c1, int mcols, const std::vector<double>& qh,
               int m) {
  const int nc = mcols - c0;
  if (nc <= 0) return;
  SdMat<T>* M = api::sdMat_create<T>(m, nc);
  for (int j = 0; j < nc; j++)
    for (int i = 0; i < m; i++) M->h(i, j) = (T)qh[(size_t)j * m + i];
  sd_to_device(M);
  {
    Mvec<T>* V = nullptr;
    api::mvec_view_block(V_, &V, c0, c1 - 1);
    gemm_times_sd_inplace<T>(V, M);
    api::mvec_delete(V);
    Mvec<T>* AV = nullptr;
    api::mvec_view_block(AV_, &AV, c0, c1 - 1);
    gemm_times_sd_inplace<T>(AV, M);
    api::mvec_delete(AV);
  }
  // H block <- M' H M (dense, host)
  std::vector<double> HM((size_t)m * nc, 0.0), out((size_t)nc * nc, 0.0);
  for (int j = 0; j < nc; j++)
    for (int i = 0; i < m; i++) {
      double sacc = 0.0;
      for (int l = 0; l < m; l++) sacc += (double)H_->h(c0 + i, c0 + l) * qh[(size_t)j * m + l];
      HM[(size_t)j * m + i] = sacc;
    }
  for (int j = 0; j < nc; j++)
    for (int i = 0; i < nc; i++) {
      double sacc = 0.0;
      for (int l = 0; l < m; l++) sacc += qh[(size_t)i * m + l] * HM[(size_t)j * m + l];
      out[(size_t)j * nc + i] = sacc;
    }
  for (int j = 0; j < nc; j++)
    for (int i = 0; i < nc; i++) H_->h(c0 + i, c0 + j) = (T)(0.5 * (out[(size_t)j * nc + i] + out[(size_t)i * nc + j]));
  // couplings of the locked vectors to the transformed block: H(0:c0, c0:) M
  // (keeps H = V'AV exact, so the Schur couplings R(0:c0, :) stay current)
  std::vector<double> row(nc);
  for (int i = 0; i < c0; i++) {
    for (int j = 0; j < nc; j++) {
      double sacc = 0.0;
      for (int l = 0; l < m; l++) sacc += (double)H_->h(i, c0 + l) * qh[(size_t)j * m + l];
      row[j] = sacc;
    }
    for (int j = 0; j < nc; j++) H_->h(i, c0 + j) = H_->h(c0 + j, i) = (T)row[j];
  }
  sd_to_device(H_);
  api::sdMat_delete(M);
}

template <typename T>
std::vector<int> ritz_order(const std::vector<double>& w, phist_EeigSort which) {
  std::vector<int> o(w.size());
  for (size_t i = 0; i < o.size(); i++) o[i] = (int)i;
  auto key = [&](int i) {
    switch (which) {
      case phist_LM: return -std::fabs(w[i]);
      case phist_SM: return std::fabs(w[i]);
      case phist_LR: return -w[i];
      default: return w[i];  // SR
    }
  };
  std::stable_sort(o.begin(), o.end(), [&](int a, int b) { return key(a) < key(b); });
  return o;
}

template <typename T>
void subspacejada(const Operator<T>& A, const Operator<float>* A32, const phist_jadaOpts& opts, Mvec<T>* Qout,
                  SdMat<T>* Rout, T* ev, T* resNormOut, int* nConvOut, int* nIterOut) {
  using MT = double;
  const MT tol = opts.convTol;
  const int nEig = opts.numEigs, bs = opts.blockSize;
  const int lookAhead = opts.lookAhead > 0 ? opts.lookAhead : 2 * bs;
  const int maxIter = opts.maxIters, minBase = opts.minBas, maxBase = opts.maxBas;
  const int innerMaxIters = opts.innerSolvMaxIters;
  const MT innerTolBase = opts.innerSolvBaseTol;
  const bool useRelTol = opts.relConvTol != 0;
  const MT initialShift = opts.initialShift_r;
  const phist_EeigSort which = opts.which;
  const MT eps = std::numeric_limits<T>::epsilon();
  const int nEigMax = nEig + bs - 1;

  PHX_REQUIRE(bs >= 1, PHIST_INVALID_INPUT, "blockSize %d", bs);
  PHX_REQUIRE(minBase >= nEigMax && minBase + bs <= maxBase && maxBase >= nEig + bs, PHIST_INVALID_INPUT,
              "basis sizes minBas=%d maxBas=%d do not fit numEigs=%d blockSize=%d", minBase, maxBase, nEig, bs);
  PHX_REQUIRE(Qout->nvec >= nEigMax && Rout->nrows == Qout->nvec && Rout->ncols == Qout->nvec, PHIST_INVALID_INPUT,
              "Q needs >= %d columns and R matching square", nEigMax);
  PHX_REQUIRE(maxBase <= 256, PHIST_INVALID_INPUT, "maxBas %d too large", maxBase);
  for (int i = 0; i < nEig; i++) resNormOut[i] = T(-1);

  const Map* map = Qout->map;
  MvBox<T> V_(map, maxBase), AV_(map, maxBase), Q_(map, nEigMax), t_(map, bs);
  const int resDim = std::min(lookAhead, nEigMax);
  MvBox<T> res_(map, resDim);
  SdBox<T> H_(maxBase, maxBase);
  // host mirror of the projected matrix H = V'AV (column-major maxBase^2)
  auto H = [&](int i, int j) -> T& { return H_->h(i, j); };
  std::vector<double> theta(maxBase, 0.0), resNorm(nEigMax, -1.0), resScaling(nEigMax, 1.0);
  std::vector<double> innerTol(nEigMax, innerTolBase), lastOuterRes(nEigMax, 0.0);
  api::sdMat_put_value<T>(H_.get(), T(0));
  api::sdMat_put_value<T>(Rout, T(0));

  // ---- start-up basis ----
  int nV = 0;
  int nv0 = 0;
  const Mvec<T>* v0in = reinterpret_cast<const Mvec<T>*>(opts.v0);
  if (v0in) nv0 = v0in->nvec;
  if (nv0 > 1 && nv0 < minBase) nv0 = 1;
  if (nv0 > maxBase) nv0 = maxBase;
  if (nv0 >= minBase) {
    nV = nv0;
    View<T> V(V_.get(), 0, nV - 1), AV(AV_.get(), 0, nV - 1);
    View<T> src(v0in, 0, nV - 1);
    blas_copy<T>(src.get(), V.get());
    A.apply_shifted(T(1), nullptr, V.get(), T(0), AV.get(), nullptr);
    SdView<T> Hv(H_.get(), 0, nV - 1, 0, nV - 1);
    gemm_gram<T>(T(1), V.get(), AV.get(), T(0), Hv.get());
  } else {
    // Arnoldi: A V(:,1:m) = V(:,1:m+1) H(1:m+1,1:m), AV kept explicitly
    nV = minBase;
    {
      View<T> v0(V_.get(), 0, 0);
      if (v0in) {
        View<T> src(v0in, 0, 0);
        blas_copy<T>(src.get(), v0.get());
      } else {
        api::mvec_random<T>(v0.get());
      }
      T nrm;
      blas_dot<T>(v0.get(), v0.get(), &nrm);
      nrm = T(1) / std::sqrt(nrm);
      blas_scale_cols<T>(v0.get(), &nrm, true);
    }
    MvBox<T> wcol(map, 1);
    for (int j = 0; j < nV; j++) {
      View<T> vj(V_.get(), j, j), avj(AV_.get(), j, j);
      A.apply_shifted(T(1), nullptr, vj.get(), T(0), avj.get(), nullptr);
      if (j + 1 >= maxBase) break;
      View<T> Vj(V_.get(), 0, j), vn(V_.get(), j + 1, j + 1);
      blas_copy<T>(avj.get(), vn.get());
      SdBox<T> R1(1, 1), R2(j + 1, 1);
      int rank = 0;
      // an invariant subspace is continued with a random orthogonal vector
      // (phist_simple_arnoldi_def.hpp:118-141)
      const int dim0 = api::orthog<T>(Vj.get(), vn.get(), R1.get(), R2.get(), 3, &rank, 131072);
      for (int i = 0; i <= j; i++) H(i, j) = R2->h(i, 0);
      H(j + 1, j) = R1->h(0, 0);
      if (dim0 == 0) {  // a more accurate subdiagonal entry: v_{j+1}' A v_j
        T h;
        blas_dot<T>(vn.get(), avj.get(), &h);
        H(j + 1, j) = h;
      }
    }
    // H = V'AV over the first nV columns (the square part of the Arnoldi
    // Hessenberg); symmetrise exactly
    for (int j = 0; j < nV; j++)
      for (int i = j + 1; i < nV; i++) H(j, i) = H(i, j);
    sd_to_device(H_.get());
  }

  // ---- main loop ----
  int nConv = 0, iter = 0;
  std::vector<double> hb, qh;  // Ritz problem of the active block
  std::vector<int> selected(bs);
  std::vector<T> sigma(bs);
  for (iter = 0; iter < maxIter; iter++) {
    const int nEig_ = std::min(nConv + lookAhead, nEigMax);
    const int m = nV - nConv;
    // Ritz pairs of H(nConv:nV, nConv:nV), sorted by `which`
    hb.assign((size_t)m * m, 0.0);
    for (int j = 0; j < m; j++)
      for (int i = 0; i < m; i++) hb[(size_t)j * m + i] = (double)H(nConv + i, nConv + j);
    std::vector<double> w(m);
    dense::sym_eig(hb.data(), m, m, w.data());
    const std::vector<int> ord = ritz_order<T>(w, which);
    qh.assign((size_t)m * m, 0.0);
    for (int j = 0; j < m; j++) {
      theta[nConv + j] = w[ord[j]];
      for (int i = 0; i < m; i++) qh[(size_t)j * m + i] = hb[(size_t)ord[j] * m + i];
    }
    const int nq = nEig_ - nConv;  // Ritz vectors formed
    SdBox<T> QH(m, nq);
    for (int j = 0; j < nq; j++)
      for (int i = 0; i < m; i++) QH->h(i, j) = (T)qh[(size_t)j * m + i];
    sd_to_device(QH.get());
    // R(0:nEig_, nConv:nEig_): Ritz values on the diagonal, couplings to the
    // locked vectors H(0:nConv, nConv:nV)*QH above it
    for (int j = nConv; j < nEig_; j++) {
      for (int i = 0; i < nEig_; i++) Rout->h(i, j) = T(0);
      Rout->h(j, j) = (T)theta[j];
      for (int i = 0; i < nConv; i++) {
        double sacc = 0.0;
        for (int l = 0; l < m; l++) sacc += (double)H(i, nConv + l) * qh[(size_t)(j - nConv) * m + l];
        Rout->h(i, j) = (T)sacc;
      }
    }
    sd_to_device(Rout);
    {
      View<T> V(V_.get(), nConv, nV - 1), AV(AV_.get(), nConv, nV - 1);
      View<T> Qq(Q_.get(), nConv, nEig_ - 1), Q(Q_.get(), 0, nEig_ - 1);
      View<T> tres(res_.get(), 0, nq - 1);
      gemm_times_sd<T>(T(1), V.get(), QH.get(), T(0), Qq.get());
      gemm_times_sd<T>(T(1), AV.get(), QH.get(), T(0), tres.get());
      // res = Q R - A Q (the negative residual)
      SdView<T> Rr(Rout, 0, nEig_ - 1, nConv, nEig_ - 1);
      gemm_times_sd<T>(T(1), Q.get(), Rr.get(), T(-1), tres.get());
      std::vector<T> nr(nq);
      blas_dot<T>(tres.get(), tres.get(), nr.data());
      for (int j = 0; j < nq; j++) resNorm[nConv + j] = std::sqrt((double)nr[j]);
    }
    for (int j = nConv; j < nEig_; j++)
      resScaling[j] = (useRelTol && std::fabs(theta[j]) > std::pow(eps, 2.0 / 3.0)) ? std::fabs(theta[j]) : 1.0;

    int nNewConv = 0;
    for (int i = nConv; i < std::min(nEig, nEig_); i++) {
      if (resNorm[i] <= tol * resScaling[i] && i == nConv + nNewConv)
        nNewConv++;
      else if (bs == 1)
        break;
    }
    if (nConv + nNewConv >= nEig) {
      nConv += nNewConv;
      break;
    }
    // lock the newly converged Ritz vectors (they become the leading
    // columns of the transformed basis)
    if (nNewConv > 0) {
      const int mcols = (nV + 2 * bs > maxBase) ? minBase : nV;
      transform(V_.get(), AV_.get(), H_.get(), nConv, nV, mcols, qh, m);
      nConv += nNewConv;
      if (nV + 2 * bs > maxBase) nV = minBase;
    }
    // choose up to bs unconverged Ritz pairs (not splitting clusters)
    int k = 0, k_ = 0;
    for (int i = 0; i < nEig_ && k < bs; i++) {
      if (i >= nEig && std::fabs(theta[i] - theta[i - 1]) > std::sqrt(tol)) break;
      if (resNorm[i] > tol) {
        k_ = i;
        sigma[k] = (T)(iter < opts.initialShiftIters ? initialShift : theta[i]);
        selected[k] = i;
        k++;
      }
    }
    while (k_ + 1 < nEig_ && std::fabs(theta[k_ + 1] - theta[k_]) < 10 * std::fabs(theta[k_ + 1]) * std::sqrt(tol))
      k_++;
    if (nV + k > maxBase) {
      if (nNewConv == 0) transform(V_.get(), AV_.get(), H_.get(), nConv, nV, minBase, qh, m);
      nV = minBase;
    }
    if (k == 0) continue;
    // inner tolerances (phist_subspacejada_def.hpp:832-853)
    for (int i = 0; i < k; i++) {
      const int idx = nConv + i;
      if (resNorm[idx] > 4 * lastOuterRes[idx]) innerTol[idx] = innerTolBase;
      if (innerTol[idx] > eps) innerTol[idx] = std::max(eps, innerTol[idx] * innerTolBase);
      innerTol[idx] = std::max(innerTol[idx], 0.1 * tol / (eps + resNorm[idx]));
      innerTol[idx] = std::min(innerTol[idx], 0.1);
      lastOuterRes[idx] = resNorm[idx];
    }
    // correction equations (I - QQ')(A - sigma I) t = QR - AQ (= -r)
    {
      const int oldConv = nConv - nNewConv;
      MvBox<T> rhs(map, k);
      for (int i = 0; i < k; i++) {
        View<T> src(res_.get(), selected[i] - oldConv, selected[i] - oldConv), dst(rhs.get(), i, i);
        blas_copy<T>(src.get(), dst.get());
      }
      const int k0 = opts.innerSolvMaxProjectionSpace >= 0 ? std::max(0, k_ - opts.innerSolvMaxProjectionSpace) : 0;
      View<T> Qtil(Q_.get(), k0, k_);
      View<T> t(t_.get(), 0, k - 1);
      std::vector<T> s(k);
      std::vector<double> tl(k);
      for (int i = 0; i < k; i++) {
        s[i] = -sigma[i];
        tl[i] = innerTol[nConv + i];
      }
      std::vector<int> inner_it(k, 0);
      std::vector<double> inner_rr(k, 0.0);
      if (opts.innerSolvType == phist_NO_LINSOLV) {
        blas_copy<T>(rhs.get(), t.get());
      } else if (A32) {
        for (double& x : tl) x = std::max(x, 1.0e-6);  // FP32 solves stagnate below this
        MinresOut<float> r = minres_lowered<T, float>(*A32, s, Qtil.get(), rhs.get(), t.get(), tl, innerMaxIters);
        inner_it = r.iters;
        inner_rr = r.relres;
      } else {
        MinresOut<T> r = minres_block<T>(A, s, Qtil.get(), rhs.get(), t.get(), tl, innerMaxIters);
        inner_it = r.iters;
        inner_rr = r.relres;
      }
      if (jada_verbose()) {
        fprintf(stderr, "[jada] it %d nConv %d nV %d k %d Qtil %d..%d:", iter, nConv, nV, k, k0, k_);
        for (int i = 0; i < k; i++)
          fprintf(stderr, " [%d th %.10g res %.3e sig %.6g tol %.2e mr %d %.2e]", selected[i], theta[selected[i]],
                  resNorm[selected[i]], (double)sigma[i], tl[i], inner_it[i], inner_rr[i]);
        fprintf(stderr, "\n");
      }
      // extend the basis: V(:, nV:nV+k) = orthonormalised t
      View<T> Vful(V_.get(), 0, nV - 1), Vv(V_.get(), nV, nV + k - 1), AVv(AV_.get(), nV, nV + k - 1);
      blas_copy<T>(t.get(), Vv.get());
      SdBox<T> R1(k, k), R2(nV, k);
      int rank = 0;
      // no rank truncation here: near convergence t is tiny, and a rank
      // test would replace it by noise (the reference passes no flags)
      api::orthog<T>(Vful.get(), Vv.get(), R1.get(), R2.get(), 3, &rank, 0);
      A.apply_shifted(T(1), nullptr, Vv.get(), T(0), AVv.get(), nullptr);
      View<T> Vall(V_.get(), 0, nV + k - 1);
      SdBox<T> Hn(nV + k, k);
      gemm_gram<T>(T(1), Vall.get(), AVv.get(), T(0), Hn.get());
      for (int j = 0; j < k; j++)
        for (int i = 0; i < nV + k; i++) {
          H(i, nV + j) = Hn->h(i, j);
          if (i < nV) H(nV + j, i) = Hn->h(i, j);  // symmetric
        }
      for (int j = 0; j < k; j++)  // symmetrise the new diagonal block
        for (int i = j + 1; i < k; i++) {
          const T a = (T)0.5 * (H(nV + i, nV + j) + H(nV + j, nV + i));
          H(nV + i, nV + j) = H(nV + j, nV + i) = a;
        }
      sd_to_device(H_.get());
      nV += k;
    }
  }
  *nIterOut = iter;
  *nConvOut = nConv;
  // Schur vectors (or the current basis when more columns were asked for)
  const int nQout = Qout->nvec;
  if (nQout <= nEigMax) {
    View<T> src(Q_.get(), 0, nQout - 1);
    blas_copy<T>(src.get(), Qout);
  } else {
    const int nb = std::min(nQout, std::min(minBase, nV));
    View<T> src(V_.get(), 0, nb - 1), dst(Qout, 0, nb - 1);
    blas_copy<T>(src.get(), dst.get());
  }
  for (int i = 0; i < nEigMax; i++) ev[i] = (T)theta[i];
  for (int i = 0; i < nEig; i++) resNormOut[i] = (T)resNorm[i];
}

}  // namespace

namespace {

template <typename T>
void wrap_sparse(typename LinOpT<T>::type* op, const void* A) {
  std::memset(op, 0, sizeof(*op));
  const SparseMat<T>* M = reinterpret_cast<const SparseMat<T>*>(A);
  op->A = A;
  op->range_map = M->range_map;
  op->domain_map = M->domain_map;
  op->apply = &LinOpT<T>::apply;
  op->applyT = &LinOpT<T>::applyT;
  op->apply_shifted = &LinOpT<T>::apply_shifted;
  op->fused_apply_mvTmv = &LinOpT<T>::fused;
  op->destroy = &LinOpT<T>::destroy;
}

template <typename T>
void run_jada(const typename LinOpT<T>::type* AB_op, const void* B_op, const phist_jadaOpts& opts, void* Q, void* R,
              T* ev, T* resNorm, int* nConv, int* nIter) {
  PHX_REQUIRE(AB_op != nullptr, PHIST_INVALID_INPUT, "no operator");
  PHX_REQUIRE(B_op == nullptr, PHIST_NOT_IMPLEMENTED, "generalized problems (B_op) are not supported");
  PHX_REQUIRE(opts.how == phist_STANDARD, PHIST_NOT_IMPLEMENTED, "only standard Ritz extraction");
  PHX_REQUIRE(opts.symmetry == phist_HERMITIAN || opts.symmetry == phist_COMPLEX_SYMMETRIC, PHIST_NOT_IMPLEMENTED,
              "only symmetric problems (opts.symmetry HERMITIAN)");
  PHX_REQUIRE(opts.which == phist_LM || opts.which == phist_SM || opts.which == phist_LR || opts.which == phist_SR,
              PHIST_NOT_IMPLEMENTED, "eigenvalue selection %d not supported", (int)opts.which);
  PHX_REQUIRE(opts.innerSolvType == phist_MINRES || opts.innerSolvType == phist_GMRES ||
                  opts.innerSolvType == phist_NO_LINSOLV,
              PHIST_NOT_IMPLEMENTED, "inner solver %d not supported", (int)opts.innerSolvType);
  PHX_REQUIRE(opts.preconOp == nullptr, PHIST_NOT_IMPLEMENTED, "preconditioned correction equations");
  Operator<T> op;
  op.A = own_matrix<T>(AB_op);
  if (!op.A) op.user = AB_op;
  Operator<float> op32;
  const Operator<float>* lowered = nullptr;
  if constexpr (std::is_same<T, double>::value) {
    if (inner_fp32() && op.A) {
      op32.A = sparse_fp32_twin(op.A);
      lowered = &op32;
    }
  }
  subspacejada<T>(op, lowered, opts, reinterpret_cast<Mvec<T>*>(Q), reinterpret_cast<SdMat<T>*>(R), ev, resNorm,
                  nConv, nIter);
}

}  // namespace
}  // namespace phx

using namespace phx;

extern "C" {

void phist_jadaOpts_setDefaults(phist_jadaOpts* opts) {
  std::memset(opts, 0, sizeof(*opts));
  opts->symmetry = phist_GENERAL;
  opts->numEigs = 6;
  opts->which = phist_LM;
  opts->how = phist_STANDARD;
  opts->maxIters = 300;
  opts->blockSize = 1;
  opts->lookAhead = -1;
  opts->minBas = 10;
  opts->maxBas = 20;
  opts->convTol = 1.0e-12;
  opts->relConvTol = 0;
  opts->v0 = nullptr;
  opts->arno = 1;
  opts->initialShift_r = 0.0;
  opts->initialShift_i = 0.0;
  opts->initialShiftIters = 0;
  opts->innerSolvType = phist_GMRES;
  opts->innerSolvMaxBas = -1;
  opts->innerSolvMaxIters = 20;
  opts->innerSolvBlockSize = -1;
  opts->innerSolvStopAfterFirstConverged = 0;
  opts->innerSolvRobust = 1;
  opts->innerSolvMaxProjectionSpace = -1;
  opts->innerSolvBaseTol = 0.1;
  opts->preconOp = nullptr;
  opts->preconType = phist_NO_PRECON;
  std::strcpy(opts->preconOpts, "none");
  opts->preconSkewProject = 1;
  opts->preconUpdate = 0;
  opts->customSolver = nullptr;
  opts->customSolver_run1 = nullptr;
  opts->customSolver_run = nullptr;
}

void phist_DlinearOp_wrap_sparseMat(phist_DlinearOp_ptr op, phist_Dconst_sparseMat_ptr A, int* iflag) {
  *iflag = 0;
  guard(iflag, "phist_DlinearOp_wrap_sparseMat", [&] {
    PHX_REQUIRE(op && A, PHIST_INVALID_INPUT, "null argument");
    wrap_sparse<double>(op, A);
  });
}
void phist_SlinearOp_wrap_sparseMat(phist_SlinearOp_ptr op, phist_Sconst_sparseMat_ptr A, int* iflag) {
  *iflag = 0;
  guard(iflag, "phist_SlinearOp_wrap_sparseMat", [&] {
    PHX_REQUIRE(op && A, PHIST_INVALID_INPUT, "null argument");
    wrap_sparse<float>(op, A);
  });
}

void phist_Dsubspacejada(phist_Dconst_linearOp_ptr AB_op, phist_Dconst_linearOp_ptr B_op, phist_jadaOpts opts,
                         phist_Dmvec_ptr Q, phist_DsdMat_ptr R, double* ev, double* resNorm, int* nConv, int* nIter,
                         int* iflag) {
  *iflag = 0;
  guard(iflag, "phist_Dsubspacejada", [&] { run_jada<double>(AB_op, B_op, opts, Q, R, ev, resNorm, nConv, nIter); });
}
void phist_Ssubspacejada(phist_Sconst_linearOp_ptr AB_op, phist_Sconst_linearOp_ptr B_op, phist_jadaOpts opts,
                         phist_Smvec_ptr Q, phist_SsdMat_ptr R, float* ev, float* resNorm, int* nConv, int* nIter,
                         int* iflag) {
  *iflag = 0;
  guard(iflag, "phist_Ssubspacejada", [&] { run_jada<float>(AB_op, B_op, opts, Q, R, ev, resNorm, nConv, nIter); });
}

void phist_amd_jada_set_inner_fp32(int enable, int* iflag) {
  *iflag = 0;
  inner_fp32();
  g_inner_fp32 = enable != 0;
}

void phist_amd_DblockedMINRES_shifted(phist_Dconst_sparseMat_ptr A, int nsys, const double shifts[],
                                      phist_Dconst_mvec_ptr B, phist_Dmvec_ptr X, const double tol[], int maxIter,
                                      int fp32, int nIter[], double relres[], int* iflag) {
  *iflag = 0;
  guard(iflag, "phist_amd_DblockedMINRES_shifted", [&] {
    const auto* M = reinterpret_cast<const SparseMat<double>*>(A);
    const auto* b = reinterpret_cast<const Mvec<double>*>(B);
    auto* x = reinterpret_cast<Mvec<double>*>(X);
    PHX_REQUIRE(M && b && x, PHIST_INVALID_INPUT, "null argument");
    PHX_REQUIRE(nsys >= 1 && nsys <= 64 && b->nvec == nsys && x->nvec == nsys, PHIST_INVALID_INPUT,
                "nsys %d must match B and X (at most 64)", nsys);
    PHX_REQUIRE(M->nrows == M->ncols && b->nrows == M->nrows && x->nrows == M->nrows, PHIST_INVALID_INPUT,
                "shapes do not match the square matrix");
    std::vector<double> s(shifts, shifts + nsys), tl(tol, tol + nsys);
    int unconv = 0;
    if (fp32) {
      Operator<float> op;
      op.A = sparse_fp32_twin(M);
      MinresOut<float> r = minres_lowered<double, float>(op, s, nullptr, b, x, tl, maxIter);
      for (int j = 0; j < nsys; j++) {
        nIter[j] = r.iters[j];
        relres[j] = r.relres[j];
      }
      unconv = r.unconverged;
    } else {
      Operator<double> op;
      op.A = M;
      MinresOut<double> r = minres_block<double>(op, s, nullptr, b, x, tl, maxIter);
      for (int j = 0; j < nsys; j++) {
        nIter[j] = r.iters[j];
        relres[j] = r.relres[j];
      }
      unconv = r.unconverged;
    }
    *iflag = unconv;
  });
}

}  // extern "C"
