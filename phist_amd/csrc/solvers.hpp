// solvers.hpp -- building blocks of the device solvers (jada.cpp, serve.cpp):
// scoped block vectors and views, the operator abstraction, and the slot-based
// blocked MINRES engine.
#pragma once
#include "objects.hpp"
#include "phist_amd_jada.h"
#include <algorithm>
#include <cmath>
#include <memory>
#include <type_traits>
#include <vector>

namespace phx {

template <typename T>
void krylov_lanczos(int64_t nrows, int ncols, int64_t ld, const T* v, const T* vold, T* w, const T* a, const T* b,
                    T* nrm_host);
template <typename T>
void krylov_minres_update(int64_t nrows, int ncols, int64_t ld, const T* v, T* vold, const T* w, const T* d1, T* d2,
                          T* x, const T* delta, const T* eps, const T* inv_gamma, const T* tau, const T* inv_beta);
template <typename TS, typename TD>
void convert_block(const TS* x, int64_t ldx, TD* y, int64_t ldy, int64_t nrows, int ncols);

// Classical Gram-Schmidt sweeps of the blocked GMRES over kk bases laid side
// by side (column l*kk + a of v: vector l of system a), one read of the basis
// each; mode 0: out(l,a) = <v(:,l*kk+a), w(:,a)>; 1: w(:,a) -= sum_l
// v(:,l*kk+a) c(l,a), then out as in mode 0 for the new w; 2: the update,
// then out(a) = ||w(:,a)||^2; 3: the update only.  c(l,a) = coef[l*kk+a]
// (device) where act[a] != 0 (act == nullptr: every system).  work holds
// gmres_cgs_grid(nrows, kk) * nout partial sums plus nout; out (device) gets
// the nout = nl*kk (mode 2: kk) sums.
int gmres_cgs_grid(int64_t nrows, int kk);
template <typename T>
void gmres_cgs_dev(int mode, int64_t nrows, const T* v, int64_t ldv, T* w, int64_t ldw, int kk, int nl,
                   const T* coef, const T* act, T* work, int64_t work_elems, T* out);

// device Givens step and basis normalisation of the blocked GMRES (krylov.hip; GmresDevState in amd_common.hpp)
template <typename T>
void gmres_givens_dev(const GmresDevState& st, int kk, int m, int i, const T* outs, T* act, T* inv, double eps1000,
                      int max_iter);
template <typename T>
void gmres_normalize_dev(int64_t nrows, int kk, const T* w, int64_t ldw, const T* inv, const T* nrm2, T* vcur,
                         int64_t ldc, T* vb, int64_t ldv);

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
  SdBox() = default;
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
  // Y = (A + diag(s)) X with the shifts s on the device and the column dots
  // <X, Y> left on the device (dots_dev[nvec + c]; dots_dev[c] = <Y, Y>);
  // sparse matrices of this library only
  // (xscale: optional per-column factors of X on the device, stride xstride)
  void apply_shifted_dev(const T* s_dev, const Mvec<T>* X, Mvec<T>* Y, T* dots_dev, const double* xscale = nullptr,
                         int64_t xstride = 1, int ncols_active = -1) const {
    PHX_REQUIRE(A != nullptr, PHIST_NOT_IMPLEMENTED, "device dots need a sparse matrix of this library");
    SpmvRequest req;
    req.has_shift = true;
    req.shifts_dev = s_dev;
    req.dot_vw = true;
    req.dots_dev = dots_dev;
    req.xscale = xscale;
    req.xscale_stride = xstride;
    req.ncols_active = ncols_active;
    spmv_launch<T>(A, T(1), nullptr, X, T(0), Y, req, nullptr, nullptr);
  }
};


// copy a block into another of any precision (same shape)
template <typename TS, typename TD>
void copy_any(const Mvec<TS>* X, Mvec<TD>* Y) {
  if constexpr (std::is_same<TS, TD>::value)
    blas_copy<TD>(X, Y);
  else
    convert_block<TS, TD>(X->dev(), X->lda, Y->dev(), Y->lda, X->nrows, X->nvec);
}

// ---------------------------------------------------------------------------
// MinresBatch: up to `cap` independent systems P(A + s_j I)P x_j = b_j
// (P = I - QQ', or I without Q), one column (slot) each, advanced together by
// MINRES with Paige & Saunders' short recurrences.  Per iteration: one
// shifted SpMV whose fused <v, Av> dots give the Lanczos alphas, the
// projection, k_lanczos (update + norms) and k_minres (direction, solution,
// next Lanczos vector).  A slot whose system has converged (|phibar| <=
// tol*|b|), broken down or reached its iteration limit stops updating; its
// solution can be stored and the slot loaded with the next system while the
// others keep iterating (the slot swap of
// phist_jadaCorrectionSolver_def.hpp:428-466).
//
// For a sparse matrix of this library the scalars live on the device
// (krylov_minres_step_dev): an iteration is a chain of launches with no host
// synchronisation, and the host reads the slot states one iteration late
// (a copy into pinned memory per iteration, waited for after the next
// iteration has been queued).  The accessors (active, iters, relres) report
// that lagged state; a loop `while (any_active()) iterate();` therefore runs
// one iteration more than needed, which changes nothing for frozen slots.
// For a user operator the scalars stay on the host, as the dots come back
// to the host anyway.
// ---------------------------------------------------------------------------
template <typename T>
void krylov_minres_step_dev(int64_t nrows, int k, int64_t ld, const T* v, T* vold, T* w, const T* d1, T* d2, T* x,
                            const T* dots, MinresState* st, T* work, const T* Q = nullptr, int64_t ldq = 0,
                            int nq = 0, const T* C = nullptr, int ldc = 0, int phase = -1, int nact = -1);

// device column states of freshly loaded systems (see krylov.hip)
template <typename T>
void krylov_minres_init_dev(MinresState* st, const T* nrm2, int j0, int n, int ph);

constexpr int kFusedProjCols = 64;  // projection bases up to this width are applied inside k_lanczos

template <typename T>
class MinresBatch {
 public:
  MinresBatch(const Operator<T>& op, const Map* map, int cap, const Mvec<T>* Q = nullptr)
      : op_(op), map_(map), cap_(cap), Q_((Q && Q->nvec > 0) ? Q : nullptr), st_(cap), load_seq_(cap, 0),
        view_seq_(cap, 0) {
    PHX_REQUIRE(cap >= 1 && cap <= 64, PHIST_INVALID_INPUT, "MINRES batch of %d systems", cap);
    // six separate blocks of `cap` columns (row stride cap each): the SpMV
    // gathers rows of V, and a V whose rows are interleaved with the other
    // blocks would waste most of every cache line it touches
    for (int b = 0; b < 6; b++) {
      blk_[b].p = api::mvec_create<T>(map, cap);
      blas_put_value<T>(blk_[b].p, T(0));
    }
    if (Q_ && Q_->nvec > 0) C_.reset(new SdBox<T>(Q_->nvec, cap));
    dev_ = op_.A != nullptr && rt().minres_device;
    if (dev_) {
      d_st_ = static_cast<MinresState*>(dev_alloc(sizeof(MinresState) * cap));
      d_dots_ = static_cast<T*>(dev_alloc(sizeof(T) * 2 * cap));
      d_work_ = static_cast<T*>(dev_alloc(sizeof(T) * 8 * cap));
      d_shift_ = static_cast<T*>(dev_alloc(sizeof(T) * cap));
      h_buf_ = static_cast<MinresState*>(pinned_alloc(sizeof(MinresState) * 3 * cap));
      h_shift_ = static_cast<T*>(pinned_alloc(sizeof(T) * cap));
      for (int e = 0; e < 3; e++) PHX_HIP(hipEventCreateWithFlags(&ev_[e], hipEventDisableTiming));
      upload(0, cap);  // all frozen
    }
  }
  ~MinresBatch() {
    if (!dev_) return;
    try {
      for (int e = 0; e < 3; e++)
        if (ev_[e]) PHX_HIP(hipEventDestroy(ev_[e]));
      pinned_release(h_buf_, sizeof(MinresState) * 3 * cap_);  // synchronises the stream first
      pinned_release(h_shift_, sizeof(T) * cap_);
      drop_graphs();
      if (garena_) dev_release(garena_, garena_bytes_);
      dev_release(d_shift_, sizeof(T) * cap_);
      dev_release(d_work_, sizeof(T) * 8 * cap_);
      dev_release(d_dots_, sizeof(T) * 2 * cap_);
      dev_release(d_st_, sizeof(MinresState) * cap_);
    } catch (...) {
    }
  }
  MinresBatch(const MinresBatch&) = delete;
  MinresBatch& operator=(const MinresBatch&) = delete;

  int capacity() const { return cap_; }
  // projector P = I - QQ' for the following systems (Q = null: none); Q must
  // stay alive while the batch iterates.  The coefficient block Q'w is kept
  // while Q keeps its width, so the captured iterations stay valid.
  void set_projection(const Mvec<T>* Q) {
    Q_ = (Q && Q->nvec > 0) ? Q : nullptr;
    if (Q_ && (!C_ || C_->get()->nrows != Q_->nvec)) C_.reset(new SdBox<T>(Q_->nvec, cap_));
  }
  int free_slot() const {
    for (int j = 0; j < cap_; j++)
      if (!st_[j].used) return j;
    return -1;
  }
  bool used(int j) const { return st_[j].used; }
  bool active(int j) const { return st_[j].m.active != 0; }
  int iters(int j) const { return st_[j].m.iters; }
  double relres(int j) const { return st_[j].m.relres; }
  bool converged(int j) const { return st_[j].m.relres <= st_[j].m.tol; }
  bool any_active() const {
    for (const Slot& s : st_)
      if (s.m.active) return true;
    return false;
  }
  int n_used() const {
    int n = 0;
    for (const Slot& s : st_) n += s.used;
    return n;
  }

  // all slots at once from the columns of B (x = 0)
  template <typename TS>
  void load_all(const Mvec<TS>* B, const std::vector<T>& shifts, const std::vector<double>& tol,
                const std::vector<int>& max_iter) {
    const int k = B->nvec;
    PHX_REQUIRE(k == cap_, PHIST_INVALID_INPUT, "load_all needs %d columns", cap_);
    drain();
    if (dev_ && B->nrows > 0) {
      // no host round trip: b stays unnormalised in v (factor 1/||b|| in the
      // column state), the norms go from the device dots into the states
      for (int r : {VOLD, D1, D2, XS}) blas_put_value<T>(cur(r), T(0));
      copy_any<TS, T>(B, cur(V));
      blas_dot_dev<T>(cur(V), cur(V), d_work_);
      for (int j = 0; j < k; j++) start(j, 1.0, shifts[j], tol[j], max_iter[j]);  // the norms follow on the device
      upload(0, k);
      krylov_minres_init_dev<T>(d_st_, d_work_, 0, k, (int)(phase_ % 3));
      return;
    }
    for (int b = 0; b < 6; b++) blas_put_value<T>(blk_[b].p, T(0));
    Mvec<T>* v = cur(V);
    copy_any<TS, T>(B, v);
    std::vector<T> nrm(k), sc(k);
    blas_dot<T>(v, v, nrm.data());
    for (int j = 0; j < k; j++) {
      nrm[j] = std::sqrt(nrm[j]);
      sc[j] = nrm[j] > T(0) ? T(1) / nrm[j] : T(0);
      start(j, (double)nrm[j], shifts[j], tol[j], max_iter[j]);
    }
    blas_scale_cols<T>(v, sc.data(), false);
    if (dev_) upload(0, k);
  }

  // slot j from column c of B (x = 0)
  template <typename TS>
  void load(int j, const Mvec<TS>* B, int c, T shift, double tol, int max_iter) {
    if (dev_ && B->nrows > 0) {  // as load_all: no host round trip
      for (int r : {VOLD, D1, D2, XS}) {
        View<T> col(cur(r), j, j);
        blas_put_value<T>(col.get(), T(0));
      }
      View<T> vj(cur(V), j, j);
      View<TS> bc(B, c, c);
      copy_any<TS, T>(bc.get(), vj.get());
      blas_dot_dev<T>(vj.get(), vj.get(), d_work_ + j);
      start(j, 1.0, shift, tol, max_iter);
      upload(j, 1);
      krylov_minres_init_dev<T>(d_st_, d_work_ + j, j, 1, (int)(phase_ % 3));
      return;
    }
    for (int b = 0; b < 6; b++) {
      View<T> col(blk_[b].p, j, j);
      blas_put_value<T>(col.get(), T(0));
    }
    View<T> vj(cur(V), j, j);
    View<TS> bc(B, c, c);
    copy_any<TS, T>(bc.get(), vj.get());
    T nrm;
    blas_dot<T>(vj.get(), vj.get(), &nrm);
    nrm = std::sqrt(nrm);
    T sc = nrm > T(0) ? T(1) / nrm : T(0);
    blas_scale_cols<T>(vj.get(), &sc, true);
    start(j, (double)nrm, shift, tol, max_iter);
    if (dev_) upload(j, 1);
  }

  // solution of slot j into column c of X (any precision)
  template <typename TD>
  void store(int j, Mvec<TD>* X, int c) const {
    View<T> xj(cur(XS), j, j);
    View<TD> xc(X, c, c);
    copy_any<T, TD>(xj.get(), xc.get());
  }
  template <typename TD>
  void store_all(Mvec<TD>* X) const {
    copy_any<T, TD>(cur(XS), X);
  }
  void release(int j) {
    st_[j] = Slot();
    if (dev_) upload(j, 1);  // frozen on the device as well
  }

  // one or more MINRES iterations of every active slot (device scalars: up to
  // the runtime's iterations per call, never more than any active system may
  // still take; when the queued ones suffice, only the host view catches up)
  void iterate() {
    if (dev_)
      iterate_dev();
    else
      iterate_host();
  }

  // bring the host's view of the slot states up to date with every queued iteration
  void drain() {
    if (dev_ && pending_ >= 0) {
      apply(pending_);
      pending_ = -1;
    }
  }

 private:
  enum { V = 0, VOLD = 1, W = 2, D1 = 3, D2 = 4, XS = 5 };
  struct Slot {
    bool used = false;
    T shift = T(0);
    MinresState m{0, 0, 1, 0, 1, 0, 0, 1, 0, 0, 0, 0, 0};
  };
  struct Blk {
    Mvec<T>* p = nullptr;
    ~Blk() {
      try {
        if (p) api::mvec_delete(p);
      } catch (...) {
      }
    }
  };
  Mvec<T>* cur(int r) const { return blk_[role_[r]].p; }
  MinresState* h_up() const { return h_buf_ + 2 * cap_; }
  void start(int j, double beta1, T shift, double tol, int max_iter) {
    Slot q;
    q.used = true;
    q.shift = shift;
    q.m.beta1 = beta1;
    q.m.phibar = beta1;
    q.m.tol = tol;
    q.m.max_iter = max_iter;
    q.m.active = beta1 > 0.0 && max_iter > 0;
    q.m.relres = beta1 > 0.0 ? 1.0 : 0.0;
    st_[j] = q;
  }
  // slot states j0..j0+n-1 to the device; copies of the device state queued
  // before this upload no longer speak for these slots
  void upload(int j0, int n) {
    if (uploads_) PHX_HIP(hipEventSynchronize(ev_[2]));  // the staging area is free again
    for (int j = j0; j < j0 + n; j++) {
      h_up()[j] = st_[j].m;
      // v is stored normalised after a load (factor 1), v_old is zero
      h_up()[j].sc[phase_ % 3] = 1.0;
      h_up()[j].sc[(phase_ + 1) % 3] = 0.0;
      h_up()[j].sc[(phase_ + 2) % 3] = 0.0;
      h_shift_[j] = st_[j].shift;
      load_seq_[j] = seq_ + 1;
      view_seq_[j] = seq_;
    }
    PHX_HIP(hipMemcpyAsync(d_st_ + j0, h_up() + j0, sizeof(MinresState) * n, hipMemcpyHostToDevice, rt().stream));
    PHX_HIP(hipMemcpyAsync(d_shift_ + j0, h_shift_ + j0, sizeof(T) * n, hipMemcpyHostToDevice, rt().stream));
    PHX_HIP(hipEventRecord(ev_[2], rt().stream));
    uploads_++;
  }
  void apply(int b) {
    PHX_HIP(hipEventSynchronize(ev_[b]));
    const MinresState* src = h_buf_ + (size_t)b * cap_;
    for (int j = 0; j < cap_; j++)
      if (copy_seq_[b] >= load_seq_[j]) {
        st_[j].m = src[j];
        view_seq_[j] = copy_seq_[b];
      }
  }

  // the kernels of one iteration on the library stream (recorded instead of
  // run while a graph is captured); the roles of the rotating blocks advance
  void enqueue_iteration() {
    // Without a projection (and outside graph capture) the vector passes skip the columns after the
    // last one the host view shows active (the same launches, so the other columns' arithmetic does not
    // change): the host view is a superset of the device's active set (a load shows active at once, a
    // deactivation after the next copy), and an inactive column is left as it is.  Slots fill from the
    // lowest free one, so a batch with few systems in flight moves few columns.
    const int k = cap_;
    int nact = -1;
    if (!Q_ && !rt().capturing) {
      nact = 0;
      for (int j = 0; j < cap_; j++)
        if (st_[j].m.active) nact = j + 1;
    }
    Mvec<T>* v = cur(V);
    Mvec<T>* w = cur(W);
    const int ph = (int)(phase_ % 3);
    op_.apply_shifted_dev(d_shift_, v, w, d_dots_, &d_st_[0].sc[ph], sizeof(MinresState) / sizeof(double), nact);
    const bool fuse = Q_ && Q_->nvec <= kFusedProjCols;  // w -= Q*C inside the Lanczos update
    if (Q_) {
      gemm_gram<T>(T(1), Q_, w, T(0), C_->get(), false);
      if (!fuse) gemm_times_sd<T>(T(-1), Q_, C_->get(), T(1), w);
    }
    krylov_minres_step_dev<T>(v->nrows, k, cap_, v->dev(), cur(VOLD)->dev(), w->dev(), cur(D1)->dev(),
                              cur(D2)->dev(), cur(XS)->dev(), d_dots_, d_st_, d_work_, fuse ? Q_->dev() : nullptr,
                              fuse ? Q_->lda : 0, fuse ? Q_->nvec : 0, fuse ? C_->get()->dev() : nullptr,
                              fuse ? C_->get()->root->ld : 0, ph, nact);
    advance_roles();
  }
  // after a device iteration: w (unnormalised) is the next v, v the next
  // v_old, the old v_old block takes the next SpMV; the direction blocks swap
  void advance_roles() {
    const int b = role_[VOLD];
    role_[VOLD] = role_[V];
    role_[V] = role_[W];
    role_[W] = b;
    std::swap(role_[D1], role_[D2]);
    phase_++;
  }

  void drop_graphs() {
    for (hipGraphExec_t& g : gexec_)
      if (g) {
        (void)hipGraphExecDestroy(g);
        g = nullptr;
      }
  }

  // glen_ iterations captured as one HIP graph, for the layout p (0..5) the
  // rotating blocks and factors are in now (they repeat every six
  // iterations); kernel arguments name only blocks, device scalars, the
  // projection basis and its coefficient block, and work space the batch owns
  void capture(int p) {
    Runtime& r = rt();
    hipStream_t st = r.stream;
    int saved[6];
    std::copy(role_, role_ + 6, saved);
    const int64_t saved_phase = phase_;
    for (int attempt = 0; attempt < 2; attempt++) {
      r.work_arena = garena_;
      r.work_need = 0;
      if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) break;
      r.capturing = true;
      bool ok = true;
      try {
        for (int i = 0; i < glen_; i++) enqueue_iteration();
      } catch (...) {
        ok = false;
      }
      r.capturing = false;
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(st, &g);
      std::copy(saved, saved + 6, role_);
      phase_ = saved_phase;
      if (!ok || e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        break;
      }
      if (r.work_need > garena_bytes_) {  // recorded with too small a work space: never launched
        (void)hipGraphDestroy(g);
        drop_graphs();  // the other layout's graph names the old work space
        if (garena_) dev_release(garena_, garena_bytes_);
        garena_bytes_ = r.work_need;
        garena_ = dev_alloc(garena_bytes_);
        continue;
      }
      hipGraphExec_t ex = nullptr;
      const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ei != hipSuccess || !ex) break;
      gexec_[p] = ex;
      r.graph_captures++;
      return;
    }
    (void)hipGetLastError();
    graph_off_ = true;  // this batch launches its kernels one by one from now on
  }

  void iterate_dev() {
    Runtime& r = rt();
    const int k = cap_;
    int64_t rem = 0;  // iterations the busiest active system may still take beyond those queued
    for (int j = 0; j < k; j++)
      if (st_[j].m.active)
        rem = std::max<int64_t>(rem, (int64_t)st_[j].m.max_iter - st_[j].m.iters - (seq_ - view_seq_[j]));
    if (rem <= 0 && pending_ >= 0) {
      drain();
      return;
    }
    // small blocks are bound by the host's launches and round trips: two
    // iterations per step there (measured marginally faster on a 64^3 JD)
    const int ipc = r.minres_iters_per_call > 0 ? r.minres_iters_per_call
                                                : (cur(V)->nrows * (int64_t)cap_ <= (int64_t(1) << 22) ? 2 : 1);
    const int n = (int)std::max<int64_t>(1, std::min<int64_t>(ipc, rem));
    const bool graphs = r.minres_graphs && !graph_off_ && !r.ktimer_on && cur(V)->nrows > 0 &&
                        (!Q_ || Q_->nvec <= kFusedProjCols);
    if (graphs) {  // the graphs hold the projection's pointers and width
      const T* qd = Q_ ? Q_->dev() : nullptr;
      const int nq = Q_ ? Q_->nvec : 0;
      const int64_t ldq = Q_ ? Q_->lda : 0;
      const T* cd = Q_ ? C_->get()->dev() : nullptr;
      const int glen = std::max(1, r.minres_graph_len);
      if (qd != gq_ || nq != gnq_ || ldq != gldq_ || cd != gc_ || glen != glen_) {
        drop_graphs();
        gq_ = qd;
        gnq_ = nq;
        gldq_ = ldq;
        gc_ = cd;
        glen_ = glen;
      }
    }
    for (int i = 0; i < n;) {
      if (graphs && !graph_off_ && n - i >= glen_) {
        const int p = (int)(phase_ % 6);  // the layout of the rotating blocks and factors
        if (!gexec_[p]) capture(p);
        if (gexec_[p]) {
          PHX_HIP(hipGraphLaunch(gexec_[p], r.stream));
          r.graph_launches++;
          for (int t = 0; t < glen_; t++) advance_roles();
          i += glen_;
          continue;
        }
      }
      enqueue_iteration();
      i++;
    }
    // a copy of the state after these iterations; the host view takes the
    // previous call's copy (lagging one call behind the queue)
    seq_ += n;
    const int b = pending_ == 0 ? 1 : 0;
    hipStream_t st = r.stream;
    PHX_HIP(hipMemcpyAsync(h_buf_ + (size_t)b * cap_, d_st_, sizeof(MinresState) * cap_, hipMemcpyDeviceToHost, st));
    PHX_HIP(hipEventRecord(ev_[b], st));
    copy_seq_[b] = seq_;
    if (pending_ >= 0) apply(pending_);
    pending_ = b;
  }

  void iterate_host() {
    const int k = cap_;
    std::vector<T> s(k), alpha(k), la(k), lb(k), bnext(k), dl(k), ep(k), ig(k), ta(k), ib(k);
    for (int j = 0; j < k; j++) s[j] = st_[j].m.active ? st_[j].shift : T(0);
    Mvec<T>* v = cur(V);
    Mvec<T>* w = cur(W);
    op_.apply_shifted(T(1), s.data(), v, T(0), w, alpha.data());
    if (Q_) {
      gemm_gram<T>(T(1), Q_, w, T(0), C_->get());
      gemm_times_sd<T>(T(-1), Q_, C_->get(), T(1), w);
    }
    for (int j = 0; j < k; j++) {
      la[j] = st_[j].m.active ? alpha[j] : T(0);
      lb[j] = st_[j].m.active && st_[j].m.iters > 0 ? (T)st_[j].m.beta : T(0);
    }
    const int64_t ld = cap_;
    krylov_lanczos<T>(v->nrows, k, ld, v->dev(), cur(VOLD)->dev(), w->dev(), la.data(), lb.data(), bnext.data());
    for (int j = 0; j < k; j++) {
      dl[j] = ep[j] = ig[j] = ta[j] = ib[j] = T(0);
      MinresState& q = st_[j].m;
      if (!q.active) continue;
      // column `iters` of the Lanczos tridiagonal: beta_k above, alpha_k on and
      // beta_{k+1} below the diagonal; two old rotations, one new one
      const double bk = q.iters > 0 ? q.beta : 0.0, ak = (double)alpha[j], bn = (double)bnext[j];
      const double eps_k = q.s2 * bk, dbar = q.c2 * bk;
      const double delta = q.c1 * dbar + q.s1 * ak;
      const double gbar = -q.s1 * dbar + q.c1 * ak;
      const double gamma = std::hypot(gbar, bn);
      if (!(gamma > 0.0) || !std::isfinite(gamma)) {  // singular or broken down: stop this system
        q.active = 0;
        continue;
      }
      const double c = gbar / gamma, sn = bn / gamma;
      const double tau = c * q.phibar;
      q.phibar = -sn * q.phibar;
      dl[j] = (T)delta;
      ep[j] = (T)eps_k;
      ig[j] = (T)(1.0 / gamma);
      ta[j] = (T)tau;
      ib[j] = bn > 0.0 ? (T)(1.0 / bn) : T(0);
      q.c2 = q.c1;
      q.s2 = q.s1;
      q.c1 = c;
      q.s1 = sn;
      q.beta = bn;
      q.iters++;
      q.relres = std::fabs(q.phibar) / q.beta1;
      if (q.relres <= q.tol || bn == 0.0 || q.iters >= q.max_iter) q.active = 0;
    }
    krylov_minres_update<T>(v->nrows, k, ld, v->dev(), cur(VOLD)->dev(), w->dev(), cur(D1)->dev(), cur(D2)->dev(),
                            cur(XS)->dev(), dl.data(), ep.data(), ig.data(), ta.data(), ib.data());
    std::swap(role_[V], role_[VOLD]);  // the next Lanczos vector was written over the old one
    std::swap(role_[D1], role_[D2]);   // the new direction was written over d_{k-2}
  }

  const Operator<T>& op_;
  const Map* map_;
  int cap_;
  const Mvec<T>* Q_;
  Blk blk_[6];  // V, VOLD, W, D1, D2, XS (roles rotate)
  int role_[6] = {0, 1, 2, 3, 4, 5};
  std::vector<Slot> st_;
  std::unique_ptr<SdBox<T>> C_;
  // device-resident scalars
  bool dev_ = false;
  MinresState* d_st_ = nullptr;
  T* d_dots_ = nullptr;
  T* d_work_ = nullptr;
  T* d_shift_ = nullptr;
  MinresState* h_buf_ = nullptr;  // pinned: two lagged state copies, then the upload staging
  T* h_shift_ = nullptr;          // pinned staging of the shifts
  hipEvent_t ev_[3] = {nullptr, nullptr, nullptr};  // two lagged state copies, the last upload
  int64_t seq_ = 0, copy_seq_[2] = {0, 0}, uploads_ = 0;  // seq_: iterations queued so far
  int pending_ = -1;
  std::vector<int64_t> load_seq_;  // slot j's state was uploaded before iteration load_seq_[j]
  std::vector<int64_t> view_seq_;  // the host view of slot j is its state after view_seq_[j] iterations
  // captured iterations, one per role layout, and what they were captured with
  hipGraphExec_t gexec_[6] = {};
  int64_t phase_ = 0;  // device iterations enqueued (rotation of the blocks and factors)
  const T* gq_ = nullptr;
  int gnq_ = 0;
  int64_t gldq_ = 0;
  const T* gc_ = nullptr;
  int glen_ = 1;  // iterations per graph
  void* garena_ = nullptr;  // device work space of the captured kernels
  size_t garena_bytes_ = 0;
  bool graph_off_ = false;
};

template <typename T>
struct MinresOut {
  std::vector<int> iters;
  std::vector<double> relres;
  int unconverged = 0;
};

// Repeated solves of k systems with changing shifts, projectors and
// right-hand sides (the correction equations of one eigensolver run): the
// batch and, for a lower inner precision TI, the converted projection basis
// are kept between solves.
template <typename T, typename TI>
class MinresSolver {
 public:
  MinresSolver(const Operator<TI>& op, const Map* map, int k) : mb_(op, map, k), map_(map) {}
  MinresOut<TI> solve(const std::vector<T>& s, const Mvec<T>* Q, const Mvec<T>* B, Mvec<T>* X,
                      const std::vector<double>& tol, int maxIter) {
    const int k = B->nvec;
    const Mvec<TI>* Qi = nullptr;
    if (Q && Q->nvec > 0) {
      // every iteration reads Q twice (Gram and projection); a view that
      // skips most of each row of its parent is first copied into a block of
      // its own width (as is a basis of another precision)
      bool copy = 2 * Q->nvec <= Q->lda;
      if constexpr (!std::is_same<T, TI>::value) copy = true;
      if (copy) {
        if (!q_.p || q_->nvec != Q->nvec) {
          MvBox<TI> fresh(map_, Q->nvec);
          std::swap(q_.p, fresh.p);
        }
        copy_any<T, TI>(Q, q_.get());
        Qi = q_.get();
      } else {
        if constexpr (std::is_same<T, TI>::value) Qi = Q;
      }
    }
    mb_.set_projection(Qi);
    std::vector<TI> si(s.begin(), s.end());
    mb_.load_all(B, si, tol, std::vector<int>(k, maxIter));
    while (mb_.any_active()) mb_.iterate();
    mb_.drain();
    mb_.store_all(X);
    mb_.set_projection(nullptr);
    MinresOut<TI> out;
    out.iters.resize(k);
    out.relres.resize(k);
    for (int j = 0; j < k; j++) {
      out.iters[j] = mb_.iters(j);
      out.relres[j] = mb_.relres(j);
      if (!mb_.converged(j)) out.unconverged++;
    }
    return out;
  }

 private:
  MinresBatch<TI> mb_;
  const Map* map_;
  MvBox<TI> q_;
};

// k systems at once in precision TI (B and X in precision T, Q converted)
template <typename T, typename TI>
MinresOut<TI> minres_solve(const Operator<TI>& op, const std::vector<T>& s, const Mvec<T>* Q, const Mvec<T>* B,
                           Mvec<T>* X, const std::vector<double>& tol, int maxIter) {
  MinresSolver<T, TI> solver(op, B->map, B->nvec);
  return solver.solve(s, Q, B, X, tol, maxIter);
}

// nsys systems (B, X with nsys columns) through a batch of `slots` columns:
// a system whose slot stops (converged, broken down or at maxIter) has its
// solution stored and the slot is loaded with the next system while the
// others keep iterating (phist_jadaCorrectionSolver_def.hpp:428-466)
template <typename T, typename TI>
MinresOut<TI> minres_refill_solve(const Operator<TI>& op, const std::vector<T>& s, const Mvec<T>* B, Mvec<T>* X,
                                  const std::vector<double>& tol, int maxIter, int slots) {
  const int nsys = B->nvec;
  MinresBatch<TI> mb(op, B->map, slots);
  std::vector<int> owner(slots, -1);
  MinresOut<TI> out;
  out.iters.assign(nsys, 0);
  out.relres.assign(nsys, 0.0);
  int next = 0, done = 0;
  auto fill = [&] {
    int j;
    while (next < nsys && (j = mb.free_slot()) >= 0) {
      mb.load(j, B, next, (TI)s[next], tol[next], maxIter);
      owner[j] = next++;
    }
  };
  fill();
  while (done < nsys) {
    mb.iterate();
    for (int j = 0; j < slots; j++) {
      if (owner[j] < 0 || mb.active(j)) continue;
      const int c = owner[j];
      mb.store(j, X, c);
      out.iters[c] = mb.iters(j);
      out.relres[c] = mb.relres(j);
      if (!mb.converged(j)) out.unconverged++;
      mb.release(j);
      owner[j] = -1;
      done++;
    }
    fill();
  }
  sync_stream();
  return out;
}

// restarted GMRES(m) on k shifted systems (gmres.cpp), x from X as the
// initial guess; iters = Arnoldi steps per system, relres = ||r|| / ||r0||
struct GmresOut {
  std::vector<int> iters;
  std::vector<double> relres;
  int unconverged = 0;
};
template <typename T, typename TI>
GmresOut gmres_solve(const Operator<TI>& op, const std::vector<T>& shifts, const Mvec<T>* B, Mvec<T>* X,
                     const std::vector<double>& tol, int maxIter, int m, const Mvec<TI>* Q = nullptr);

// resumable block Jacobi-Davidson on a sparse matrix of this library
// (defined in jada.cpp): start builds the start-up basis, step runs one outer
// iteration (true when finished), finish writes the outputs
struct JadaHandle;
// throws -99 for options outside what subspacejada supports, -55 for inconsistent sizes;
// matrix_symmetric: the operator is a matrix of this library, symmetric entry by entry
// (then the reference's default symmetry, phist_GENERAL, is accepted)
void jada_validate(const phist_jadaOpts& opts, bool matrix_symmetric);
template <typename T> bool sparse_is_symmetric(const SparseMat<T>* A);
JadaHandle* jada_start(const SparseMat<double>* A, bool fp32, const phist_jadaOpts& opts, Mvec<double>* Q,
                       SdMat<double>* R, double* ev, double* resNorm);
bool jada_step(JadaHandle* h);
void jada_finish(JadaHandle* h, int* nConv, int* nIter);
void jada_free(JadaHandle* h);

}  // namespace phx
