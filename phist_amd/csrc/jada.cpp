// jada.cpp -- device-resident eigensolver and shifted linear solvers.
//
//   MinresBatch      k shifted symmetric systems P(A + s_j I)P x_j = b_j at
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
//                    go to MinresBatch, in FP64 or on an FP32 copy of A.
//   linearOp_wrap_sparseMat, jadaOpts_setDefaults: the reference's operator
//                    wrapper (src/core/phist_operator_def.hpp:84-106) and
//                    option defaults (src/jada/phist_jadaOpts.cpp:19-58).
#include "solvers.hpp"
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

namespace phx {

namespace dense {
void sym_eig(double* a, int n, int lda, double* w);
void sym_eig_tridiag(double* a, int n, int lda, double* w);
}

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

// Block Jacobi-Davidson as a resumable object: the constructor builds the
// start-up basis, step() runs one outer iteration (true when finished:
// converged or maxIters reached), finish() writes Q, R, ev and resNorm.
template <typename T>
class JadaRun {
 public:
  using MT = double;
  JadaRun(const Operator<T>& A_, const Operator<float>* A32_, const phist_jadaOpts& opts_, Mvec<T>* Qout_,
          SdMat<T>* Rout_, T* ev_, T* resNormOut_)
      : A(A_),
        A32(A32_),
        opts(opts_),
        Qout(Qout_),
        Rout(Rout_),
        ev(ev_),
        resNormOut(resNormOut_),
        tol(opts_.convTol),
        nEig(opts_.numEigs),
        bs(opts_.blockSize),
        lookAhead(opts_.lookAhead > 0 ? opts_.lookAhead : 2 * opts_.blockSize),
        maxIter(opts_.maxIters),
        minBase(opts_.minBas),
        maxBase(opts_.maxBas),
        innerMaxIters(opts_.innerSolvMaxIters),
        innerTolBase(opts_.innerSolvBaseTol),
        useRelTol(opts_.relConvTol != 0),
        initialShift(opts_.initialShift_r),
        which(opts_.which),
        eps(std::numeric_limits<T>::epsilon()),
        nEigMax(opts_.numEigs + opts_.blockSize - 1) {
    PHX_REQUIRE(bs >= 1, PHIST_INVALID_INPUT, "blockSize %d", bs);
    PHX_REQUIRE(minBase >= nEigMax && minBase + bs <= maxBase && maxBase >= nEig + bs, PHIST_INVALID_INPUT,
                "basis sizes minBas=%d maxBas=%d do not fit numEigs=%d blockSize=%d", minBase, maxBase, nEig, bs);
    PHX_REQUIRE(Qout->nvec >= nEigMax && Rout->nrows == Qout->nvec && Rout->ncols == Qout->nvec, PHIST_INVALID_INPUT,
                "Q needs >= %d columns and R matching square", nEigMax);
    PHX_REQUIRE(maxBase <= 256, PHIST_INVALID_INPUT, "maxBas %d too large", maxBase);
    for (int i = 0; i < nEig; i++) resNormOut[i] = T(-1);

    const Map* map = Qout->map;
    V_.p = api::mvec_create<T>(map, maxBase);
    AV_.p = api::mvec_create<T>(map, maxBase);
    Q_.p = api::mvec_create<T>(map, nEigMax);
    t_.p = api::mvec_create<T>(map, bs);
    tback_.p = api::mvec_create<T>(map, bs);
    oC_.p = api::sdMat_create<T>(maxBase, bs);
    oC2_.p = api::sdMat_create<T>(maxBase, bs);
    oG_.p = api::sdMat_create<T>(bs, bs);
    oR_.p = api::sdMat_create<T>(bs, bs);
    oF_.p = api::sdMat_create<T>(2, 1);
    if (const char* e = getenv("PHIST_AMD_JADA_ORTH")) orth_mode_ = atoi(e);
    res_.p = api::mvec_create<T>(map, std::min(lookAhead, nEigMax));
    H_.p = api::sdMat_create<T>(maxBase, maxBase);
    solve_.resize(bs + 1);
    solve32_.resize(bs + 1);
    theta.assign(maxBase, 0.0);
    resNorm.assign(nEigMax, -1.0);
    resScaling.assign(nEigMax, 1.0);
    innerTol.assign(nEigMax, innerTolBase);
    lastOuterRes.assign(nEigMax, 0.0);
    selected.assign(bs, 0);
    sigma.assign(bs, T(0));
    api::sdMat_put_value<T>(H_.get(), T(0));
    api::sdMat_put_value<T>(Rout, T(0));
    // ---- start-up basis ----
    nV = 0;
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
      // v_j and A v_j live in contiguous single-column blocks (wcol, acol):
      // the SpMV gathers rows of its input, and the orthogonalisation reads
      // and rewrites the new vector several times, where a column of the wide
      // basis would cost a whole cache line per row for 8 useful bytes; each
      // is copied into the basis once
      MvBox<T> wcol(map, 1), acol(map, 1), ncol(map, 1);
      // first pass without host round trips: every step's Gram-Schmidt
      // coefficients and squared norm stay on the device until the end; a
      // step whose A v_j lies numerically in the basis sends the whole
      // start-up through the step-by-step robust variant below
      bool fast_ok = true;
      {
        SdBox<T> hall(maxBase, 2 * maxBase), n2all(maxBase, 1);
        {
          View<T> v0(V_.get(), 0, 0);
          blas_copy<T>(v0.get(), wcol.get());
        }
        int nsteps = 0;
        for (int j = 0; j < nV; j++) {
          View<T> avj(AV_.get(), j, j);
          A.apply_shifted(T(1), nullptr, wcol.get(), T(0), acol.get(), nullptr);
          blas_copy<T>(acol.get(), avj.get());
          if (j + 1 >= maxBase) break;
          View<T> Vj(V_.get(), 0, j), vn(V_.get(), j + 1, j + 1);
          blas_copy<T>(acol.get(), ncol.get());
          // classical Gram-Schmidt twice against the orthonormal V(:, 0:j)
          SdView<T> h1(hall.get(), 0, j, 2 * j, 2 * j), h2(hall.get(), 0, j, 2 * j + 1, 2 * j + 1);
          gemm_gram<T>(T(1), Vj.get(), ncol.get(), T(0), h1.get(), false);
          gemm_times_sd<T>(T(-1), Vj.get(), h1.get(), T(1), ncol.get());
          gemm_gram<T>(T(1), Vj.get(), ncol.get(), T(0), h2.get(), false);
          gemm_times_sd<T>(T(-1), Vj.get(), h2.get(), T(1), ncol.get());
          blas_dot_dev<T>(ncol.get(), ncol.get(), n2all->dev() + j);
          blas_scale_inv_norm_dev<T>(ncol.get(), n2all->dev() + j);
          blas_copy<T>(ncol.get(), vn.get());
          std::swap(wcol.p, ncol.p);  // v_{j+1} is the next input
          nsteps = j + 1;
        }
        sd_from_device_async<T>(hall.get());
        sd_from_device<T>(n2all.get());  // one synchronisation for the whole start-up
        for (int j = 0; j < nsteps && fast_ok; j++) {
          double hn2 = 0.0;
          for (int i = 0; i <= j; i++) {
            const double h = (double)hall->h(i, 2 * j) + (double)hall->h(i, 2 * j + 1);
            hn2 += h * h;
          }
          const double n2 = (double)n2all->h(j, 0), nrm = std::sqrt(n2);
          if (!(nrm > 1e-12 * std::sqrt(hn2 + n2)) || !std::isfinite(nrm)) {
            fast_ok = false;
            break;
          }
          for (int i = 0; i <= j; i++) H(i, j) = hall->h(i, 2 * j) + hall->h(i, 2 * j + 1);
          H(j + 1, j) = (T)nrm;
        }
      }
      if (!fast_ok) {
        SdBox<T> hc(maxBase, 2);  // the two Gram-Schmidt passes' coefficients (device, then host)
        {
          View<T> v0(V_.get(), 0, 0);
          blas_copy<T>(v0.get(), wcol.get());
        }
        for (int j = 0; j < nV; j++) {
          View<T> avj(AV_.get(), j, j);
          A.apply_shifted(T(1), nullptr, wcol.get(), T(0), acol.get(), nullptr);
          blas_copy<T>(acol.get(), avj.get());
          if (j + 1 >= maxBase) break;
          View<T> Vj(V_.get(), 0, j), vn(V_.get(), j + 1, j + 1);
          blas_copy<T>(acol.get(), ncol.get());
          // classical Gram-Schmidt twice against the orthonormal V(:, 0:j), the
          // coefficients kept on the device: one host round trip per step (the
          // norm), and the robust orthogonalisation where the norm is too small
          SdView<T> h1(hc.get(), 0, j, 0, 0), h2(hc.get(), 0, j, 1, 1), hh(hc.get(), 0, j, 0, 1);
          gemm_gram<T>(T(1), Vj.get(), ncol.get(), T(0), h1.get(), false);
          gemm_times_sd<T>(T(-1), Vj.get(), h1.get(), T(1), ncol.get());
          gemm_gram<T>(T(1), Vj.get(), ncol.get(), T(0), h2.get(), false);
          gemm_times_sd<T>(T(-1), Vj.get(), h2.get(), T(1), ncol.get());
          sd_from_device_async<T>(hh.get());
          T n2;
          blas_dot<T>(ncol.get(), ncol.get(), &n2);  // synchronises: hh is on the host too
          double hn2 = 0.0;
          for (int i = 0; i <= j; i++) {
            const double h = (double)hh->h(i, 0) + (double)hh->h(i, 1);
            hn2 += h * h;
          }
          const double nrm = std::sqrt((double)n2);
          if (nrm > 1e-12 * std::sqrt(hn2 + (double)n2) && std::isfinite(nrm)) {
            const T sc = (T)(1.0 / nrm);
            blas_scale_cols<T>(ncol.get(), &sc, true);
            for (int i = 0; i <= j; i++) H(i, j) = hh->h(i, 0) + hh->h(i, 1);
            H(j + 1, j) = (T)nrm;
          } else {
            // A v_j (numerically) in span(V): the robust orthogonalisation continues
            // with a random orthogonal vector (phist_simple_arnoldi_def.hpp:118-141)
            blas_copy<T>(acol.get(), ncol.get());
            SdBox<T> R1(1, 1), R2(j + 1, 1);
            int rank = 0;
            const int dim0 = api::orthog<T>(Vj.get(), ncol.get(), R1.get(), R2.get(), 3, &rank, 131072);
            for (int i = 0; i <= j; i++) H(i, j) = R2->h(i, 0);
            H(j + 1, j) = R1->h(0, 0);
            if (dim0 == 0) {  // a more accurate subdiagonal entry: v_{j+1}' A v_j
              T h;
              blas_dot<T>(ncol.get(), acol.get(), &h);
              H(j + 1, j) = h;
            }
          }
          blas_copy<T>(ncol.get(), vn.get());
          std::swap(wcol.p, ncol.p);  // v_{j+1} is the next input
        }
      }
      // H = V'AV over the first nV columns (the square part of the Arnoldi
      // Hessenberg); symmetrise exactly
      for (int j = 0; j < nV; j++)
        for (int i = j + 1; i < nV; i++) H(j, i) = H(i, j);
      sd_to_device(H_.get());
    }

  }

  int iterations() const { return iter; }
  int converged() const { return nConv; }
  bool done() const { return done_; }

  bool step() {
    if (done_) return true;
    if (iter >= maxIter) {
      done_ = true;
      return true;
    }
    const int nEig_ = std::min(nConv + lookAhead, nEigMax);
    const int m = nV - nConv;
    // Ritz pairs of H(nConv:nV, nConv:nV), sorted by `which`
    hb.assign((size_t)m * m, 0.0);
    for (int j = 0; j < m; j++)
      for (int i = 0; i < m; i++) hb[(size_t)j * m + i] = (double)H(nConv + i, nConv + j);
    std::vector<double> w(m);
    dense::sym_eig_tridiag(hb.data(), m, m, w.data());
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
      done_ = true;
      return true;
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
    if (k == 0) {
      iter++;
      return false;
    }
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
      MvBox<T> rhs(Qout->map, k);
      bool run = true;  // selected pairs are usually consecutive: one block copy
      for (int i = 1; i < k; i++) run = run && selected[i] == selected[0] + i;
      if (run) {
        View<T> src(res_.get(), selected[0] - oldConv, selected[0] - oldConv + k - 1);
        blas_copy<T>(src.get(), rhs.get());
      } else {
        for (int i = 0; i < k; i++) {
          View<T> src(res_.get(), selected[i] - oldConv, selected[i] - oldConv), dst(rhs.get(), i, i);
          blas_copy<T>(src.get(), dst.get());
        }
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
      } else if (opts.innerSolvType == phist_GMRES) {
        // blocked GMRES on P (A - sigma I) P, P = I - QQ' (phist_jadaCorrectionSolver_def.hpp:499-512),
        // from t = 0, basis size innerSolvMaxBas (innerSolvMaxIters when unset, :37-39): no restart
        // within the iteration limit by default.  FP32 operator and basis on the FP32 copy of A.
        const int m = std::max(1, std::min(opts.innerSolvMaxBas > 0 ? opts.innerSolvMaxBas : innerMaxIters, 256));
        blas_put_value<T>(t.get(), T(0));
        GmresOut r;
        if (A32) {
          for (double& x : tl) x = std::max(x, 1.0e-6);  // FP32 solves stagnate below this
          const Mvec<float>* q32 = nullptr;
          if (Qtil->nvec > 0) {
            if (!q32_.p || q32_->nvec != Qtil->nvec) {
              MvBox<float> fresh(Qout->map, Qtil->nvec);
              std::swap(q32_.p, fresh.p);
            }
            copy_any<T, float>(Qtil.get(), q32_.get());
            q32 = q32_.get();
          }
          r = gmres_solve<T, float>(*A32, s, rhs.get(), t.get(), tl, innerMaxIters, m, q32);
        } else {
          r = gmres_solve<T, T>(A, s, rhs.get(), t.get(), tl, innerMaxIters, m, Qtil.get());
        }
        inner_it = r.iters;
        inner_rr = r.relres;
      } else if (A32) {
        for (double& x : tl) x = std::max(x, 1.0e-6);  // FP32 solves stagnate below this
        if (!solve32_[k]) solve32_[k].reset(new MinresSolver<T, float>(*A32, Qout->map, k));
        MinresOut<float> r = solve32_[k]->solve(s, Qtil.get(), rhs.get(), t.get(), tl, innerMaxIters);
        inner_it = r.iters;
        inner_rr = r.relres;
      } else {
        if (!solve_[k]) solve_[k].reset(new MinresSolver<T, T>(A, Qout->map, k));
        MinresOut<T> r = solve_[k]->solve(s, Qtil.get(), rhs.get(), t.get(), tl, innerMaxIters);
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
      View<T> tb(tback_.get(), 0, k - 1);
      // orthonormalise the narrow block t (row stride bs) and copy it into
      // the basis afterwards: the Gram and in-place updates of a k-column
      // view of the wide basis (row stride maxBas) would move a whole cache
      // line per row for k*8 useful bytes.  The device-resident variant needs
      // no host round trip; if its Cholesky finds t (numerically) rank
      // deficient, the robust orthogonalisation redoes the step from a copy.
      // No rank truncation: near convergence t is tiny, and a rank test would
      // replace it by noise (the reference passes no flags)
      blas_copy<T>(t.get(), tb.get());
      if (orth_mode_ != 0) orth_dev(Vful.get(), t.get(), k);
      blas_copy<T>(t.get(), Vv.get());
      // the SpMV gathers rows of its input: take them from t as well
      A.apply_shifted(T(1), nullptr, t.get(), T(0), AVv.get(), nullptr);
      View<T> Vall(V_.get(), 0, nV + k - 1);
      SdBox<T> Hn(nV + k, k);
      sd_from_device_async<T>(oF_.get());
      gemm_gram<T>(T(1), Vall.get(), AVv.get(), T(0), Hn.get());  // synchronises: the flags are on the host
      if (orth_mode_ != 1 || oF_->h(0, 0) != T(0) || oF_->h(1, 0) != T(0)) {
        blas_copy<T>(tb.get(), t.get());
        SdBox<T> R1(k, k), R2(nV, k);
        int rank = 0;
        api::orthog<T>(Vful.get(), t.get(), R1.get(), R2.get(), 3, &rank, 0);
        blas_copy<T>(t.get(), Vv.get());
        A.apply_shifted(T(1), nullptr, t.get(), T(0), AVv.get(), nullptr);
        gemm_gram<T>(T(1), Vall.get(), AVv.get(), T(0), Hn.get());
      }
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
    iter++;
    return false;
  }

  void finish(int* nConvOut, int* nIterOut) {
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

 private:
  T& H(int i, int j) { return H_->h(i, j); }

  const Operator<T>& A;
  const Operator<float>* A32;
  const phist_jadaOpts opts;
  Mvec<T>* Qout;
  SdMat<T>* Rout;
  T* ev;
  T* resNormOut;
  const MT tol;
  const int nEig, bs, lookAhead, maxIter, minBase, maxBase, innerMaxIters;
  const MT innerTolBase;
  const bool useRelTol;
  const MT initialShift;
  const phist_EeigSort which;
  const MT eps;
  const int nEigMax;
  MvBox<T> V_, AV_, Q_, t_, res_;
  SdBox<T> H_;
  // device-resident orthonormalisation of the correction block: a copy of
  // the block for the robust path, coefficient and Gram work space, flags
  MvBox<T> tback_;
  SdBox<T> oC_, oC2_, oG_, oR_, oF_;
  // PHIST_AMD_JADA_ORTH: 1 (default) device-resident, robust path when a
  // Cholesky fails; 0 robust path only; 2 both, the robust result kept (for
  // tests of the fallback)
  int orth_mode_ = 1;

  // t (k columns) against the orthonormal V: classical Gram-Schmidt twice,
  // then Cholesky QR twice, all on the device; oF_ holds one flag per
  // Cholesky (1 = a pivot too small for t to be taken as full rank)
  void orth_dev(const Mvec<T>* V, Mvec<T>* t, int k) {
    const int m = V->nvec;
    SdView<T> C(oC_.get(), 0, m - 1, 0, k - 1), C2(oC2_.get(), 0, m - 1, 0, k - 1), G(oG_.get(), 0, k - 1, 0, k - 1),
        R(oR_.get(), 0, k - 1, 0, k - 1);
    // the first pass's update and the second pass's Gram share one sweep over V
    gemm_gram<T>(T(1), V, t, T(0), C.get(), false);
    gemm_sub_gram_dev<T>(V, C.get(), t, C2.get());
    gemm_times_sd<T>(T(-1), V, C2.get(), T(1), t);
    for (int pass = 0; pass < 2; pass++) {
      gemm_gram<T>(T(1), t, t, T(0), G.get(), false);
      chol_inv_dev<T>(G.get(), R.get(), oF_->dev() + pass);
      gemm_times_sd_inplace<T>(t, R.get());
    }
  }
  // correction-equation solvers by number of systems (kept between outer iterations)
  std::vector<std::unique_ptr<MinresSolver<T, T>>> solve_;
  std::vector<std::unique_ptr<MinresSolver<T, float>>> solve32_;
  MvBox<float> q32_;  // FP32 copy of the projection basis for FP32 GMRES correction solves
  int nV = 0, nConv = 0, iter = 0;
  bool done_ = false;
  std::vector<double> theta, resNorm, resScaling, innerTol, lastOuterRes, hb, qh;
  std::vector<int> selected;
  std::vector<T> sigma;
};

template <typename T>
void subspacejada(const Operator<T>& A, const Operator<float>* A32, const phist_jadaOpts& opts, Mvec<T>* Qout,
                  SdMat<T>* Rout, T* ev, T* resNormOut, int* nConvOut, int* nIterOut) {
  JadaRun<T> run(A, A32, opts, Qout, Rout, ev, resNormOut);
  while (!run.step()) {
  }
  run.finish(nConvOut, nIterOut);
}

}  // namespace

// A == A' entry by entry for a square matrix of this library (the sums of
// duplicate entries compared, rows of the host CRS sorted by column); the
// answer is kept on the matrix
template <typename T>
bool sparse_is_symmetric(const SparseMat<T>* A) {
  if (A->symmetric >= 0) return A->symmetric == 1;
  bool sym = A->nrows == A->ncols;
  const int64_t n = A->nrows;
  const int64_t* ptr = A->crs_ptr.data();
  const int32_t* col = A->crs_col.data();
  const T* val = A->crs_val.data();
  // sum of the entries (r, c) of row r (0 when absent): binary search in the sorted row
  auto entry = [&](int64_t r, int32_t c) {
    const int32_t* lo = std::lower_bound(col + ptr[r], col + ptr[r + 1], c);
    T s = T(0);
    for (const int32_t* p = lo; p < col + ptr[r + 1] && *p == c; p++) s += val[p - col];
    return s;
  };
  if (sym) {
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : bad)
    for (int64_t i = 0; i < n; i++) {
      for (int64_t p = ptr[i]; p < ptr[i + 1] && !bad; p++) {
        if (p > ptr[i] && col[p] == col[p - 1]) continue;  // a duplicate's sum was compared with its first entry
        if (!(entry(i, col[p]) == entry(col[p], (int32_t)i))) bad = 1;
      }
    }
    sym = bad == 0;
  }
  A->symmetric = sym ? 1 : 0;
  return sym;
}
template bool sparse_is_symmetric<double>(const SparseMat<double>*);
template bool sparse_is_symmetric<float>(const SparseMat<float>*);

// The option checks of phist_Xsubspacejada, shared by the direct call and
// the service loop: every combination outside the symmetric standard
// problem with MINRES, GMRES or no correction solver returns -99
// (PHIST_NOT_IMPLEMENTED) instead of running something else.  The
// reference's default symmetry (phist_GENERAL, phist_jadaOpts.cpp:21) is
// accepted when the operator is a matrix of this library that is symmetric
// entry by entry (matrix_symmetric): the symmetric algorithm then computes
// the same partial Schur form the general one would (R diagonal up to the
// couplings of locked pairs, Q orthonormal).
void jada_validate(const phist_jadaOpts& opts, bool matrix_symmetric) {
  PHX_REQUIRE(opts.how == phist_STANDARD, PHIST_NOT_IMPLEMENTED, "only standard Ritz extraction");
  PHX_REQUIRE(opts.symmetry == phist_HERMITIAN || opts.symmetry == phist_COMPLEX_SYMMETRIC ||
                  (opts.symmetry == phist_GENERAL && matrix_symmetric),
              PHIST_NOT_IMPLEMENTED,
              "only symmetric problems (opts.symmetry HERMITIAN, or GENERAL on a matrix of this library that is "
              "symmetric)");
  PHX_REQUIRE(opts.which == phist_LM || opts.which == phist_SM || opts.which == phist_LR || opts.which == phist_SR,
              PHIST_NOT_IMPLEMENTED, "eigenvalue selection %d not supported", (int)opts.which);
  PHX_REQUIRE(opts.innerSolvType == phist_MINRES || opts.innerSolvType == phist_GMRES ||
                  opts.innerSolvType == phist_NO_LINSOLV,
              PHIST_NOT_IMPLEMENTED, "inner solver %d not supported", (int)opts.innerSolvType);
  PHX_REQUIRE(opts.preconOp == nullptr, PHIST_NOT_IMPLEMENTED, "preconditioned correction equations");
  PHX_REQUIRE(opts.numEigs >= 1 && opts.blockSize >= 1 && opts.minBas >= 1 && opts.maxBas >= opts.minBas,
              PHIST_INVALID_INPUT, "numEigs %d, blockSize %d, minBas %d, maxBas %d", opts.numEigs, opts.blockSize,
              opts.minBas, opts.maxBas);
}

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
  Operator<T> op;
  op.A = own_matrix<T>(AB_op);
  if (!op.A) op.user = AB_op;
  jada_validate(opts, op.A && opts.symmetry == phist_GENERAL && sparse_is_symmetric(op.A));
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

struct JadaHandle {
  Operator<double> op;
  Operator<float> op32;
  std::unique_ptr<JadaRun<double>> run;
};

JadaHandle* jada_start(const SparseMat<double>* A, bool fp32, const phist_jadaOpts& opts, Mvec<double>* Q,
                       SdMat<double>* R, double* ev, double* resNorm) {
  jada_validate(opts, opts.symmetry == phist_GENERAL && sparse_is_symmetric(A));
  std::unique_ptr<JadaHandle> h(new JadaHandle());
  h->op.A = A;
  const Operator<float>* lowered = nullptr;
  if (fp32) {
    h->op32.A = sparse_fp32_twin(A);
    lowered = &h->op32;
  }
  h->run.reset(new JadaRun<double>(h->op, lowered, opts, Q, R, ev, resNorm));
  return h.release();
}
bool jada_step(JadaHandle* h) { return h->run->step(); }
void jada_finish(JadaHandle* h, int* nConv, int* nIter) { h->run->finish(nConv, nIter); }
void jada_free(JadaHandle* h) { delete h; }

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

}  // extern "C"

namespace {

// FP32 iterations refined in FP64 (mixed-precision iterative refinement):
// after an FP32 solve the true residuals r_j = b_j - (A + s_j) x_j are
// formed in FP64, and the systems still above tol[j] solve
// (A + s_j) d_j = r_j again in FP32, x_j += d_j.  An FP32 recurrence's own
// residual estimate parts from the true residual once it reaches the FP32
// attainable accuracy (about eps32 times the condition number of the shifted
// system), so an FP32 solve alone can report convergence far above tol.
// A round asks the FP32 solve for the reduction still needed, but not below
// kRefineInnerTol; refinement stops when every system is below tol, the
// iteration budget is spent, or a round improves no unconverged system by
// at least 10 %.  iters sums the rounds; relres is the true FP64 residual.
constexpr double kRefineInnerTol = 1e-3;
constexpr int kRefineRounds = 30;
template <class Solve>
MinresOut<double> fp32_refined(const SparseMat<double>* M, const std::vector<double>& s, const Mvec<double>* b,
                               Mvec<double>* x, const std::vector<double>& tol, int maxIter, Solve&& solve) {
  const int k = b->nvec;
  MinresOut<double> out;
  out.iters.assign(k, 0);
  out.relres.assign(k, 0.0);
  {
    const auto r0 = solve(b, x, tol, maxIter);
    for (int j = 0; j < k; j++) out.iters[j] = r0.iters[j];
  }
  Mvec<double>* R = api::mvec_create<double>(b->map, k);
  Mvec<double>* D = api::mvec_create<double>(b->map, k);
  try {
    std::vector<double> nb(k), nr(k), prev(k, 1e300), tin(k);
    blas_dot<double>(b, b, nb.data());
    for (auto& v : nb) v = std::sqrt(v);
    SpmvRequest rq;
    rq.has_shift = true;
    rq.has_beta = true;
    for (int round = 1;; round++) {
      blas_copy<double>(b, R);
      spmv_launch<double>(M, -1.0, s.data(), x, 1.0, R, rq, nullptr, nullptr);
      blas_dot<double>(R, R, nr.data());
      int used = 0;
      bool more = false, progress = false;
      for (int j = 0; j < k; j++) {
        out.relres[j] = nb[j] > 0.0 ? std::sqrt(nr[j]) / nb[j] : 0.0;
        const bool conv = out.relres[j] <= tol[j];
        if (!conv && out.iters[j] < maxIter) {
          more = true;
          used = std::max(used, out.iters[j]);
          if (out.relres[j] < 0.9 * prev[j]) progress = true;
        }
        prev[j] = out.relres[j];
        tin[j] = conv ? 2.0 : std::min(0.5, std::max(kRefineInnerTol, tol[j] / out.relres[j]));
      }
      if (!more || round > kRefineRounds || (round > 1 && !progress)) break;
      blas_put_value<double>(D, 0.0);
      const auto r = solve(R, D, tin, maxIter - used);
      for (int j = 0; j < k; j++)
        if (out.relres[j] > tol[j]) out.iters[j] += r.iters[j];
      const double one = 1.0;
      blas_axpby<double>(&one, true, D, 1.0, x);
    }
  } catch (...) {
    api::mvec_delete<double>(R);
    api::mvec_delete<double>(D);
    throw;
  }
  api::mvec_delete<double>(R);
  api::mvec_delete<double>(D);
  out.unconverged = 0;
  for (int j = 0; j < k; j++) out.unconverged += out.relres[j] > tol[j] ? 1 : 0;
  return out;
}

}  // namespace

extern "C" {

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
      const MinresOut<double> r =
          fp32_refined(M, s, b, x, tl, maxIter, [&](const Mvec<double>* rhs, Mvec<double>* sol,
                                                   const std::vector<double>& t, int budget) {
            return minres_solve<double, float>(op, s, nullptr, rhs, sol, t, budget);
          });
      for (int j = 0; j < nsys; j++) {
        nIter[j] = r.iters[j];
        relres[j] = r.relres[j];
      }
      unconv = r.unconverged;
    } else {
      Operator<double> op;
      op.A = M;
      MinresOut<double> r = minres_solve<double, double>(op, s, nullptr, b, x, tl, maxIter);
      for (int j = 0; j < nsys; j++) {
        nIter[j] = r.iters[j];
        relres[j] = r.relres[j];
      }
      unconv = r.unconverged;
    }
    *iflag = unconv;
  });
}

void phist_amd_DblockedMINRES_shifted_refill(phist_Dconst_sparseMat_ptr A, int nsys, const double shifts[],
                                             phist_Dconst_mvec_ptr B, phist_Dmvec_ptr X, const double tol[],
                                             int maxIter, int fp32, int slots, int nIter[], double relres[],
                                             int* iflag) {
  *iflag = 0;
  guard(iflag, "phist_amd_DblockedMINRES_shifted_refill", [&] {
    const auto* M = reinterpret_cast<const SparseMat<double>*>(A);
    const auto* b = reinterpret_cast<const Mvec<double>*>(B);
    auto* x = reinterpret_cast<Mvec<double>*>(X);
    PHX_REQUIRE(M && b && x && shifts && tol && nIter && relres, PHIST_INVALID_INPUT, "null argument");
    PHX_REQUIRE(nsys >= 1 && b->nvec == nsys && x->nvec == nsys, PHIST_INVALID_INPUT,
                "nsys %d must match B and X", nsys);
    PHX_REQUIRE(slots >= 1 && slots <= 64, PHIST_INVALID_INPUT, "slots %d not in 1..64", slots);
    PHX_REQUIRE(M->nrows == M->ncols && b->nrows == M->nrows && x->nrows == M->nrows, PHIST_INVALID_INPUT,
                "shapes do not match the square matrix");
    std::vector<double> s(shifts, shifts + nsys), tl(tol, tol + nsys);
    int unconv = 0;
    auto put = [&](const auto& r) {
      for (int j = 0; j < nsys; j++) {
        nIter[j] = r.iters[j];
        relres[j] = r.relres[j];
      }
      unconv = r.unconverged;
    };
    if (fp32) {
      Operator<float> op;
      op.A = sparse_fp32_twin(M);
      put(fp32_refined(M, s, b, x, tl, maxIter, [&](const Mvec<double>* rhs, Mvec<double>* sol,
                                                   const std::vector<double>& t, int budget) {
        return minres_refill_solve<double, float>(op, s, rhs, sol, t, budget, std::min(slots, nsys));
      }));
    } else {
      Operator<double> op;
      op.A = M;
      put(minres_refill_solve<double, double>(op, s, b, x, tl, maxIter, std::min(slots, nsys)));
    }
    *iflag = unconv;
  });
}

}  // extern "C"
