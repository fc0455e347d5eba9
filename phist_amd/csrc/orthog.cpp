// orthog.cpp -- orthogonalisation: CholQR, SVQB and the orthog driver, with
// the k x k factorisations on the host in FP64.
//
// Device work per CholQR: one Gram pass (V'V) and one in-place pass
// (V <- V*R^-1); the factorisation of the k x k Gram matrix is done in double
// precision on the host whatever the scalar type of V.
//
// Reference behaviour followed:
//   sdMat_cholesky      src/core/phist_sdFact_kernels_def.hpp:27-130 -- rank-
//                       revealing Cholesky with diagonal pivoting: pick the
//                       largest remaining diagonal, stop when it drops below
//                       rankTol * max|diag|; M = R'R with R(:,perm) upper
//                       triangular
//   backward/forwardSubst  :133-210 -- permuted triangular solves, columns past
//                       the rank set to zero
//   sdMat_qb            :211-330 -- SVQB factor: D^-1/2 scaling, symmetric
//                       eigendecomposition, eigenvalues below rankTol*max
//                       dropped, eigenvector order reversed
//   chol_QR(p)          src/core/phist_chol_QR_def.hpp:18-143
//   svqb                src/core/phist_svqb_def.hpp:18-39
//   orthog              src/core/phist_orthog_def.hpp:17-240 and
//                       phist_orthogrr_def.hpp:83-336 (block CGS + CholQR
//                       sweeps until ||W'V|| and ||V'V - I|| reach eps)
// The symmetric eigendecomposition is a cyclic Jacobi iteration (the reference
// calls LAPACK dsyevd); eigenvectors are unique only up to sign and rotation
// inside degenerate eigenspaces, so SVQB outputs are checked by invariants.
#include "objects.hpp"
#include <algorithm>
#include <cmath>
#include <limits>

namespace phx {
namespace dense {

// column-major n x n, lda
void cholesky(double* a, int n, int lda, int* perm, int* rank, double rankTol) {
  std::vector<int> p(n);
  std::vector<double> L((size_t)n * n, 0.0), d(n);
  for (int i = 0; i < n; i++) p[i] = i;
  double dmax = 0.0;
  for (int i = 0; i < n; i++) {
    d[i] = a[(size_t)i * lda + i];
    dmax = std::max(dmax, std::fabs(d[i]));
  }
  if (dmax <= rankTol) dmax = 1.0;  // all diagonals tiny: absolute criterion
  int r = 0;
  while (r < n) {
    double rest = 0.0;
    for (int i = r; i < n; i++) rest = std::max(rest, std::fabs(d[p[i]]));
    if (rest <= rankTol * dmax) break;
    int piv = r;
    for (int j = r + 1; j < n; j++)
      if (std::fabs(d[p[j]]) > std::fabs(d[p[piv]])) piv = j;
    std::swap(p[piv], p[r]);
    const int pr = p[r];
    const double lrr = std::sqrt(d[pr]);
    L[(size_t)pr * n + r] = lrr;
    const double inv = 1.0 / lrr;
    for (int i = r + 1; i < n; i++) {
      const int pi = p[i];
      double s = a[(size_t)pi * lda + pr];
      for (int j = 0; j < r; j++) s -= L[(size_t)pr * n + j] * L[(size_t)pi * n + j];
      s *= inv;
      L[(size_t)pi * n + r] = s;
      d[pi] -= s * s;
    }
    r++;
  }
  *rank = r;
  for (int i = 0; i < n; i++) {
    perm[i] = p[i];
    for (int j = 0; j < n; j++) a[(size_t)i * lda + j] = L[(size_t)i * n + j];
  }
}

void backward_subst(const double* R, int n, int ldr, const int* p, int rank, double* x, int k, int ldx) {
  std::vector<double> nx(n);
  for (int l = 0; l < k; l++) {
    double* xl = x + (size_t)l * ldx;
    for (int i = n - 1; i >= rank; i--) nx[i] = 0.0;
    for (int i = rank - 1; i >= 0; i--) {
      const double rinv = 1.0 / R[(size_t)p[i] * ldr + i];
      for (int j = i + 1; j < n; j++) xl[i] -= R[(size_t)p[j] * ldr + i] * nx[j];
      nx[i] = xl[i] * rinv;
    }
    for (int i = 0; i < rank; i++) xl[p[i]] = nx[i];
    for (int i = rank; i < n; i++) xl[p[i]] = 0.0;
  }
}

void forward_subst(const double* R, int n, int ldr, const int* p, int rank, double* x, int k, int ldx) {
  std::vector<double> nx(n);
  for (int l = 0; l < k; l++) {
    double* xl = x + (size_t)l * ldx;
    for (int i = 0; i < rank; i++) {
      double s = xl[p[i]];
      for (int j = 0; j < i; j++) s -= R[(size_t)p[i] * ldr + j] * nx[j];
      nx[i] = s / R[(size_t)p[i] * ldr + i];
    }
    for (int i = 0; i < n; i++) xl[i] = i < rank ? nx[i] : xl[i];
    for (int i = rank; i < n; i++) xl[p[i]] = 0.0;
  }
}

// symmetric eigendecomposition by cyclic Jacobi rotations: a (n x n, col-major)
// is overwritten by the eigenvectors (columns), w gets the eigenvalues in
// ascending order
void sym_eig(double* a, int n, int lda, double* w) {
  std::vector<double> A((size_t)n * n), Vv((size_t)n * n, 0.0);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) A[(size_t)j * n + i] = 0.5 * (a[(size_t)j * lda + i] + a[(size_t)i * lda + j]);
  for (int i = 0; i < n; i++) Vv[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        const double v = A[(size_t)j * n + i] * A[(size_t)j * n + i];
        tot += v;
        if (i != j) off += v;
      }
    if (off <= 1e-32 * tot || off == 0.0) break;
    for (int pp = 0; pp < n - 1; pp++)
      for (int qq = pp + 1; qq < n; qq++) {
        const double apq = A[(size_t)qq * n + pp];
        if (apq == 0.0) continue;
        const double app = A[(size_t)pp * n + pp], aqq = A[(size_t)qq * n + qq];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; k++) {  // columns p, q
          const double akp = A[(size_t)pp * n + k], akq = A[(size_t)qq * n + k];
          A[(size_t)pp * n + k] = c * akp - s * akq;
          A[(size_t)qq * n + k] = s * akp + c * akq;
        }
        for (int k = 0; k < n; k++) {  // rows p, q
          const double apk = A[(size_t)k * n + pp], aqk = A[(size_t)k * n + qq];
          A[(size_t)k * n + pp] = c * apk - s * aqk;
          A[(size_t)k * n + qq] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; k++) {
          const double vkp = Vv[(size_t)pp * n + k], vkq = Vv[(size_t)qq * n + k];
          Vv[(size_t)pp * n + k] = c * vkp - s * vkq;
          Vv[(size_t)qq * n + k] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int x, int y) { return A[(size_t)x * n + x] < A[(size_t)y * n + y]; });
  for (int j = 0; j < n; j++) {
    w[j] = A[(size_t)order[j] * n + order[j]];
    for (int i = 0; i < n; i++) a[(size_t)j * lda + i] = Vv[(size_t)order[j] * n + i];
  }
}

// Symmetric eigendecomposition by Householder reduction to tridiagonal form
// and the implicit QL iteration with Wilkinson shifts (Golub & Van Loan,
// Matrix Computations, 4th ed., algorithms 8.3.1 and 8.3.3; the QL form runs
// the same bulge chase from the bottom of the matrix upwards).  About 9 n^3
// flops against 3 n^3 per sweep of the Jacobi iteration above; used for the
// Rayleigh-Ritz problems of the eigensolver, where absolute accuracy is what
// counts.  Same interface as sym_eig.
void sym_eig_tridiag(double* a, int n, int lda, double* w) {
  if (n <= 0) return;
  std::vector<double> A((size_t)n * n), Q((size_t)n * n, 0.0), d(n), e(n, 0.0);
  auto at = [&](int i, int j) -> double& { return A[(size_t)j * n + i]; };
  auto qt = [&](int i, int j) -> double& { return Q[(size_t)j * n + i]; };
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) at(i, j) = 0.5 * (a[(size_t)j * lda + i] + a[(size_t)i * lda + j]);
  // ---- Householder: A <- H_k A H_k, H_k = I - beta v v' acting on rows/cols k+1..n-1 ----
  std::vector<double> v(n), p(n), wv(n), betas(n, 0.0);
  std::vector<double> V((size_t)n * n, 0.0);  // v of step k in column k (rows k+1..)
  for (int k = 0; k + 2 < n; k++) {
    const int m = n - k - 1;  // length of the column part below the diagonal
    double sig = 0.0;
    for (int i = 1; i < m; i++) sig += at(k + 1 + i, k) * at(k + 1 + i, k);
    const double x0 = at(k + 1, k);
    if (sig == 0.0) {  // already tridiagonal in this column
      e[k] = x0;
      continue;
    }
    const double mu = std::sqrt(x0 * x0 + sig);
    const double alpha = x0 <= 0.0 ? mu : -mu;  // H x = alpha e1, no cancellation in v0
    const double v0 = x0 - alpha;
    v[0] = 1.0;
    for (int i = 1; i < m; i++) v[i] = at(k + 1 + i, k) / v0;
    const double beta = -v0 / alpha;  // 2 / (v'v) for this scaling
    betas[k] = beta;
    for (int i = 0; i < m; i++) V[(size_t)k * n + k + 1 + i] = v[i];
    // p = beta * A22 v ; wv = p - (beta/2)(p'v) v ; A22 -= v wv' + wv v'
    for (int i = 0; i < m; i++) {
      double sacc = 0.0;
      for (int j = 0; j < m; j++) sacc += at(k + 1 + i, k + 1 + j) * v[j];
      p[i] = beta * sacc;
    }
    double pv = 0.0;
    for (int i = 0; i < m; i++) pv += p[i] * v[i];
    for (int i = 0; i < m; i++) wv[i] = p[i] - 0.5 * beta * pv * v[i];
    for (int j = 0; j < m; j++)
      for (int i = 0; i < m; i++) at(k + 1 + i, k + 1 + j) -= v[i] * wv[j] + wv[i] * v[j];
    e[k] = alpha;
  }
  if (n >= 2) e[n - 2] = at(n - 1, n - 2);
  for (int i = 0; i < n; i++) d[i] = at(i, i);
  e[n - 1] = 0.0;
  // ---- Q = H_0 H_1 ... H_{n-3}, accumulated backwards ----
  for (int i = 0; i < n; i++) qt(i, i) = 1.0;
  for (int k = n - 3; k >= 0; k--) {
    const double beta = betas[k];
    if (beta == 0.0) continue;
    const int m = n - k - 1;
    for (int j = k + 1; j < n; j++) {  // columns of the trailing block
      double sacc = 0.0;
      for (int i = 0; i < m; i++) sacc += V[(size_t)k * n + k + 1 + i] * qt(k + 1 + i, j);
      sacc *= beta;
      for (int i = 0; i < m; i++) qt(k + 1 + i, j) -= sacc * V[(size_t)k * n + k + 1 + i];
    }
  }
  // ---- implicit QL on (d, e): e[i] couples i and i+1 ----
  const double eps = std::numeric_limits<double>::epsilon();
  for (int l = 0; l < n; l++) {
    for (int iter = 0; iter < 60; iter++) {
      int m = l;
      for (; m < n - 1; m++) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= eps * dd) break;
      }
      if (m == l) break;
      // Wilkinson-type shift from the leading 2x2 block at l
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? r : -r));
      double s = 1.0, c = 1.0, pacc = 0.0;
      bool early = false;
      for (int i = m - 1; i >= l; i--) {
        const double f = s * e[i], b = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {  // split: recover and restart this eigenvalue
          d[i + 1] -= pacc;
          e[m] = 0.0;
          early = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - pacc;
        r = (d[i] - g) * s + 2.0 * c * b;
        pacc = s * r;
        d[i + 1] = g + pacc;
        g = c * r - b;
        for (int k = 0; k < n; k++) {  // rotate columns i, i+1 of Q
          const double qi = qt(k, i), qi1 = qt(k, i + 1);
          qt(k, i + 1) = s * qi + c * qi1;
          qt(k, i) = c * qi - s * qi1;
        }
      }
      if (early) continue;
      d[l] -= pacc;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int x, int y) { return d[x] < d[y]; });
  for (int j = 0; j < n; j++) {
    w[j] = d[order[j]];
    for (int i = 0; i < n; i++) a[(size_t)j * lda + i] = qt(i, order[j]);
  }
}

// SVQB factor of the Gram matrix a (n x n): on exit a = B with V*B
// orthonormal (up to rank), bi (optional) its left pseudo-inverse, D the
// column norms sqrt(diag(a))
void qb(double* a, double* bi, double* D, int n, int lda, int* rank, double rankTol) {
  std::vector<double> d(n), di(n), w(n), wi(n);
  for (int i = 0; i < n; i++) {
    const double aii = a[(size_t)i * lda + i];
    if (aii <= rankTol) {
      d[i] = di[i] = 0.0;
    } else {
      d[i] = std::sqrt(aii);
      di[i] = 1.0 / d[i];
    }
    if (D) D[i] = d[i];
  }
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) a[(size_t)j * lda + i] *= di[i] * di[j];
  sym_eig(a, n, lda, w.data());
  *rank = n;
  const double emax = n ? std::fabs(w[n - 1]) : 0.0;
  if (emax <= rankTol) {
    for (int i = 0; i < n; i++) w[i] = wi[i] = 0.0;
    *rank = 0;
  } else {
    for (int i = 0; i < n; i++) {
      if (w[i] <= rankTol * emax) {
        (*rank)--;
        w[i] = wi[i] = 0.0;
      } else {
        w[i] = std::sqrt(w[i]);
        wi[i] = 1.0 / w[i];
      }
    }
  }
  if (bi)
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) bi[(size_t)i * lda + j] = a[(size_t)j * lda + i] * d[i] * w[j];
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) a[(size_t)j * lda + i] *= di[i] * wi[j];
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n / 2; j++) {
      const int k = n - j - 1;
      std::swap(a[(size_t)j * lda + i], a[(size_t)k * lda + i]);
      if (bi) std::swap(bi[(size_t)i * lda + j], bi[(size_t)i * lda + k]);
    }
}

}  // namespace dense

namespace api {

namespace {
// host copy of an sdMat as a packed double matrix and back
template <typename T>
std::vector<double> to_dbl(const SdMat<T>* M) {
  std::vector<double> a((size_t)M->nrows * M->ncols);
  for (int j = 0; j < M->ncols; j++)
    for (int i = 0; i < M->nrows; i++) a[(size_t)j * M->nrows + i] = (double)M->h(i, j);
  return a;
}
template <typename T>
void from_dbl(SdMat<T>* M, const std::vector<double>& a) {
  for (int j = 0; j < M->ncols; j++)
    for (int i = 0; i < M->nrows; i++) M->h(i, j) = (T)a[(size_t)j * M->nrows + i];
  sd_to_device(M);
}
template <typename T>
double normF(const SdMat<T>* M) {
  double s = 0.0;
  for (int j = 0; j < M->ncols; j++)
    for (int i = 0; i < M->nrows; i++) s += (double)M->h(i, j) * (double)M->h(i, j);
  return std::sqrt(s);
}
template <typename T>
T rank_tol(bool robust) {
  return robust ? (sizeof(T) == 8 ? (T)1.0e-25 : (T)1.0e-12) : (T)(10.0 * std::numeric_limits<T>::epsilon());
}
template <typename T>
struct SdOwner {
  SdMat<T>* p;
  explicit SdOwner(SdMat<T>* q) : p(q) {}
  ~SdOwner() {
    try {
      sdMat_delete(p);
    } catch (...) {
    }
  }
};
}  // namespace

template <typename T>
void sdMat_cholesky(SdMat<T>* M, int* perm, int* rank, T rankTol) {
  PHX_REQUIRE(M->nrows == M->ncols, PHIST_INVALID_INPUT, "cholesky needs a square sdMat");
  const int n = M->nrows;
  std::vector<double> a = to_dbl(M);
  dense::cholesky(a.data(), n, n, perm, rank, (double)rankTol);
  from_dbl(M, a);
}

template <typename T>
void sdMat_subst(const SdMat<T>* R, const int* perm, int rank, SdMat<T>* X, bool backward) {
  const int n = R->nrows;
  PHX_REQUIRE(R->ncols == n && X->nrows == n, PHIST_INVALID_INPUT, "triangular solve shapes do not match");
  std::vector<double> r = to_dbl(R), x = to_dbl(X);
  if (backward)
    dense::backward_subst(r.data(), n, n, perm, rank, x.data(), X->ncols, n);
  else
    dense::forward_subst(r.data(), n, n, perm, rank, x.data(), X->ncols, n);
  from_dbl(X, x);
}

template <typename T>
void sdMat_qb(SdMat<T>* B, SdMat<T>* B_1, T* nrms, int* rank, T rankTol) {
  const int n = B->nrows;
  PHX_REQUIRE(B->ncols == n && (!B_1 || (B_1->nrows == n && B_1->ncols == n)), PHIST_INVALID_INPUT,
              "qb needs square sdMats");
  std::vector<double> a = to_dbl(B), bi((size_t)n * n), D(n);
  dense::qb(a.data(), B_1 ? bi.data() : nullptr, D.data(), n, n, rank, (double)rankTol);
  if (nrms)
    for (int i = 0; i < n; i++) nrms[i] = (T)D[i];
  from_dbl(B, a);
  if (B_1) from_dbl(B_1, bi);
}

template <typename T>
int chol_QRp(Mvec<T>* V, SdMat<T>* R, int* perm, bool robust) {
  const int m = V->nvec;
  PHX_REQUIRE(R->nrows == m && R->ncols == m, PHIST_INVALID_INPUT, "R must be %d x %d", m, m);
  const T rankTol = rank_tol<T>(robust);
  for (int i = 0; i < m; i++) perm[i] = i;
  if (m == 1) {
    T nrm;
    blas_dot<T>(V, V, &nrm);
    nrm = std::sqrt(nrm);
    int rank = 1;
    if (nrm < rankTol) {
      mvec_random<T>(V);
      blas_dot<T>(V, V, &nrm);
      nrm = std::sqrt(nrm);
      T s = T(1) / nrm;
      blas_scale_cols<T>(V, &s, true);
      sdMat_put_value<T>(R, T(0));
      rank = 0;
    } else {
      T s = T(1) / nrm;
      blas_scale_cols<T>(V, &s, true);
      sdMat_put_value<T>(R, nrm);
    }
    return 1 - rank;
  }
  gemm_gram<T>(T(1), V, V, T(0), R);
  int rank = 0;
  sdMat_cholesky<T>(R, perm, &rank, rankTol);
  SdMat<T>* R_1 = sdMat_create<T>(m, m);
  SdOwner<T> own(R_1);
  sdMat_identity<T>(R_1);
  sdMat_subst<T>(R, perm, rank, R_1, true);
  gemm_times_sd_inplace<T>(V, R_1);
  if (rank < m) {
    // randomise the null space and orthogonalise it against the range
    Mvec<T>* Vorth = nullptr;
    Mvec<T>* Vn = nullptr;
    if (rank) mvec_view_block<T>(V, &Vorth, 0, rank - 1);
    mvec_view_block<T>(V, &Vn, rank, m - 1);
    mvec_random<T>(Vn);
    SdMat<T>* R1 = sdMat_create<T>(m - rank, m - rank);
    SdOwner<T> o1(R1);
    SdMat<T>* R2 = rank ? sdMat_create<T>(rank, m - rank) : nullptr;
    SdOwner<T> o2(R2);
    std::vector<int> p2(m - rank);
    for (int sweep = 0; sweep < 2; sweep++) {
      int def = chol_QRp<T>(Vn, R1, p2.data(), robust);
      PHX_REQUIRE(def == 0, PHIST_FUNCTIONAL_ERROR, "catastrophic breakdown in chol_QR");
      if (rank) {
        gemm_gram<T>(T(1), Vorth, Vn, T(0), R2);
        gemm_times_sd<T>(T(-1), Vorth, R2, T(1), Vn);
      }
    }
    if (Vorth) mvec_delete<T>(Vorth);
    mvec_delete<T>(Vn);
  }
  return m - rank;
}

template <typename T>
int svqb(Mvec<T>* V, SdMat<T>* B, T* e, bool robust) {
  const int m = V->nvec;
  PHX_REQUIRE(B->nrows == m && B->ncols == m, PHIST_INVALID_INPUT, "B must be %d x %d", m, m);
  const T rankTol = rank_tol<T>(robust);
  gemm_gram<T>(T(1), V, V, T(0), B);
  int rank = 0;
  sdMat_qb<T>(B, nullptr, e, &rank, rankTol);
  gemm_times_sd_inplace<T>(V, B);
  return m - rank;
}

// orthogrr sweeps (no B operator): Wq is orthogonalised against the
// orthonormal block Vo; returns the null-space dimension of [Vo Wq].
template <typename T>
int orthogrr(const Mvec<T>* Vo, Mvec<T>* Wq, SdMat<T>* R2, SdMat<T>* R1, SdMat<T>* WtW, double desiredEps, int maxIter,
             T rankTol, bool robust) {
  const int m = Wq->nvec;
  const int k = Vo ? Vo->nvec : 0;
  SdMat<T>* R = sdMat_create<T>(m, m);
  SdOwner<T> oR(R);
  SdMat<T>* R1t = sdMat_create<T>(m, m);
  SdOwner<T> oR1t(R1t);
  SdMat<T>* Ri = sdMat_create<T>(m, m);
  SdOwner<T> oRi(Ri);
  SdMat<T>* VtW = k ? sdMat_create<T>(k, m) : nullptr;
  SdOwner<T> oVtW(VtW);
  SdMat<T>* Iden = sdMat_create<T>(m, m);
  SdOwner<T> oI(Iden);
  std::vector<int> perm(m);

  auto err_to_identity = [&](const SdMat<T>* G, int rk) {  // ||G - I_rk||_F
    double s = 0.0;
    for (int j = 0; j < m; j++)
      for (int i = 0; i < m; i++) {
        const double v = (double)G->h(i, j) - ((i == j && i < rk) ? 1.0 : 0.0);
        s += v * v;
      }
    return std::sqrt(s);
  };
  auto cholrr = [&](int* rk) {  // R <- chol(R) (permuted), Ri <- R^-1
    sdMat_cholesky<T>(R, perm.data(), rk, rankTol);
    sdMat_identity<T>(Ri);
    sdMat_subst<T>(R, perm.data(), *rk, Ri, true);
  };

  int rank = m;
  double VtV_err = err_to_identity(WtW, m);
  if (R1) sdMat_identity<T>(R1);
  if (R2) sdMat_put_value<T>(R2, T(0));
  int iter = 0;
  bool WtW_updated = true;
  for (; iter < maxIter; iter++) {
    int r0 = 0;
    sdMat_add<T>(T(1), WtW, false, T(0), R);
    cholrr(&r0);
    rank = std::min(rank, r0);
    if (k == 0) {
      iter++;
      break;
    }
    // Wq <- Wq*R^-1 and VtW = Vo'Wq in one call
    gemm_fused_mvsdi_mvTmv<T>(T(1), Vo, Wq, Ri, T(0), VtW);
    if (R1) {
      sdMat_add<T>(T(1), R1, false, T(0), R1t);
      sdMat_gemm<T>(T(1), R, false, R1t, false, T(0), R1);
    }
    VtV_err = robust ? std::max(VtV_err * std::numeric_limits<T>::epsilon(), (double)std::numeric_limits<T>::epsilon())
                     : std::max(VtV_err * std::sqrt((double)std::numeric_limits<T>::epsilon()),
                                (double)std::numeric_limits<T>::epsilon());
    double VtW_err = normF(VtW);
    VtV_err += VtW_err;
    if (VtW_err <= desiredEps && VtW_err * VtV_err <= desiredEps * desiredEps) {
      WtW_updated = false;
      break;
    }
    if (R2) sdMat_gemm<T>(T(1), VtW, false, R, false, T(1), R2);
    // Wq <- Wq - Vo*VtW, WtW = Wq'Wq in one call
    gemm_fused_mvsd_mvTmv<T>(T(-1), Vo, VtW, T(1), Wq, WtW);
    int r1 = 0;
    sdMat_add<T>(T(1), WtW, false, T(0), R);
    cholrr(&r1);
    rank = std::min(rank, r1);
    VtW_err = normF(Ri) * std::numeric_limits<T>::epsilon();
    VtV_err = err_to_identity(WtW, rank);
    if (VtW_err <= desiredEps) {
      iter++;
      break;
    }
  }
  for (; iter <= maxIter; iter++) {
    if (VtV_err <= desiredEps) break;
    if (!WtW_updated) {
      gemm_gram<T>(T(1), Wq, Wq, T(0), WtW);
      WtW_updated = true;
    } else {
      gemm_fused_mvsdi_mvTmv<T>(T(1), Wq, Wq, Ri, T(0), WtW);
      if (R1) {
        sdMat_add<T>(T(1), R1, false, T(0), R1t);
        sdMat_gemm<T>(T(1), R, false, R1t, false, T(0), R1);
      }
    }
    int r2 = 0;
    sdMat_add<T>(T(1), WtW, false, T(0), R);
    cholrr(&r2);
    rank = std::min(rank, r2);
    VtV_err = err_to_identity(WtW, rank);
  }
  return m - rank;
}

template <typename T>
int orthog(const Mvec<T>* V, Mvec<T>* W, SdMat<T>* R1, SdMat<T>* R2, int numSweeps, int* rankVW, int flags) {
  const bool robust = flags & PHIST_ROBUST_REDUCTIONS;
  const bool randomize = flags & 131072;  // PHIST_ORTHOG_RANDOMIZE_NULLSPACE (phist_core_flags.h:37)
  PHX_REQUIRE(!(flags & 262144), PHIST_NOT_IMPLEMENTED, "PHIST_ORTHOG_TRIANGULAR_R1 is not supported");
  const int k = W->nvec;
  const int m = V ? V->nvec : 0;
  if (V) PHX_REQUIRE(V->nrows == W->nrows, PHIST_INVALID_INPUT, "V and W have different lengths");
  PHX_REQUIRE(!R1 || (R1->nrows == k && R1->ncols == k), PHIST_INVALID_INPUT, "R1 must be %d x %d", k, k);
  PHX_REQUIRE(!R2 || !V || (R2->nrows == m && R2->ncols == k), PHIST_INVALID_INPUT, "R2 must be %d x %d", m, k);
  const T rankTol = randomize ? rank_tol<T>(robust) : T(0);
  double eps = std::numeric_limits<T>::epsilon();
  if (robust) eps *= std::sqrt((double)std::numeric_limits<T>::epsilon());
  SdMat<T>* WtW = sdMat_create<T>(k, k);
  SdOwner<T> oW(WtW);
  gemm_gram<T>(T(1), W, W, T(0), WtW);
  int dim0 = orthogrr<T>(V, W, R2, R1, WtW, eps, numSweeps, rankTol, robust);
  *rankVW = m + k - dim0;
  int attempts = 0;
  const int max_attempts = 5;
  if (randomize && dim0 > 0) {
    SdMat<T>* R1p = sdMat_create<T>(k, k);
    SdOwner<T> o1(R1p);
    SdMat<T>* R1pp = sdMat_create<T>(k, k);
    SdOwner<T> o2(R1pp);
    SdMat<T>* R2p = m ? sdMat_create<T>(m, k) : nullptr;
    SdOwner<T> o3(R2p);
    while (dim0 > 0 && attempts++ < max_attempts) {
      const int rankW = k - dim0;
      Mvec<T>* Wr = nullptr;
      mvec_view_block<T>(W, &Wr, rankW, k - 1);
      mvec_random<T>(Wr);
      mvec_delete<T>(Wr);
      gemm_gram<T>(T(1), W, W, T(0), WtW);
      dim0 = orthogrr<T>(V, W, R2p, R1p, WtW, eps, numSweeps, rankTol, robust);
      if (m > 0 && R2 && R1) {
        for (int j = rankW; j < k; j++)
          for (int i = 0; i < m; i++) R2p->h(i, j) = T(0);
        sd_to_device(R2p);
        sdMat_gemm<T>(T(1), R2p, false, R1, false, T(1), R2);
      }
      if (R1) {
        sdMat_add<T>(T(1), R1, false, T(0), R1pp);
        sdMat_gemm<T>(T(1), R1p, false, R1pp, false, T(0), R1);
        for (int j = rankW; j < k; j++)
          for (int i = 0; i < k; i++) R1->h(i, j) = T(0);
        sd_to_device(R1);
      }
    }
  }
  if (attempts == max_attempts) return -8;
  return *rankVW < m + k ? 1 : 0;
}

#define PHX_INST(T)                                                                                           \
  template void sdMat_cholesky<T>(SdMat<T>*, int*, int*, T);                                                  \
  template void sdMat_subst<T>(const SdMat<T>*, const int*, int, SdMat<T>*, bool);                            \
  template void sdMat_qb<T>(SdMat<T>*, SdMat<T>*, T*, int*, T);                                               \
  template int chol_QRp<T>(Mvec<T>*, SdMat<T>*, int*, bool);                                                  \
  template int svqb<T>(Mvec<T>*, SdMat<T>*, T*, bool);                                                        \
  template int orthog<T>(const Mvec<T>*, Mvec<T>*, SdMat<T>*, SdMat<T>*, int, int*, int);
PHX_INST(double)
PHX_INST(float)
#undef PHX_INST

}  // namespace api
}  // namespace phx

// host-only check hook: symmetric eigendecomposition (method 0: Jacobi, 1:
// tridiagonal QL), col-major a overwritten by the eigenvectors
extern "C" void phist_amd_dense_sym_eig(double* a, int n, int lda, double* w, int method, int* iflag) {
  phx::guard(iflag, __func__, [&] {
    PHX_REQUIRE(n >= 0 && lda >= std::max(1, n), PHIST_INVALID_INPUT, "bad size");
    if (method == 1)
      phx::dense::sym_eig_tridiag(a, n, lda, w);
    else
      phx::dense::sym_eig(a, n, lda, w);
    *iflag = 0;
  });
}
