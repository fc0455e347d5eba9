// gmres.cpp -- restarted GMRES(m) on k shifted systems at once, for general
// (nonsymmetric) matrices of this library.
//
// The systems (A + s_j I) x_j = b_j share one launch per step but keep their
// own Krylov bases and Hessenberg matrices, as in PHIST's blocked GMRES
// (src/krylov/phist_blockedgmres_def.hpp:453-1010, restartedBlockedGMRES
// :1029-1200): the basis vectors of all systems for step l are the k
// columns of block V_l, stored side by side in one block of (m+1)*k columns
// so one pass reads every basis vector of a row.  Per step:
//   w = (A + diag(s)) v_i                               one shifted SpMV
//   h(0:i) = V(0:i)' w, w -= V(0:i) h  (twice, CGS2)     Gram + row mix on the
//                                                        whole basis; only the
//                                                        blocks' diagonals are
//                                                        used (the other
//                                                        entries ride along in
//                                                        the memory-bound pass)
//   h(i+1) = ||w||, v_{i+1} = w / h(i+1)
//   Givens rotations of every system's Hessenberg column on the host; the
//   implicit residual |g_{i+1}| decides convergence.
// Convergence as the reference (:600-611, :968-988): ||r|| / ||r0|| < tol or
// ||r|| < 1000 eps (absolute); a system stops when its iterations reach
// maxIter.  At the end of a cycle x_j += V_j y_j (y from the triangular
// solve); a restart recomputes the true residual b - (A + sI) x.
#include "solvers.hpp"
#include <cmath>
#include <limits>

namespace phx {

template <typename T, typename TI>
GmresOut gmres_solve(const Operator<TI>& op, const std::vector<T>& shifts, const Mvec<T>* B, Mvec<T>* X,
                     const std::vector<double>& tol, int maxIter, int m) {
  const int k = B->nvec;
  const Map* map = B->map;
  PHX_REQUIRE(m >= 1 && m <= 256, PHIST_INVALID_INPUT, "GMRES basis size %d (1..256)", m);
  PHX_REQUIRE((int64_t)(m + 1) * k <= 4096, PHIST_INVALID_INPUT, "GMRES basis (m+1)*k = %d columns too wide",
              (m + 1) * k);
  const double eps = std::numeric_limits<double>::epsilon();
  // working precision copies of b and x
  MvBox<TI> b(map, k), x(map, k), w(map, k), Vb(map, (m + 1) * k);
  copy_any<T, TI>(B, b.get());
  copy_any<T, TI>(X, x.get());
  std::vector<TI> s(shifts.begin(), shifts.end());

  GmresOut out;
  out.iters.assign(k, 0);
  out.relres.assign(k, 1.0);
  std::vector<double> normR0(k, -1.0), normR(k, 0.0);
  std::vector<int> status(k, 1);  // 1 iterating, 0 converged, 3 failed
  // Hessenberg (column-major (m+1) x m per system), rotations, rhs
  std::vector<double> H((size_t)k * (m + 1) * m), cs((size_t)k * m), sn((size_t)k * m), g((size_t)k * (m + 1));
  auto h = [&](int j, int r, int c) -> double& { return H[((size_t)j * m + c) * (m + 1) + r]; };

  auto residual = [&](std::vector<double>& nrm) {  // w = b - (A + sI) x, norms
    op.apply_shifted(TI(-1), s.data(), x.get(), TI(0), w.get(), nullptr);
    std::vector<TI> one(k, TI(1));
    blas_axpby<TI>(one.data(), true, b.get(), TI(1), w.get());
    std::vector<TI> d(k);
    blas_dot<TI>(w.get(), w.get(), d.data());
    for (int j = 0; j < k; j++) nrm[j] = std::sqrt((double)d[j]);
  };

  std::vector<double> rn(k);
  residual(rn);
  for (;;) {
    // ---- (re)start: v_0 = r / ||r|| for the systems still iterating ----
    std::vector<TI> sc(k);
    for (int j = 0; j < k; j++) {
      normR[j] = rn[j];
      if (normR0[j] < 0.0) normR0[j] = rn[j];
      if (status[j] == 1) {
        const double rel = normR0[j] > 0.0 ? rn[j] / normR0[j] : 0.0;
        out.relres[j] = rel;
        if (rn[j] < 1000 * eps || rel < tol[j]) status[j] = 0;
        else if (out.iters[j] >= maxIter) status[j] = 3;
      }
      sc[j] = (status[j] == 1 && rn[j] > 0.0) ? TI(1.0 / rn[j]) : TI(0);
      for (int r = 0; r <= m; r++) g[(size_t)j * (m + 1) + r] = 0.0;
      g[(size_t)j * (m + 1)] = status[j] == 1 ? rn[j] : 0.0;
    }
    bool any = false;
    for (int j = 0; j < k; j++) any = any || status[j] == 1;
    if (!any) break;
    {
      View<TI> v0(Vb.get(), 0, k - 1);
      blas_axpby<TI>(sc.data(), false, w.get(), TI(0), v0.get());
    }
    std::vector<int> active(k), steps(k, 0);
    for (int j = 0; j < k; j++) active[j] = status[j] == 1;
    int i = 0;
    for (; i < m; i++) {
      bool live = false;
      for (int j = 0; j < k; j++) live = live || active[j];
      if (!live) break;
      View<TI> vi(Vb.get(), i * k, (i + 1) * k - 1);
      View<TI> Vi(Vb.get(), 0, (i + 1) * k - 1);
      op.apply_shifted(TI(1), s.data(), vi.get(), TI(0), w.get(), nullptr);
      std::vector<double> hc((size_t)(i + 1) * k, 0.0);
      for (int pass = 0; pass < 2; pass++) {  // classical Gram-Schmidt, twice
        SdBox<TI> G((i + 1) * k, k);
        gemm_gram<TI>(TI(1), Vi.get(), w.get(), TI(0), G.get());
        for (int j = 0; j < k; j++) {
          for (int l = 0; l <= i; l++) {
            const TI c = active[j] ? G->h(l * k + j, j) : TI(0);
            hc[(size_t)l * k + j] += (double)c;
          }
        }
        SdBox<TI> C((i + 1) * k, k);
        for (int j = 0; j < k; j++)
          for (int l = 0; l <= i; l++) C->h(l * k + j, j) = active[j] ? G->h(l * k + j, j) : TI(0);
        sd_to_device(C.get());
        gemm_times_sd<TI>(TI(-1), Vi.get(), C.get(), TI(1), w.get());
      }
      std::vector<TI> d(k), inv(k);
      blas_dot<TI>(w.get(), w.get(), d.data());
      for (int j = 0; j < k; j++) {
        const double hn = std::sqrt((double)d[j]);
        inv[j] = (active[j] && hn > 0.0 && std::isfinite(hn)) ? TI(1.0 / hn) : TI(0);
        if (!active[j]) continue;
        if (!std::isfinite(hn)) {  // broken down: keep the last solution of this cycle
          active[j] = 0;
          status[j] = 3;
          continue;
        }
        // Hessenberg column i: h(0:i, i) and h(i+1, i); earlier rotations, then a new one
        for (int l = 0; l <= i; l++) h(j, l, i) = hc[(size_t)l * k + j];
        h(j, i + 1, i) = hn;
        for (int l = 0; l < i; l++) {
          const double c = cs[(size_t)j * m + l], sv = sn[(size_t)j * m + l];
          const double a = h(j, l, i), bb = h(j, l + 1, i);
          h(j, l, i) = c * a + sv * bb;
          h(j, l + 1, i) = -sv * a + c * bb;
        }
        const double a = h(j, i, i), bb = h(j, i + 1, i);
        const double r = std::hypot(a, bb);
        const double c = r > 0.0 ? a / r : 1.0, sv = r > 0.0 ? bb / r : 0.0;
        cs[(size_t)j * m + i] = c;
        sn[(size_t)j * m + i] = sv;
        h(j, i, i) = r;
        h(j, i + 1, i) = 0.0;
        double* gj = &g[(size_t)j * (m + 1)];
        gj[i + 1] = -sv * gj[i];
        gj[i] = c * gj[i];
        steps[j] = i + 1;
        out.iters[j]++;
        normR[j] = std::fabs(gj[i + 1]);
        const double rel = normR[j] / normR0[j];
        out.relres[j] = rel;
        if (normR[j] < 1000 * eps || rel < tol[j] || hn == 0.0) {
          active[j] = 0;
          status[j] = 0;
        } else if (out.iters[j] >= maxIter) {
          active[j] = 0;
          status[j] = 3;
        }
      }
      View<TI> vn(Vb.get(), (i + 1) * k, (i + 2) * k - 1);
      blas_axpby<TI>(inv.data(), false, w.get(), TI(0), vn.get());
      for (int j = 0; j < k; j++)  // a non-finite column must not reach the other systems through 0 * NaN
        if (!std::isfinite((double)d[j])) {
          View<TI> col(vn.get(), j, j);
          blas_put_value<TI>(col.get(), TI(0));
        }
    }
    // ---- x += V y per system (y from the triangular systems R y = g) ----
    int L = 0;
    for (int j = 0; j < k; j++) L = std::max(L, steps[j]);
    if (L > 0) {
      SdBox<TI> Y(L * k, k);
      for (int j = 0; j < k; j++) {
        const int n = steps[j];
        std::vector<double> y(n);
        for (int r = n - 1; r >= 0; r--) {
          double acc = g[(size_t)j * (m + 1) + r];
          for (int c = r + 1; c < n; c++) acc -= h(j, r, c) * y[c];
          y[r] = h(j, r, r) != 0.0 ? acc / h(j, r, r) : 0.0;
        }
        for (int r = 0; r < n; r++) Y->h(r * k + j, j) = (TI)y[r];
      }
      sd_to_device(Y.get());
      View<TI> VL(Vb.get(), 0, L * k - 1);
      gemm_times_sd<TI>(TI(1), VL.get(), Y.get(), TI(1), x.get());
    }
    // ---- restart with the true residual (or finish) ----
    bool again = false;
    for (int j = 0; j < k; j++) again = again || status[j] == 1;
    residual(rn);
    if (!again) {
      for (int j = 0; j < k; j++)
        if (normR0[j] > 0.0) out.relres[j] = rn[j] / normR0[j];
      break;
    }
    // a system the implicit residual declared converged keeps that status
    for (int j = 0; j < k; j++)
      if (status[j] != 1) rn[j] = 0.0;
  }
  copy_any<TI, T>(x.get(), X);
  out.unconverged = 0;
  for (int j = 0; j < k; j++) out.unconverged += status[j] != 0;
  return out;
}

template GmresOut gmres_solve<double, double>(const Operator<double>&, const std::vector<double>&,
                                              const Mvec<double>*, Mvec<double>*, const std::vector<double>&, int,
                                              int);
template GmresOut gmres_solve<double, float>(const Operator<float>&, const std::vector<double>&, const Mvec<double>*,
                                             Mvec<double>*, const std::vector<double>&, int, int);

}  // namespace phx

using namespace phx;

extern "C" {

void phist_amd_DblockedGMRES_shifted(phist_Dconst_sparseMat_ptr A, int nsys, const double shifts[],
                                     phist_Dconst_mvec_ptr B, phist_Dmvec_ptr X, const double tol[], int maxIter,
                                     int restart, int fp32, int nIter[], double relres[], int* iflag) {
  *iflag = 0;
  guard(iflag, "phist_amd_DblockedGMRES_shifted", [&] {
    const auto* M = reinterpret_cast<const SparseMat<double>*>(A);
    const auto* b = reinterpret_cast<const Mvec<double>*>(B);
    auto* x = reinterpret_cast<Mvec<double>*>(X);
    PHX_REQUIRE(M && b && x && shifts && tol && nIter && relres, PHIST_INVALID_INPUT, "null argument");
    PHX_REQUIRE(nsys >= 1 && nsys <= 64 && b->nvec == nsys && x->nvec == nsys, PHIST_INVALID_INPUT,
                "nsys %d must match B and X (at most 64)", nsys);
    PHX_REQUIRE(M->nrows == M->ncols && b->nrows == M->nrows && x->nrows == M->nrows, PHIST_INVALID_INPUT,
                "shapes do not match the square matrix");
    std::vector<double> s(shifts, shifts + nsys), tl(tol, tol + nsys);
    GmresOut r;
    if (fp32) {
      Operator<float> op;
      op.A = sparse_fp32_twin(M);
      r = gmres_solve<double, float>(op, s, b, x, tl, maxIter, restart);
    } else {
      Operator<double> op;
      op.A = M;
      r = gmres_solve<double, double>(op, s, b, x, tl, maxIter, restart);
    }
    for (int j = 0; j < nsys; j++) {
      nIter[j] = r.iters[j];
      relres[j] = r.relres[j];
    }
    *iflag = r.unconverged;
  });
}

}  // extern "C"
