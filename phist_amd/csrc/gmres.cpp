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
//   h(0:i) = V(0:i)' w, w -= V(0:i) h  (twice, CGS2)     three sweeps over the
//                                                        basis (gmres_cgs_dev:
//                                                        dots; update + dots;
//                                                        update + norms), each
//                                                        system meeting only
//                                                        its own columns
//   h(i+1) = ||w||, v_{i+1} = w / h(i+1)
// (for more than 64 systems or m > 32: a Gram and a row mix of the whole
// basis per pass, of which only the blocks' diagonals are used).
//   Givens rotations of every system's Hessenberg column on the device (one
//   thread per system; on the host for the whole-basis passes); the
//   implicit residual |g_{i+1}| decides convergence.
// Convergence as the reference (:600-611, :968-988): ||r|| / ||r0|| < tol or
// ||r|| < 1000 eps (absolute); a system stops when its iterations reach
// maxIter.  At the end of a cycle x_j += V_j y_j (y from the triangular
// solve); a restart recomputes the true residual b - (A + sI) x, and the
// systems that finished leave: the next cycle runs on a block of the others
// only (the basis shrinks with them).  With a projection basis Q the
// operator is P (A + s I), P = I - QQ' (the Jacobi-Davidson correction
// equations; right-hand sides orthogonal to Q keep the iterates so).
#include "solvers.hpp"
#include <cmath>
#include <limits>

namespace phx {

template <typename T, typename TI>
GmresOut gmres_solve(const Operator<TI>& op, const std::vector<T>& shifts, const Mvec<T>* B, Mvec<T>* X,
                     const std::vector<double>& tol, int maxIter, int m, const Mvec<TI>* Q) {
  const int k = B->nvec;
  const Map* map = B->map;
  if (Q && Q->nvec == 0) Q = nullptr;
  // the projected operator P (A + s I) with P = I - QQ' (the correction equations of
  // Jacobi-Davidson, phist_jadaOp_def.hpp:192-268): w -= Q (Q'w) after every product
  auto project = [&](Mvec<TI>* w) {
    if (!Q) return;
    SdBox<TI> C(Q->nvec, w->nvec);
    gemm_gram<TI>(TI(1), Q, w, TI(0), C.get(), false);
    gemm_times_sd<TI>(TI(-1), Q, C.get(), TI(1), w);
  };
  PHX_REQUIRE(m >= 1 && m <= 256, PHIST_INVALID_INPUT, "GMRES basis size %d (1..256)", m);
  PHX_REQUIRE((int64_t)(m + 1) * k <= 4096, PHIST_INVALID_INPUT, "GMRES basis (m+1)*k = %d columns too wide",
              (m + 1) * k);
  const double eps = std::numeric_limits<double>::epsilon();
  // working precision copies of b and x (all systems); the cycles run on
  // compact blocks of the systems still iterating (act: their indices)
  MvBox<TI> b(map, k), x(map, k);
  copy_any<T, TI>(B, b.get());
  copy_any<T, TI>(X, x.get());

  GmresOut out;
  out.iters.assign(k, 0);
  out.relres.assign(k, 1.0);
  std::vector<double> normR0(k, -1.0);
  std::vector<int> status(k, 1);  // 1 iterating, 0 converged, 3 failed

  std::vector<int> act(k);
  for (int j = 0; j < k; j++) act[j] = j;
  int kk = k;
  auto cols = [&](Mvec<TI>* src, const std::vector<int>& idx, Mvec<TI>* dst) {  // dst(:, a) = src(:, idx[a])
    for (size_t a = 0; a < idx.size(); a++) {
      View<TI> from(src, idx[a], idx[a]), to(dst, (int)a, (int)a);
      blas_copy<TI>(from.get(), to.get());
    }
  };
  std::unique_ptr<MvBox<TI>> bc(new MvBox<TI>(map, k)), xc(new MvBox<TI>(map, k)), wc(new MvBox<TI>(map, k));
  blas_copy<TI>(b.get(), bc->get());
  blas_copy<TI>(x.get(), xc->get());
  std::vector<TI> s(shifts.begin(), shifts.end());

  auto residual = [&](std::vector<double>& nrm) {  // w = b - (A + sI) x on the compact blocks, norms
    op.apply_shifted(TI(-1), s.data(), xc->get(), TI(0), wc->get(), nullptr);
    project(wc->get());
    std::vector<TI> one(kk, TI(1));
    blas_axpby<TI>(one.data(), true, bc->get(), TI(1), wc->get());
    std::vector<TI> d(kk);
    blas_dot<TI>(wc->get(), wc->get(), d.data());
    nrm.resize(kk);
    for (int a = 0; a < kk; a++) nrm[a] = std::sqrt((double)d[a]);
  };

  std::vector<double> rn;
  residual(rn);
  for (;;) {
    // ---- convergence at the true residual, then the systems that go on ----
    std::vector<int> keep;
    for (int a = 0; a < kk; a++) {
      const int j = act[a];
      if (normR0[j] < 0.0) normR0[j] = rn[a];
      const double rel = normR0[j] > 0.0 ? rn[a] / normR0[j] : 0.0;
      out.relres[j] = rel;
      if (status[j] == 1) {
        if (rn[a] < 1000 * eps || rel < tol[j]) status[j] = 0;
        else if (out.iters[j] >= maxIter) status[j] = 3;
      }
      if (status[j] == 1) keep.push_back(a);
    }
    if ((int)keep.size() < kk) {
      // finished systems leave with their solutions; the others move to narrower blocks
      std::vector<int> done;
      for (int a = 0, p = 0; a < kk; a++) {
        if (p < (int)keep.size() && keep[p] == a) {
          p++;
          continue;
        }
        done.push_back(a);
      }
      for (int a : done) {
        View<TI> from(xc->get(), a, a), to(x.get(), act[a], act[a]);
        blas_copy<TI>(from.get(), to.get());
      }
      if (keep.empty()) break;
      const int kn = (int)keep.size();
      std::unique_ptr<MvBox<TI>> bn(new MvBox<TI>(map, kn)), xn(new MvBox<TI>(map, kn)), wn(new MvBox<TI>(map, kn));
      cols(bc->get(), keep, bn->get());
      cols(xc->get(), keep, xn->get());
      cols(wc->get(), keep, wn->get());
      std::vector<int> an(kn);
      std::vector<TI> sn_(kn);
      std::vector<double> rnn(kn);
      for (int a = 0; a < kn; a++) {
        an[a] = act[keep[a]];
        sn_[a] = s[keep[a]];
        rnn[a] = rn[keep[a]];
      }
      bc.swap(bn);
      xc.swap(xn);
      wc.swap(wn);
      act.swap(an);
      s.swap(sn_);
      rn.swap(rnn);
      kk = kn;
    }

    // ---- one GMRES(m) cycle on the kk systems: v_0 = r / ||r|| ----
    // vcur: the current basis vectors as a compact block (row stride kk), the
    // SpMV's input; gathering them from the side-by-side basis (row stride
    // (m+1)*kk) would fetch a cache line per entry for a few bytes of it
    MvBox<TI> w(map, kk), Vb(map, (m + 1) * kk), vcur(map, kk);
    std::vector<double> H((size_t)kk * (m + 1) * m), cs((size_t)kk * m), sn((size_t)kk * m), g((size_t)kk * (m + 1));
    auto h = [&](int a, int r, int c) -> double& { return H[((size_t)a * m + c) * (m + 1) + r]; };
    std::vector<TI> sc(kk);
    for (int a = 0; a < kk; a++) {
      sc[a] = rn[a] > 0.0 ? TI(1.0 / rn[a]) : TI(0);
      g[(size_t)a * (m + 1)] = rn[a];
    }
    {
      View<TI> v0(Vb.get(), 0, kk - 1);
      blas_axpby<TI>(sc.data(), false, wc->get(), TI(0), vcur.get());
      blas_copy<TI>(vcur.get(), v0.get());
    }
    std::vector<int> active(kk, 1), steps(kk, 0);
    // per-system sweeps (each system meets only its own basis): device work for the partial sums, the
    // coefficients of both passes and the squared norms (D1 | D2 | N, read back together), the mask of
    // the systems still iterating and the solution coefficients
    const bool cgs = rt().gmres_cgs && kk <= kGmresCgsMaxKk && m <= kGmresCgsMaxNl;
    const int64_t nrows = w->nrows, mk = (int64_t)m * kk;
    const int64_t work_elems = cgs ? (int64_t)gmres_cgs_grid(nrows, kk) * mk + mk : 0;
    const size_t dev_elems = cgs ? (size_t)(work_elems + 2 * mk + kk + kk + mk) : 0;
    TI* dbuf = cgs ? static_cast<TI*>(dev_alloc(sizeof(TI) * dev_elems)) : nullptr;
    struct Release {
      void* p;
      size_t n;
      ~Release() { dev_release(p, n); }
    } release{dbuf, sizeof(TI) * dev_elems};
    TI* d_outs = dbuf + work_elems;
    TI* d_act = d_outs + 2 * mk + kk;
    TI* d_y = d_act + kk;
    auto upload = [&](TI* dst, const TI* src, int64_t n) {
      const size_t bytes = sizeof(TI) * n;
      if (!stage_upload_2d(dst, bytes, src, bytes, bytes, 1)) {
        PHX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, rt().stream));
        PHX_HIP(hipStreamSynchronize(rt().stream));
      }
    };
    std::vector<TI> hact(kk, TI(1));
    if (cgs) upload(d_act, hact.data(), kk);
    void *sig_h = nullptr, *sig_d = nullptr;
    const bool devmode = cgs && rt().gmres_device && kk <= 64 && signal_buffer(&sig_h, &sig_d);
    int i = 0;
    if (devmode) {
      // ---- the cycle with the Givens rotations on the device: per step the product, the
      // projection, the three sweeps, the rotations and the new basis vector are enqueued
      // without waiting; the host then reads the active flags of the step before (the device
      // works on this step meanwhile) and stops once no system is left ----
      const size_t m1 = (size_t)m + 1, nhd = (size_t)kk * m1 + 3 * (size_t)kk;
      const size_t nd = nhd + (size_t)kk * m * m1 + 2 * (size_t)kk * m;
      const size_t sbytes = nd * sizeof(double) + 3 * (size_t)kk * sizeof(int);
      const size_t inv_off = (sbytes + 15) / 16 * 16, abytes = inv_off + sizeof(TI) * (size_t)kk;
      char* sbuf = static_cast<char*>(dev_alloc(abytes));
      struct ReleaseState {
        void* p;
        size_t n;
        ~ReleaseState() { dev_release(p, n); }
      } release_state{sbuf, abytes};
      GmresDevState ds;
      double* dp = reinterpret_cast<double*>(sbuf);
      ds.g = dp;
      ds.normR0 = dp + (size_t)kk * m1;
      ds.tol = ds.normR0 + kk;
      ds.relres = ds.tol + kk;
      ds.H = ds.relres + kk;
      ds.cs = ds.H + (size_t)kk * m * m1;
      ds.sn = ds.cs + (size_t)kk * m;
      int* ip = reinterpret_cast<int*>(dp + nd);
      ds.steps = ip;
      ds.status = ip + kk;
      ds.iters0 = ip + 2 * kk;
      TI* d_inv = reinterpret_cast<TI*>(sbuf + inv_off);
      std::vector<double> hd(nhd, 0.0);
      std::vector<int> hi(3 * (size_t)kk, 0);
      for (int a = 0; a < kk; a++) {
        hd[(size_t)a * m1] = rn[a];
        hd[(size_t)kk * m1 + a] = normR0[act[a]];
        hd[(size_t)kk * m1 + kk + a] = tol[act[a]];
        hd[(size_t)kk * m1 + 2 * kk + a] = out.relres[act[a]];
        hi[kk + a] = 1;
        hi[2 * kk + a] = out.iters[act[a]];
      }
      auto upload_bytes = [&](void* dst, const void* src, size_t bytes) {
        if (!stage_upload_2d(dst, bytes, src, bytes, bytes, 1)) {
          PHX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, rt().stream));
          PHX_HIP(hipStreamSynchronize(rt().stream));
        }
      };
      upload_bytes(ds.g, hd.data(), sizeof(double) * nhd);
      upload_bytes(ds.steps, hi.data(), sizeof(int) * hi.size());
      uint32_t seqs[2] = {0, 0};
      for (; i < m; i++) {
        op.apply_shifted(TI(1), s.data(), vcur.get(), TI(0), w.get(), nullptr);
        project(w.get());
        const int nl = i + 1, nc = nl * kk;
        TI* D1 = d_outs;
        TI* D2 = d_outs + nc;
        TI* N = d_outs + 2 * nc;
        const TI* v = Vb->dev();
        const int64_t ldv = Vb->lda;
        gmres_cgs_dev<TI>(0, nrows, v, ldv, w->dev(), w->lda, kk, nl, nullptr, nullptr, dbuf, work_elems, D1);
        gmres_cgs_dev<TI>(1, nrows, v, ldv, w->dev(), w->lda, kk, nl, D1, d_act, dbuf, work_elems, D2);
        gmres_cgs_dev<TI>(2, nrows, v, ldv, w->dev(), w->lda, kk, nl, D2, d_act, dbuf, work_elems, N);
        gmres_givens_dev<TI>(ds, kk, m, i, d_outs, d_act, d_inv, 1000 * eps, maxIter);
        View<TI> vn(Vb.get(), (i + 1) * kk, (i + 2) * kk - 1);
        gmres_normalize_dev<TI>(nrows, kk, w->dev(), w->lda, d_inv, N, vcur->dev(), vcur->lda, vn->dev(), vn->lda);
        seqs[i % 2] = publish_async(d_act, sizeof(TI) * kk, 1 + i % 2);
        PHX_REQUIRE(seqs[i % 2] != 0, PHIST_FUNCTIONAL_ERROR, "GMRES: the active flags could not be published");
        if (i >= 1) {
          const TI* f = static_cast<const TI*>(published_data(seqs[(i - 1) % 2], 1 + (i - 1) % 2));
          bool any = false;
          for (int a = 0; a < kk; a++) any = any || f[a] != TI(0);
          if (!any) break;  // step i was enqueued already: for stopped systems it changes nothing
        }
      }
      // the cycle's recurrences back to the host (one synchronisation per cycle)
      std::vector<char> hs(sbytes);
      PHX_HIP(hipMemcpyAsync(hs.data(), sbuf, sbytes, hipMemcpyDeviceToHost, rt().stream));
      PHX_HIP(hipStreamSynchronize(rt().stream));
      const double* hg = reinterpret_cast<const double*>(hs.data());
      const double* hrel = hg + (size_t)kk * m1 + 2 * kk;
      const double* hH = hrel + kk;
      const int* hsteps = reinterpret_cast<const int*>(hg + nd);
      const int* hstat = hsteps + kk;
      for (int a = 0; a < kk; a++) {
        const int j = act[a];
        steps[a] = hsteps[a];
        out.iters[j] += hsteps[a];
        if (hsteps[a] > 0) out.relres[j] = hrel[a];
        if (hstat[a] != 1) status[j] = hstat[a];
        active[a] = hstat[a] == 1;
        std::copy(hg + (size_t)a * m1, hg + (size_t)(a + 1) * m1, &g[(size_t)a * m1]);
        std::copy(hH + (size_t)a * m * m1, hH + (size_t)(a + 1) * m * m1, &H[(size_t)a * m * m1]);
      }
    } else
    for (; i < m; i++) {
      bool live = false;
      for (int a = 0; a < kk; a++) live = live || active[a];
      if (!live) break;
      View<TI> Vi(Vb.get(), 0, (i + 1) * kk - 1);
      op.apply_shifted(TI(1), s.data(), vcur.get(), TI(0), w.get(), nullptr);
      project(w.get());
      std::vector<double> hc((size_t)(i + 1) * kk, 0.0);
      std::vector<TI> d(kk), inv(kk);
      if (cgs) {  // classical Gram-Schmidt twice in three sweeps, the norms in the last
        bool changed = false;
        for (int a = 0; a < kk; a++) {
          const TI f = active[a] ? TI(1) : TI(0);
          changed = changed || f != hact[a];
          hact[a] = f;
        }
        if (changed) upload(d_act, hact.data(), kk);
        const int nl = i + 1, nc = nl * kk;
        TI* D1 = d_outs;
        TI* D2 = d_outs + nc;
        TI* N = d_outs + 2 * nc;
        const TI* v = Vb->dev();
        const int64_t ldv = Vb->lda;
        gmres_cgs_dev<TI>(0, nrows, v, ldv, w->dev(), w->lda, kk, nl, nullptr, nullptr, dbuf, work_elems, D1);
        gmres_cgs_dev<TI>(1, nrows, v, ldv, w->dev(), w->lda, kk, nl, D1, d_act, dbuf, work_elems, D2);
        gmres_cgs_dev<TI>(2, nrows, v, ldv, w->dev(), w->lda, kk, nl, D2, d_act, dbuf, work_elems, N);
        TI* hb = static_cast<TI*>(host_work(sizeof(TI) * (2 * nc + kk)));
        if (!fetch_to_host(hb, d_outs, sizeof(TI) * (2 * nc + kk))) {
          PHX_HIP(hipMemcpyAsync(hb, d_outs, sizeof(TI) * (2 * nc + kk), hipMemcpyDeviceToHost, rt().stream));
          PHX_HIP(hipStreamSynchronize(rt().stream));
        }
        for (int a = 0; a < kk; a++) {
          if (active[a])
            for (int l = 0; l <= i; l++) hc[(size_t)l * kk + a] = (double)hb[l * kk + a] + (double)hb[nc + l * kk + a];
          d[a] = hb[2 * nc + a];
        }
      }
      for (int pass = 0; pass < 2 && !cgs; pass++) {  // classical Gram-Schmidt, twice
        SdBox<TI> G((i + 1) * kk, kk);
        gemm_gram<TI>(TI(1), Vi.get(), w.get(), TI(0), G.get());
        for (int a = 0; a < kk; a++) {
          for (int l = 0; l <= i; l++) {
            const TI c = active[a] ? G->h(l * kk + a, a) : TI(0);
            hc[(size_t)l * kk + a] += (double)c;
          }
        }
        SdBox<TI> C((i + 1) * kk, kk);
        for (int a = 0; a < kk; a++)
          for (int l = 0; l <= i; l++) C->h(l * kk + a, a) = active[a] ? G->h(l * kk + a, a) : TI(0);
        sd_to_device(C.get());
        gemm_times_sd<TI>(TI(-1), Vi.get(), C.get(), TI(1), w.get());
      }
      if (!cgs) blas_dot<TI>(w.get(), w.get(), d.data());
      for (int a = 0; a < kk; a++) {
        const int j = act[a];
        const double hn = std::sqrt((double)d[a]);
        inv[a] = (active[a] && hn > 0.0 && std::isfinite(hn)) ? TI(1.0 / hn) : TI(0);
        if (!active[a]) continue;
        if (!std::isfinite(hn)) {  // broken down: keep the last solution of this cycle
          active[a] = 0;
          status[j] = 3;
          continue;
        }
        // Hessenberg column i: h(0:i, i) and h(i+1, i); earlier rotations, then a new one
        for (int l = 0; l <= i; l++) h(a, l, i) = hc[(size_t)l * kk + a];
        h(a, i + 1, i) = hn;
        for (int l = 0; l < i; l++) {
          const double c = cs[(size_t)a * m + l], sv = sn[(size_t)a * m + l];
          const double p0 = h(a, l, i), p1 = h(a, l + 1, i);
          h(a, l, i) = c * p0 + sv * p1;
          h(a, l + 1, i) = -sv * p0 + c * p1;
        }
        const double p0 = h(a, i, i), p1 = h(a, i + 1, i);
        const double r = std::hypot(p0, p1);
        const double c = r > 0.0 ? p0 / r : 1.0, sv = r > 0.0 ? p1 / r : 0.0;
        cs[(size_t)a * m + i] = c;
        sn[(size_t)a * m + i] = sv;
        h(a, i, i) = r;
        h(a, i + 1, i) = 0.0;
        double* ga = &g[(size_t)a * (m + 1)];
        ga[i + 1] = -sv * ga[i];
        ga[i] = c * ga[i];
        steps[a] = i + 1;
        out.iters[j]++;
        const double nr = std::fabs(ga[i + 1]);
        const double rel = nr / normR0[j];
        out.relres[j] = rel;
        if (nr < 1000 * eps || rel < tol[j] || hn == 0.0) {
          active[a] = 0;
          status[j] = 0;
        } else if (out.iters[j] >= maxIter) {
          active[a] = 0;
          status[j] = 3;
        }
      }
      View<TI> vn(Vb.get(), (i + 1) * kk, (i + 2) * kk - 1);
      blas_axpby<TI>(inv.data(), false, w.get(), TI(0), vcur.get());
      for (int a = 0; a < kk; a++)  // a non-finite column must not reach the other systems through 0 * NaN
        if (!std::isfinite((double)d[a])) {
          View<TI> col(vcur.get(), a, a);
          blas_put_value<TI>(col.get(), TI(0));
        }
      blas_copy<TI>(vcur.get(), vn.get());
    }
    // ---- x += V y per system (y from the triangular systems R y = g) ----
    int L = 0;
    for (int a = 0; a < kk; a++) L = std::max(L, steps[a]);
    auto ysolve = [&](int a) {  // back substitution with the rotated Hessenberg matrix of system a
      const int n = steps[a];
      std::vector<double> y(n);
      for (int r = n - 1; r >= 0; r--) {
        double acc = g[(size_t)a * (m + 1) + r];
        for (int c = r + 1; c < n; c++) acc -= h(a, r, c) * y[c];
        y[r] = h(a, r, r) != 0.0 ? acc / h(a, r, r) : 0.0;
      }
      return y;
    };
    if (L > 0 && cgs) {  // x -= V c with c(l, a) = -y_a(l): one sweep over the basis
      std::vector<TI> yc((size_t)L * kk, TI(0));
      for (int a = 0; a < kk; a++) {
        const int n = steps[a];
        const std::vector<double> y = ysolve(a);
        for (int r = 0; r < n; r++) yc[(size_t)r * kk + a] = (TI)(-y[r]);
      }
      upload(d_y, yc.data(), (int64_t)L * kk);
      gmres_cgs_dev<TI>(3, nrows, Vb->dev(), Vb->lda, xc->get()->dev(), xc->get()->lda, kk, L, d_y, nullptr, nullptr,
                        0, nullptr);
    } else if (L > 0) {
      SdBox<TI> Y(L * kk, kk);
      for (int a = 0; a < kk; a++) {
        const int n = steps[a];
        const std::vector<double> y = ysolve(a);
        for (int r = 0; r < n; r++) Y->h(r * kk + a, a) = (TI)y[r];
      }
      sd_to_device(Y.get());
      View<TI> VL(Vb.get(), 0, L * kk - 1);
      gemm_times_sd<TI>(TI(1), VL.get(), Y.get(), TI(1), xc->get());
    }
    // ---- the true residual: restart, or the final residuals of the systems that stopped ----
    residual(rn);
  }
  copy_any<TI, T>(x.get(), X);
  out.unconverged = 0;
  for (int j = 0; j < k; j++) out.unconverged += status[j] != 0;
  return out;
}

template GmresOut gmres_solve<double, double>(const Operator<double>&, const std::vector<double>&,
                                              const Mvec<double>*, Mvec<double>*, const std::vector<double>&, int,
                                              int, const Mvec<double>*);
template GmresOut gmres_solve<double, float>(const Operator<float>&, const std::vector<double>&, const Mvec<double>*,
                                             Mvec<double>*, const std::vector<double>&, int, int,
                                             const Mvec<float>*);
template GmresOut gmres_solve<float, float>(const Operator<float>&, const std::vector<float>&, const Mvec<float>*,
                                            Mvec<float>*, const std::vector<double>&, int, int, const Mvec<float>*);

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
