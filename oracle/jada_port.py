"""CPU port of PHIST's symmetric block Jacobi-Davidson with blocked-GMRES
correction solves, for the CPU reference time of the JD time-to-solution
configuration (3D Laplacian, 8 smallest pairs, blockSize 4, minBas 16,
maxBas 32, at most 10 GMRES iterations per correction, tol 1e-8).

TEST INFRASTRUCTURE AND CPU REFERENCE TIMING ONLY.  bench.py's cpu_baseline leg times
it and tests/ check it; the product library never imports, links or runs it.

The reference's own JD (src/jada/phist_subspacejada_def.hpp) needs MPI and
the Fortran modules of the builtin kernel library, which do not build here.
This port restates its control flow for the symmetric, standard-Ritz case:
  start-up: single-vector Arnoldi to minBas columns, classical Gram-Schmidt
    twice (phist_subspacejada_def.hpp:335-380, phist_simple_arnoldi_def.hpp)
  Ritz pairs of H = V'AV sorted by `which`, R couplings to the locked
    vectors, residuals QR - AQ and their norms          :509-630
  convergence and locking, thick restart to minBas (transform_searchSpace,
    src/core/phist_transform_searchspace_def.hpp:11-45)  :710-833
  selection of up to blockSize unconverged pairs without splitting
    clusters, shrinking before the basis overflows        :847-910
  inner tolerances                                        :933-943
  correction equations (I - QQ')(A - theta_j I) t_j = -r_j
    (phist_jadaCorrectionSolver_def.hpp:499-512) by blocked GMRES, one Krylov
    basis per system, classical Gram-Schmidt twice, Givens rotations, no
    restart within innerSolvMaxIters, the true residual at the end
    (src/krylov/phist_blockedgmres_def.hpp:453-1010)
  extension: t orthogonalised against V (classical Gram-Schmidt twice) and
    orthonormalised by Cholesky QR twice (phist_orthog_def.hpp), H extended
    by V'(At)                                             :955-1005
The sparse products are the reference's own builtin kernels
(spmvm_kernels.F90 dspmvm_1/2/4/8/generic, OpenMP, flang-built into
oracle/_ref) when those are built, else the oracle's C restatement; block
products and Gram matrices use numpy (BLAS); the small dense eigenproblems
LAPACK.  All vectors are row-major (n, k) blocks as in builtin.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
REF_SO = os.path.join(_HERE, "_ref", "libbuiltin_leaf.so")
ORACLE_SO = os.path.join(_HERE, "liboracle.so")


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class CrsOperator:
    """y = (A + diag(shifts)) x on a 1-based builtin CRS (int64 row offsets,
    int32 column indices); `kind` says which kernels run"""

    def __init__(self, nrows, row_offset, col_idx, val, prefer_reference=True):
        self.n = nrows
        self.ro = np.ascontiguousarray(row_offset, np.int64)
        self.ci = np.ascontiguousarray(col_idx, np.int32)
        self.va = np.ascontiguousarray(val, np.float64)
        self.halo_ptr = np.ascontiguousarray(self.ro[1:])  # no halo: row ends
        self.nnz = int(self.ro[-1] - 1)
        self.ref = ctypes.CDLL(REF_SO) if prefer_reference and os.path.exists(REF_SO) else None
        self.kind = "reference" if self.ref is not None else "port"
        if self.ref is None:
            self.orc = ctypes.CDLL(ORACLE_SO)
        self.products = 0

    def apply(self, x, shifts=None):
        x = np.ascontiguousarray(x, np.float64)
        n, k = x.shape
        y = np.zeros((n, k))  # the builtin kernels form alpha*tmp + beta*y: y must hold numbers
        sh = np.zeros(k) if shifts is None else np.ascontiguousarray(shifts, np.float64)
        self.products += 1
        I, L, D = ctypes.c_int, ctypes.c_int64, ctypes.c_double
        if self.ref is not None:
            halo = np.zeros(k)
            args = [ctypes.byref(I(n)), ctypes.byref(I(0)), ctypes.byref(I(n)), ctypes.byref(L(self.nnz)),
                    ctypes.byref(D(1.0)), _p(self.ro), _p(self.halo_ptr), _p(self.ci), _p(self.va), _p(sh), _p(x)]
            if k in (1, 2, 4, 8):
                getattr(self.ref, f"dspmvm_{k}_")(*args, _p(halo), ctypes.byref(D(0.0)), _p(y))
            else:
                self.ref.dspmvm_generic_(ctypes.byref(I(k)), *args[:10], _p(x), ctypes.byref(I(k)), _p(halo),
                                         ctypes.byref(D(0.0)), _p(y), ctypes.byref(I(k)))
        else:
            self.orc.oracle_spmv(L(n), _p(self.ro), _p(self.ci), _p(self.va), I(k), D(1.0), _p(sh), _p(x), L(k),
                                 D(0.0), _p(y), L(k), I(1))
        return y


def _cgs2(V, w):
    """w -= V (V'w) twice; returns the summed coefficients (V orthonormal columns)"""
    h = V.T @ w
    w -= V @ h
    h2 = V.T @ w
    w -= V @ h2
    return h + h2


def _cgs2_rows(Vr, w):
    """as _cgs2 with the basis vectors as the contiguous rows of Vr"""
    h = Vr @ w
    w -= Vr.T @ h
    h2 = Vr @ w
    w -= Vr.T @ h2
    return h + h2


def _orthonormalize(V, t):
    """t against the orthonormal V (classical Gram-Schmidt twice), then Cholesky QR
    twice; a Cholesky failure (t numerically rank deficient) falls back to QR"""
    if V.shape[1]:
        _cgs2(V, t)
    for _ in range(2):
        G = t.T @ t
        try:
            R = np.linalg.cholesky(G).T
        except np.linalg.LinAlgError:
            q, _r = np.linalg.qr(t)
            t[:] = q
            return t
        t[:] = np.linalg.solve(R.T, t.T).T
    return t


def blocked_gmres(op, shifts, B, Q, tol, max_iter, m):
    """k systems P (A + s_j I) x_j = b_j, P = I - QQ', from x = 0, each with its own
    basis of at most m + 1 vectors (src/krylov/phist_blockedgmres_def.hpp:453-1010):
    restarted GMRES(m) with classical Gram-Schmidt twice and Givens rotations;
    stops a system at ||r|| / ||r0|| < tol_j or ||r|| < 1000 eps or after max_iter
    steps.  Returns (x, iterations, relative residuals)."""
    n, k = B.shape
    eps = np.finfo(np.float64).eps

    def proj(w):
        if Q is not None and Q.shape[1]:
            w -= Q @ (Q.T @ w)
        return w

    X = np.zeros((n, k))
    iters = np.zeros(k, int)
    relres = np.ones(k)
    r0 = None
    status = np.ones(k, bool)  # still iterating
    R = B.copy()  # r = b - P (A + sI) x with x = 0
    while True:
        rn = np.linalg.norm(R, axis=0)
        if r0 is None:
            r0 = rn.copy()
        for a in range(k):
            if not status[a]:
                continue
            relres[a] = rn[a] / r0[a] if r0[a] > 0 else 0.0
            if rn[a] < 1000 * eps or relres[a] < tol[a] or iters[a] >= max_iter:
                status[a] = False
        act = np.nonzero(status)[0]
        if act.size == 0:
            break
        # one GMRES(m) cycle on the active systems
        kk = act.size
        basis = np.zeros((kk, m + 1, n))  # system b_'s basis vectors: contiguous rows
        H = np.zeros((kk, m + 1, m))
        cs, sn = np.zeros((kk, m)), np.zeros((kk, m))
        g = np.zeros((kk, m + 1))
        for b_, a in enumerate(act):
            basis[b_, 0] = R[:, a] / rn[a]
            g[b_, 0] = rn[a]
        live = np.ones(kk, bool)
        steps = np.zeros(kk, int)
        s = np.array([shifts[a] for a in act])
        for i in range(m):
            if not live.any():
                break
            W = proj(op.apply(np.ascontiguousarray(basis[:, i].T), s))
            for b_ in range(kk):
                if not live[b_]:
                    continue
                a = act[b_]
                w = np.ascontiguousarray(W[:, b_])
                h = _cgs2_rows(basis[b_, :i + 1], w)
                hn = float(np.linalg.norm(w))
                H[b_, :i + 1, i] = h
                H[b_, i + 1, i] = hn
                for l in range(i):  # earlier rotations, then a new one
                    c, sv = cs[b_, l], sn[b_, l]
                    p0, p1 = H[b_, l, i], H[b_, l + 1, i]
                    H[b_, l, i], H[b_, l + 1, i] = c * p0 + sv * p1, -sv * p0 + c * p1
                p0, p1 = H[b_, i, i], H[b_, i + 1, i]
                r = np.hypot(p0, p1)
                c, sv = (p0 / r, p1 / r) if r > 0 else (1.0, 0.0)
                cs[b_, i], sn[b_, i] = c, sv
                H[b_, i, i], H[b_, i + 1, i] = r, 0.0
                g[b_, i + 1] = -sv * g[b_, i]
                g[b_, i] = c * g[b_, i]
                steps[b_] = i + 1
                iters[a] += 1
                nr = abs(g[b_, i + 1])
                relres[a] = nr / r0[a]
                if nr < 1000 * eps or relres[a] < tol[a] or hn == 0.0 or iters[a] >= max_iter:
                    live[b_] = False
                    status[a] = False  # done: converged by the implicit residual, or out of steps
                if hn > 0:
                    basis[b_, i + 1] = w / hn
        for b_, a in enumerate(act):  # x += V y, R y = g by back substitution
            ns = steps[b_]
            y = np.zeros(ns)
            for r in range(ns - 1, -1, -1):
                acc = g[b_, r] - H[b_, r, r + 1:ns] @ y[r + 1:]
                y[r] = acc / H[b_, r, r] if H[b_, r, r] != 0.0 else 0.0
            if ns:
                X[:, a] += basis[b_, :ns].T @ y
        if not status.any():
            break
        # true residual (restart), then the test at the top
        R = B - proj(op.apply(X, shifts))
    return X, iters, relres


def subspacejada(op, v0, num_eigs, block_size, min_bas, max_bas, tol, inner_max_iters, max_iters=300,
                 inner_base_tol=0.1, which="SR"):
    """symmetric block JD (standard Ritz values, B = I, no preconditioner, blocked
    GMRES correction solves).  v0: start vector (n,).  Returns (eigenvalues,
    residual norms, converged, outer iterations, Q)."""
    n = v0.shape[0]
    eps = np.finfo(np.float64).eps
    bs, nEig = block_size, num_eigs
    nEigMax = nEig + bs - 1
    look_ahead = 2 * bs
    V = np.zeros((n, max_bas))
    AV = np.zeros((n, max_bas))
    H = np.zeros((max_bas, max_bas))
    Q = np.zeros((n, nEigMax))
    Rm = np.zeros((nEigMax, nEigMax))
    theta = np.zeros(max_bas)
    resNorm = np.full(nEigMax, -1.0)
    innerTol = np.full(nEigMax, inner_base_tol)
    lastOuterRes = np.zeros(nEigMax)

    def order(w):
        keys = {"SR": w, "LR": -w, "SM": np.abs(w), "LM": -np.abs(w)}[which]
        return np.argsort(keys, kind="stable")

    # start-up Arnoldi: A V(:, 0:m) = V(:, 0:m+1) H(0:m+1, 0:m), AV kept explicitly
    V[:, 0] = v0 / np.linalg.norm(v0)
    nV = min_bas
    for j in range(nV):
        w = op.apply(V[:, j:j + 1])[:, 0]
        AV[:, j] = w
        if j + 1 >= max_bas:
            break
        h = _cgs2(V[:, :j + 1], w)
        nrm = np.linalg.norm(w)
        H[:j + 1, j] = h
        H[j + 1, j] = nrm
        V[:, j + 1] = w / nrm
    for j in range(nV):
        for i in range(j + 1, nV):
            H[j, i] = H[i, j]

    def transform(c0, c1, mcols, qh, m):
        """V(:, c0:c0+nc) <- V(:, c0:c1) qh(:, 0:nc), same for AV; H block <- qh'H qh
        and the couplings of the locked vectors (transform_searchSpace)"""
        nc = mcols - c0
        if nc <= 0:
            return
        M = qh[:, :nc]
        V[:, c0:c0 + nc] = V[:, c0:c1] @ M
        AV[:, c0:c0 + nc] = AV[:, c0:c1] @ M
        Hb = M.T @ H[c0:c1, c0:c1] @ M
        H[c0:c0 + nc, c0:c0 + nc] = 0.5 * (Hb + Hb.T)
        cpl = H[:c0, c0:c1] @ M
        H[:c0, c0:c0 + nc] = cpl
        H[c0:c0 + nc, :c0] = cpl.T

    nConv, it = 0, 0
    while it < max_iters:
        nEig_ = min(nConv + look_ahead, nEigMax)
        m = nV - nConv
        w, Y = np.linalg.eigh(H[nConv:nV, nConv:nV])
        o = order(w)
        theta[nConv:nV] = w[o]
        qh = Y[:, o]
        nq = nEig_ - nConv
        QH = qh[:, :nq]
        for j in range(nConv, nEig_):
            Rm[:nEig_, j] = 0.0
            Rm[j, j] = theta[j]
            Rm[:nConv, j] = H[:nConv, nConv:nV] @ qh[:, j - nConv]
        Q[:, nConv:nEig_] = V[:, nConv:nV] @ QH
        res = Q[:, :nEig_] @ Rm[:nEig_, nConv:nEig_] - AV[:, nConv:nV] @ QH  # QR - AQ
        resNorm[nConv:nEig_] = np.linalg.norm(res, axis=0)
        nNew = 0
        for i in range(nConv, min(nEig, nEig_)):
            if resNorm[i] <= tol and i == nConv + nNew:
                nNew += 1
            elif bs == 1:
                break
        if nConv + nNew >= nEig:
            nConv += nNew
            break
        if nNew > 0:  # lock: the converged Ritz vectors lead the transformed basis
            mcols = min_bas if nV + 2 * bs > max_bas else nV
            transform(nConv, nV, mcols, qh, m)
            nConv += nNew
            if nV + 2 * bs > max_bas:
                nV = min_bas
        k, k_ = 0, 0
        sigma, selected = [], []
        for i in range(nEig_):
            if k >= bs:
                break
            if i >= nEig and abs(theta[i] - theta[i - 1]) > np.sqrt(tol):
                break
            if resNorm[i] > tol:
                k_ = i
                sigma.append(theta[i])
                selected.append(i)
                k += 1
        while k_ + 1 < nEig_ and abs(theta[k_ + 1] - theta[k_]) < 10 * abs(theta[k_ + 1]) * np.sqrt(tol):
            k_ += 1
        if nV + k > max_bas:
            if nNew == 0:
                transform(nConv, nV, min_bas, qh, m)
            nV = min_bas
        if k == 0:
            it += 1
            continue
        for i in range(k):
            idx = nConv + i
            if resNorm[idx] > 4 * lastOuterRes[idx]:
                innerTol[idx] = inner_base_tol
            if innerTol[idx] > eps:
                innerTol[idx] = max(eps, innerTol[idx] * inner_base_tol)
            innerTol[idx] = max(innerTol[idx], 0.1 * tol / (eps + resNorm[idx]))
            innerTol[idx] = min(innerTol[idx], 0.1)
            lastOuterRes[idx] = resNorm[idx]
        oldConv = nConv - nNew
        rhs = np.ascontiguousarray(res[:, [s - oldConv for s in selected]])
        Qtil = np.ascontiguousarray(Q[:, :k_ + 1])
        t, _its, _rr = blocked_gmres(op, [-s for s in sigma], rhs, Qtil, innerTol[nConv:nConv + k],
                                     inner_max_iters, inner_max_iters)
        t = _orthonormalize(V[:, :nV], t)
        V[:, nV:nV + k] = t
        AV[:, nV:nV + k] = op.apply(t)
        Hn = V[:, :nV + k].T @ AV[:, nV:nV + k]
        H[:nV + k, nV:nV + k] = Hn
        H[nV:nV + k, :nV] = Hn[:nV].T
        blk = H[nV:nV + k, nV:nV + k]
        H[nV:nV + k, nV:nV + k] = 0.5 * (blk + blk.T)
        nV += k
        it += 1
    return theta[:nEigMax].copy(), resNorm[:nEig].copy(), nConv, it, Q
