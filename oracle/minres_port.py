"""CPU port of blocked shifted MINRES, for the CPU reference time of the
batched shifted MINRES configuration (honeycomb sheet, 32 shifts, 32 unit
right-hand sides, tol 1e-6).

TEST INFRASTRUCTURE AND CPU REFERENCE TIMING ONLY.  bench.py's cpu_baseline leg times
it and tests/ check it; the product library never imports, links or runs it.

k systems (A + s_j I) x_j = b_j iterate together, each with its own Lanczos
recurrence and Paige-Saunders rotations (the Krylov iterates of the
reference's blocked MINRES, src/krylov/phist_blockedminres_def.hpp:201-440,
without its basis ring buffer): per iteration one shifted SpMV of the block
(the reference's flang-built OpenMP dspmvm when built, else the oracle's C
SpMV), then the Lanczos update w -= a v + b v_old with the new norms, the
rotations, the direction d = (v - delta d1 - eps d2) / gamma and x += tau d.
A system stops at |phibar| / ||b|| <= tol (the recurrence's residual
estimate, as the reference's) or at max_iter.  The vector passes run on
row blocks in a thread pool (numpy releases the interpreter lock in its
loops), so every pass uses the host's cores as builtin's OpenMP loops do.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np


def _host_cpus():
    """the CPUs this process may run on (the cgroup's set; the calling thread's own
    affinity can be narrower: an OpenMP runtime with OMP_PROC_BIND binds its
    initial thread to one core, and threads started later inherit that)"""
    try:
        with open("/sys/fs/cgroup/cpuset.cpus.effective") as f:
            cpus = set()
            for part in f.read().strip().split(","):
                if part:
                    a, _, b = part.partition("-")
                    cpus.update(range(int(a), int(b or a) + 1))
            if cpus:
                return cpus
    except OSError:
        pass
    return set(range(os.cpu_count() or 1))


def _unpin():
    try:
        os.sched_setaffinity(0, _host_cpus())
    except OSError:
        pass


class RowPool:
    """elementwise passes and column dots over row blocks of (n, k) arrays; the
    workers may run on every CPU of the process"""

    def __init__(self, n, threads):
        self.threads = max(1, threads)
        step = -(-n // self.threads)
        self.blocks = [(r, min(n, r + step)) for r in range(0, n, step)]
        self.ex = ThreadPoolExecutor(self.threads, initializer=_unpin) if self.threads > 1 else None

    def run(self, f):
        """f(r0, r1) on every row block; the results in block order"""
        if self.ex is None:
            return [f(r0, r1) for r0, r1 in self.blocks]
        return list(self.ex.map(lambda b: f(*b), self.blocks))

    def colsum(self, f):
        """sum over the row blocks of f(r0, r1) (per-column partial sums, added in block order)"""
        out = None
        for part in self.run(f):
            out = part if out is None else out + part
        return out

    def close(self):
        if self.ex is not None:
            self.ex.shutdown()


def blocked_minres(op, shifts, B, tol, max_iter, threads=None, on_iteration=None):
    """solves (A + s_j I) x_j = b_j for the k columns of B from x = 0.  Returns (X,
    iterations, recurrence relative residuals).  on_iteration(k) is called after
    every iteration (timing of a bounded sample)."""
    n, k = B.shape
    threads = threads or int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    pool = RowPool(n, threads)
    sh = np.asarray(shifts, np.float64)
    beta1 = np.sqrt(pool.colsum(lambda r0, r1: np.einsum("ij,ij->j", B[r0:r1], B[r0:r1])))
    active = (beta1 > 0) & (max_iter > 0)
    inv = np.where(beta1 > 0, 1.0 / np.where(beta1 > 0, beta1, 1.0), 0.0)
    V = np.empty((n, k))
    Vold = np.zeros((n, k))
    D1 = np.zeros((n, k))
    D2 = np.zeros((n, k))
    X = np.zeros((n, k))
    pool.run(lambda r0, r1: np.multiply(B[r0:r1], inv, out=V[r0:r1]))
    bk = np.zeros(k)
    c1, s1, c2, s2 = np.ones(k), np.zeros(k), np.ones(k), np.zeros(k)
    phibar = beta1.copy()
    iters = np.zeros(k, int)
    relres = np.where(beta1 > 0, 1.0, 0.0)
    it = 0
    while active.any() and it < max_iter:
        W = op.apply(V, sh)
        mask = active.astype(np.float64)

        def lanczos_a(r0, r1):  # w -= b v_old, then <v, w>
            W[r0:r1] -= (bk * mask) * Vold[r0:r1]
            return np.einsum("ij,ij->j", V[r0:r1], W[r0:r1])

        a = pool.colsum(lanczos_a)

        def lanczos_b(r0, r1):  # w -= a v, then ||w||^2
            W[r0:r1] -= (a * mask) * V[r0:r1]
            return np.einsum("ij,ij->j", W[r0:r1], W[r0:r1])

        bn = np.sqrt(pool.colsum(lanczos_b))
        # the Paige-Saunders rotations of every column
        eps_k = s2 * bk
        dbar = c2 * bk
        delta = c1 * dbar + s1 * a
        gbar = -s1 * dbar + c1 * a
        gamma = np.hypot(gbar, bn)
        ok = active & (gamma > 0) & np.isfinite(gamma)
        gs = np.where(ok, gamma, 1.0)
        cs, sn = np.where(ok, gbar / gs, 1.0), np.where(ok, bn / gs, 0.0)
        tau = np.where(ok, cs * phibar, 0.0)
        ig = np.where(ok, 1.0 / gs, 0.0)
        dl, ep = np.where(ok, delta, 0.0), np.where(ok, eps_k, 0.0)
        ib = np.where(ok & (bn > 0), 1.0 / np.where(bn > 0, bn, 1.0), 0.0)

        def update(r0, r1):  # d = (v - delta d1 - eps d2) / gamma, x += tau d, next v = w / beta
            d = (V[r0:r1] - dl * D1[r0:r1] - ep * D2[r0:r1]) * ig
            X[r0:r1] += tau * d
            D2[r0:r1] = D1[r0:r1]
            D1[r0:r1] = d
            Vold[r0:r1] = V[r0:r1]
            np.multiply(W[r0:r1], ib, out=V[r0:r1])

        pool.run(update)
        phibar = np.where(ok, -sn * phibar, phibar)
        c2, s2 = np.where(ok, c1, c2), np.where(ok, s1, s2)
        c1, s1 = np.where(ok, cs, c1), np.where(ok, sn, s1)
        bk = np.where(ok, bn, bk)
        iters += ok
        relres = np.where(ok, np.abs(phibar) / np.where(beta1 > 0, beta1, 1.0), relres)
        active = ok & (relres > tol) & (bn > 0) & (iters < max_iter)
        it += 1
        if on_iteration is not None:
            on_iteration(it)
    pool.close()
    return X, iters, relres
