"""CPU: the blocked shifted MINRES port (oracle/minres_port.py), which bench.py
times as the CPU reference of the batched MINRES configuration.

Pins: on a 3D Laplacian with definite and indefinite shifts every system's
true relative residual ||b - (A + sI) x|| / ||b|| equals the recurrence
estimate it reports (to 1e-6 relative) and is <= tol for the systems that
converge; the solution agrees with a sparse direct solve to tol times the
condition number; a zero right-hand side stays zero with no iterations; the
row-block thread pool (its dots summed by row block) gives the single-thread
solutions to the tolerance and iteration counts within 2 %."""
import os
import sys

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import jada_port as jp  # noqa: E402
import minres_port as mp  # noqa: E402
from test_jada_port import lap3d_crs  # noqa: E402


def test_minres_port_shifted_laplacian():
    n, ro, ci, va = lap3d_crs(12, 13, 14)
    A = sp.csr_matrix((va, ci - 1, ro - 1), shape=(n, n))
    op = jp.CrsOperator(n, ro, ci, va, prefer_reference=False)
    rng = np.random.default_rng(5)
    B = rng.standard_normal((n, 5))
    B[:, 3] = 0.0
    shifts = np.array([0.0, 0.5, -0.6, 1.0, -1.3])  # smallest eigenvalue of A: 0.26
    tol = 1e-8
    X, its, rr = mp.blocked_minres(op, shifts, B, tol, 4000, threads=1)
    assert its[3] == 0 and np.all(X[:, 3] == 0.0)
    for j in (0, 1, 2, 4):
        M = (A + shifts[j] * sp.identity(n)).tocsc()
        true = np.linalg.norm(B[:, j] - M @ X[:, j]) / np.linalg.norm(B[:, j])
        assert rr[j] <= tol and true <= 1.01 * tol, (j, its[j], rr[j], true)
        assert abs(true - rr[j]) <= 1e-6 * tol + 1e-3 * rr[j]
        xd = spla.spsolve(M, B[:, j])
        ev = np.abs(np.linalg.eigvalsh(M.toarray()))
        assert np.linalg.norm(X[:, j] - xd) <= 10 * tol * (ev.max() / ev.min()) * np.linalg.norm(xd)
    assert its[1] < its[0] < its[2]  # a larger definite shift converges faster, the indefinite ones slower
    X4, its4, rr4 = mp.blocked_minres(op, shifts, B, tol, 4000, threads=4)
    assert np.all(np.abs(its4 - its) <= np.maximum(2, 0.02 * its))
    for j in (0, 1, 2, 4):
        assert np.linalg.norm(X4[:, j] - X[:, j]) <= 1e-6 * np.linalg.norm(X[:, j])
