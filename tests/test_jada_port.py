"""CPU: the numpy port of the block Jacobi-Davidson (oracle/jada_port.py),
which bench.py times as the CPU reference of the JD time-to-solution
configuration.

Pins: on a 3D Dirichlet Laplacian box with a simple spectrum the 8 smallest
Ritz values equal the closed form 2 (3 - cos(pi a/(nx+1)) - cos(pi b/(ny+1))
- cos(pi c/(nz+1))) to 1e-12, every residual is below tol, and the partial
Schur form holds (Q'Q = I, ||AQ - QR|| <= 10 tol per column: the checks of
test/jada/TestSubspaceJaDa_def.hpp); the same with the reference's flang-built
SpMV where that is built.  The port's blocked GMRES reaches its tolerance on
projected shifted systems (true residual of P (A + sI) x = b) and reports
honest iteration counts.
"""
import os
import sys

import numpy as np
import pytest

import oracle_lib as ol

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import jada_port as jp  # noqa: E402


def lap3d_crs(nx, ny, nz):
    """7-point Dirichlet Laplacian (6 on the diagonal, -1 off), rows in lexicographic order,
    columns ascending; 1-based builtin CRS"""
    n = nx * ny * nz
    idx = np.arange(n)
    i, j, k = idx % nx, (idx // nx) % ny, idx // (nx * ny)
    offs = [(-nx * ny, k > 0), (-nx, j > 0), (-1, i > 0), (0, np.ones(n, bool)), (1, i < nx - 1),
            (nx, j < ny - 1), (nx * ny, k < nz - 1)]
    C = np.stack([idx + o for o, _ in offs], 1)
    V = np.stack([np.full(n, 6.0 if o == 0 else -1.0) for o, _ in offs], 1)
    M = np.stack([m for _, m in offs], 1)
    ro = np.concatenate([[1], 1 + np.cumsum(M.sum(1))]).astype(np.int64)
    return n, ro, (C[M] + 1).astype(np.int32), V[M]


def dense(n, ro, ci, va):
    A = np.zeros((n, n))
    for r in range(n):
        A[r, ci[ro[r] - 1:ro[r + 1] - 1] - 1] = va[ro[r] - 1:ro[r + 1] - 1]
    return A


def closed_form(nx, ny, nz, k):
    ev = [2 * (3 - np.cos(np.pi * a / (nx + 1)) - np.cos(np.pi * b / (ny + 1)) - np.cos(np.pi * c / (nz + 1)))
          for a in range(1, nx + 1) for b in range(1, ny + 1) for c in range(1, nz + 1)]
    return np.sort(ev)[:k]


@pytest.mark.parametrize("reference", [False, True])
def test_jada_port_laplacian(reference):
    if reference and not os.path.exists(jp.REF_SO):
        pytest.skip("the reference's leaf kernels are not built (make -C oracle ref)")
    nx, ny, nz = 9, 11, 13  # simple spectrum
    n, ro, ci, va = lap3d_crs(nx, ny, nz)
    op = jp.CrsOperator(n, ro, ci, va, prefer_reference=reference)
    assert op.kind == ("reference" if reference else "port")
    v0 = ol.mvec_random_sequence([(n, 1)])[0][:, 0]
    tol = 1e-8
    ev, res, nconv, it, Q = jp.subspacejada(op, v0, 8, 4, 16, 32, tol, 10)
    assert nconv == 8 and 0 < it < 100
    assert np.all(res <= tol)
    assert np.max(np.abs(np.sort(ev[:8]) - closed_form(nx, ny, nz, 8))) <= 1e-12
    A = dense(n, ro, ci, va)
    Qc = Q[:, :8]
    assert np.max(np.abs(Qc.T @ Qc - np.eye(8))) <= 1e-12
    R = Qc.T @ A @ Qc
    assert np.max(np.linalg.norm(A @ Qc - Qc @ R, axis=0)) <= 10 * tol


def test_jada_port_gmres_projected_systems():
    n, ro, ci, va = lap3d_crs(8, 9, 10)
    op = jp.CrsOperator(n, ro, ci, va, prefer_reference=False)
    A = dense(n, ro, ci, va)
    rng = np.random.default_rng(3)
    Q, _ = np.linalg.qr(rng.standard_normal((n, 3)))
    B = rng.standard_normal((n, 4))
    B -= Q @ (Q.T @ B)
    shifts = [-0.2, 0.0, 0.3, 1.0]  # A + sI definite (smallest eigenvalue of A: 0.30)
    tol = np.array([1e-6, 1e-8, 1e-4, 1e-7])
    X, its, rr = jp.blocked_gmres(op, shifts, B, Q, tol, 400, 30)
    P = np.eye(n) - Q @ Q.T
    for a in range(4):
        r = B[:, a] - P @ ((A + shifts[a] * np.eye(n)) @ X[:, a])
        assert np.linalg.norm(r) <= 1.01 * tol[a] * np.linalg.norm(B[:, a]), (a, its[a])
        assert abs(np.linalg.norm(r) / np.linalg.norm(B[:, a]) - rr[a]) <= 1e-3 * tol[a] + 1e-12
        assert 0 < its[a] <= 400
    # looser tolerances take fewer steps
    assert its[2] < its[1]
