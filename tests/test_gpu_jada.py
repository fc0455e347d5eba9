"""GPU: blocked shifted MINRES and the symmetric block Jacobi-Davidson.

Pins: MINRES solutions by their true residual ||b - (A + s I) x|| computed on
the host from the oracle's CRS product; eigenvalues of the 3D Dirichlet
Laplacian against the closed form sum_d (2 - 2 cos(pi i_d / (n+1))); spin
chain eigenvalues against dense diagonalisation of the oracle's matrix; the
returned partial Schur form by ||A Q - Q R|| and Q'Q = I (the invariants
test/jada/TestSubspaceJaDa_def.hpp checks).  The FP32-inner (mixed) runs must
reach the same FP64 tolerance.  The eigensolver tests run with both
correction-equation solvers the reference offers (innerSolvType MINRES and
GMRES, phist_jadaCorrectionSolver_def.hpp:499-512; GMRES is the reference's
default, phist_jadaOpts.cpp:40).
"""
import itertools
import os

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def laplace_eigs(nx, ny, nz, k):
    lam = [[2 - 2 * np.cos(np.pi * i / (n + 1)) for i in range(1, n + 1)] for n in (nx, ny, nz)]
    ev = sorted(a + b + c for a, b, c in itertools.product(*lam))
    return np.array(ev[:k])


def crs_of(A):
    ro, ci, va = A.crs()
    return ol.Crs(A.nrows, A.ncols, ro, ci, va.astype(float))


@pytest.mark.parametrize("fp32", [False, True])
def test_minres_shifted_laplace(be64, fp32):
    n = 12
    A = be64.laplace3d(n, n, n)
    N = n ** 3
    rng = np.random.default_rng(5)
    shifts = np.array([0.0, -0.5, -1.3, 2.0, -3.0, 0.7, -6.5, 1.0])  # several indefinite
    k = len(shifts)
    Bh = rng.uniform(-1, 1, (N, k))
    m = be64.map(N)
    B, X = be64.mvec(m, k, Bh), be64.mvec(m, k)
    tol = 1e-5 if fp32 else 1e-10
    it, rr, unconv = be64.minres_shifted(A, shifts, B, X, tol, 2000, fp32=fp32)
    assert unconv == 0, (it, rr)
    Xh = X.get()
    AX = ol.spmv(crs_of(A), 1.0, Xh, 0.0, np.zeros((N, k)))
    true = np.linalg.norm(Bh - AX - shifts * Xh, axis=0) / np.linalg.norm(Bh, axis=0)
    assert np.all(true <= (20 * tol if fp32 else 1e-8)), true
    assert np.all(rr <= tol)
    for o in (B, X, A):
        o.delete()


def test_minres_shifted_large_block(be64):
    """12 systems on a 66^3 Laplacian (287496 rows), all shifts positive:
    the device-resident iteration on a block wider than 8 columns"""
    n = 66
    A = be64.laplace3d(n, n, n)
    N = n ** 3
    rng = np.random.default_rng(9)
    shifts = np.linspace(0.2, 3.0, 12)
    k = len(shifts)
    Bh = rng.uniform(-1, 1, (N, k))
    m = be64.map(N)
    B, X = be64.mvec(m, k, Bh), be64.mvec(m, k)
    tol = 1e-9
    it, rr, unconv = be64.minres_shifted(A, shifts, B, X, tol, 3000)
    assert unconv == 0, (it, rr)
    Xh = X.get()
    AX = ol.spmv(crs_of(A), 1.0, Xh, 0.0, np.zeros((N, k)))
    true = np.linalg.norm(Bh - AX - shifts * Xh, axis=0) / np.linalg.norm(Bh, axis=0)
    assert np.all(true <= 1e-8), true
    assert np.all(rr <= tol)
    for o in (B, X, A):
        o.delete()


def check_found(got, w, which, tol=1e-7):
    """every converged Ritz value is an eigenvalue, and the distinct values
    found are the first distinct eigenvalues in `which` order (a search space
    grown from one vector may hold fewer copies of a multiple eigenvalue)"""
    got = np.asarray(got, float)
    assert np.min(np.abs(got[:, None] - w[None, :]), axis=1).max() <= tol, (got, w)
    key = {"SR": lambda x: x, "LR": lambda x: -x, "LM": lambda x: -np.abs(x), "SM": np.abs}[which]
    dw = []
    for x in sorted(w, key=key):
        if not dw or abs(x - dw[-1]) > 1e-6:
            dw.append(x)
    dg = []
    for x in sorted(got, key=key):
        if not dg or abs(x - dg[-1]) > 1e-6:
            dg.append(x)
    assert np.allclose(dg, dw[:len(dg)], atol=tol), (dg, dw[:len(dg) + 2])


def check_schur(be, A, Q, R, nconv, tol):
    Qh = Q.get()[:, :nconv]
    Rh = R.get()[:nconv, :nconv]
    AQ = ol.spmv(crs_of(A), 1.0, Qh, 0.0, np.zeros_like(Qh))
    res = np.linalg.norm(AQ - Qh @ Rh, axis=0)
    assert np.all(res <= 10 * tol), res
    assert np.allclose(Qh.T @ Qh, np.eye(nconv), atol=1e-10)


INNER = ["MINRES", "GMRES"]


@pytest.mark.parametrize("inner", INNER)
@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("bs", [1, 2, 4])
def test_jada_laplace_smallest(be64, bs, fp32, inner):
    """box of incommensurate sides: the smallest eigenvalues are simple (a
    Krylov space grown from one start vector holds one copy of a multiple
    eigenvalue in exact arithmetic, so a cube is not a fair test)"""
    nx, ny, nz = 9, 11, 13
    A = be64.laplace3d(nx, ny, nz)
    k = 6
    tol = 1e-8
    ev, res, nconv, niter, Q, R = be64.subspacejada(A, k, "SR", tol, block_size=bs, min_bas=12 + bs,
                                                    max_bas=24 + 2 * bs, max_iters=400, inner_max_iters=20,
                                                    inner_fp32=fp32, inner=inner)
    assert nconv >= k, (nconv, niter, res)
    want = laplace_eigs(nx, ny, nz, k)
    assert np.allclose(np.sort(ev[:k]), want, rtol=0, atol=1e-7), (np.sort(ev[:k]), want)
    assert np.all(res[:k] <= tol)
    check_schur(be64, A, Q, R, k, tol)
    for o in (Q, R, A):
        o.delete()


@pytest.mark.parametrize("inner", INNER)
@pytest.mark.parametrize("which", ["SR", "LR", "LM"])
def test_jada_spin_chain(be64, which, inner):
    L = 10
    A = be64.spinchain_sz(L, L // 2)
    dense = ol.rows_to_dense(ol.spin_sz_rows(L, L // 2, 0), A.nrows)
    w = np.linalg.eigvalsh(dense)
    k = 4
    ev, res, nconv, niter, Q, R = be64.subspacejada(A, k, which, 1e-8, block_size=2, min_bas=12, max_bas=24,
                                                    max_iters=400, inner=inner)
    assert nconv >= k, (nconv, niter, res)
    check_found(ev[:k], w, which)
    check_schur(be64, A, Q, R, k, 1e-8)
    for o in (Q, R, A):
        o.delete()


@pytest.mark.parametrize("inner", INNER)
@pytest.mark.parametrize("fp32", [False, True])
def test_jada_spin_chain_L14(be64, fp32, inner):
    """the reference's integration case (drivers/integration_tests.cmake:9-10:
    subspacejada on spinSZ14 to 1e-8), smallest eigenvalues vs dense"""
    L = 14
    A = be64.spinchain_sz(L, L // 2)
    dense = ol.rows_to_dense(ol.spin_sz_rows(L, L // 2, 0), A.nrows)
    w = np.linalg.eigvalsh(dense)
    k = 4
    ev, res, nconv, niter, Q, R = be64.subspacejada(A, k, "SR", 1e-8, block_size=4, min_bas=16, max_bas=32,
                                                    max_iters=200, inner_max_iters=10, inner_fp32=fp32,
                                                    inner=inner)
    assert nconv >= k, (nconv, niter, res)
    check_found(ev[:k], w, "SR")
    check_schur(be64, A, Q, R, k, 1e-8)
    for o in (Q, R, A):
        o.delete()


@pytest.mark.parametrize("fp32", [False, True])
def test_jada_integration_config_spinSZ14(be64, fp32):
    """drivers/integration_tests.cmake:9-10 exactly: phist_Dsubspacejada spinSZ14 I 1 10 SR 1e-8 28 4 16 32
    4 8 -- 10 smallest, symmetric, tol 1e-8, at most 28 outer iterations, blockSize 4, minBas 16,
    maxBas 32, inner MINRES with at most 8 iterations (innerSolvMaxBas 8)"""
    L = 14
    A = be64.spinchain_sz(L, L // 2)
    dense = ol.rows_to_dense(ol.spin_sz_rows(L, L // 2, 0), A.nrows)
    w = np.linalg.eigvalsh(dense)
    k = 10
    ev, res, nconv, niter, Q, R = be64.subspacejada(A, k, "SR", 1e-8, block_size=4, min_bas=16, max_bas=32,
                                                    max_iters=28, inner_max_iters=8, inner_fp32=fp32)
    print(f"spinSZ14 integration config: {nconv} converged in {niter} iterations")
    assert nconv >= k and niter <= 28, (nconv, niter, res)
    check_found(ev[:k], w, "SR")
    check_schur(be64, A, Q, R, k, 1e-8)
    for o in (Q, R, A):
        o.delete()


def _run_defaults(be64, A, nq):
    """phist_Dsubspacejada with the reference's default options (phist_jadaOpts_setDefaults)"""
    import ctypes
    import phist_amd as pa
    opts = pa.JadaOpts()
    be64.L.phist_jadaOpts_setDefaults(ctypes.byref(opts))
    op = pa.LinearOp()
    fl = ctypes.c_int(0)
    be64.L.phist_DlinearOp_wrap_sparseMat(ctypes.byref(op), A.h, ctypes.byref(fl))
    Q, R = be64.mvec(A.domain_map, nq), be64.sdmat(nq, nq)
    ev, res = np.zeros(nq), np.zeros(nq)
    nc, ni = ctypes.c_int(0), ctypes.c_int(0)
    be64.L.phist_Dsubspacejada(ctypes.byref(op), None, opts, Q.h, R.h, ev.ctypes.data_as(ctypes.c_void_p),
                               res.ctypes.data_as(ctypes.c_void_p), ctypes.byref(nc), ctypes.byref(ni),
                               ctypes.byref(fl))
    return opts, fl.value, ev, res, nc.value, Q, R


def test_jada_reference_defaults(be64):
    """The reference's defaults (phist_jadaOpts.cpp: 6 largest-magnitude eigenpairs, symmetry GENERAL,
    GMRES correction solves, tol 1e-12, maxBas 20): on a matrix of this library that is symmetric entry by entry the
    symmetric algorithm runs and finds the closed-form Laplacian eigenvalues; on a nonsymmetric matrix
    the general (Schur) algorithm is not implemented: -99, no work done."""
    import ctypes
    import phist_amd as pa
    A = be64.laplace3d(5, 6, 7)
    opts, fl, ev, res, nc, Q, R = _run_defaults(be64, A, 6)
    assert opts.numEigs == 6 and opts.maxBas == 20 and abs(opts.convTol - 1e-12) < 1e-20  # reference defaults
    assert opts.symmetry != pa.HERMITIAN and opts.innerSolvType == pa.LINSOLV["GMRES"]
    assert fl == 0 and nc >= 6, (fl, nc, res)
    want = laplace_eigs(5, 6, 7, 5 * 6 * 7)[-6:]  # which = LM (the default): the six largest
    assert np.allclose(np.sort(ev[:6]), want, rtol=0, atol=1e-10), (np.sort(ev[:6]), want)
    for o in (Q, R, A):
        o.delete()
    B = be64.read_mm(os.path.join(ol.MATRICES, "Dsprandn25.mm"))  # nonsymmetric
    _, fl, _, _, _, Q, R = _run_defaults(be64, B, 6)
    assert fl == -99
    for o in (Q, R, B):
        o.delete()


@pytest.mark.parametrize("mode", ["0", "2"])
def test_jada_orthogonalisation_paths(be64, mode, monkeypatch):
    """the robust orthogonalisation of the correction block alone (mode 0)
    and after the device-resident one (mode 2, the fallback taken as if a
    Cholesky had failed) reach the same eigenvalues as the default"""
    monkeypatch.setenv("PHIST_AMD_JADA_ORTH", mode)
    nx, ny, nz = 9, 11, 13
    A = be64.laplace3d(nx, ny, nz)
    ev, res, nconv, niter, Q, R = be64.subspacejada(A, 6, "SR", 1e-8, block_size=2, min_bas=14, max_bas=28,
                                                    max_iters=400, inner_max_iters=20)
    assert nconv >= 6, (nconv, niter, res)
    assert np.allclose(np.sort(ev[:6]), laplace_eigs(nx, ny, nz, 6), rtol=0, atol=1e-7)
    check_schur(be64, A, Q, R, 6, 1e-8)
    for o in (Q, R, A):
        o.delete()



def test_jada_invariant_start_space(be64, tmp_path):
    """a diagonal matrix with three distinct values: the Krylov space of the
    start vector is invariant after three steps, so the start-up Arnoldi must
    continue with random orthogonal vectors (the step-by-step robust path)"""
    n = 600
    d = np.array([1.0, 2.0, 3.0])[np.arange(n) % 3]
    f = tmp_path / "diag3.mm"
    with open(f, "w") as fh:
        fh.write("%%MatrixMarket matrix coordinate real general\n")
        fh.write(f"{n} {n} {n}\n")
        for i in range(n):
            fh.write(f"{i + 1} {i + 1} {d[i]:.1f}\n")
    A = be64.read_mm(str(f))
    ev, res, nconv, niter, Q, R = be64.subspacejada(A, 2, "SR", 1e-8, block_size=1, min_bas=8, max_bas=16,
                                                    max_iters=200)
    assert nconv >= 2, (nconv, niter, res)
    assert np.allclose(ev[:2], 1.0, atol=1e-7), ev[:2]
    Qh = Q.get()[:, :2]
    assert np.allclose(Qh.T @ Qh, np.eye(2), atol=1e-10)
    assert np.allclose(d[:, None] * Qh, Qh * 1.0, atol=1e-7)
    for o in (Q, R, A):
        o.delete()


@pytest.mark.parametrize("inner", INNER)
@pytest.mark.parametrize("fp32", [False, True])
def test_jada_laplace_config1_shape(be64, fp32, inner):
    """the CPU-reference JD configuration's solver settings on a smaller box: 8 smallest, blockSize 4,
    minBas 16, maxBas 32, at most 10 inner iterations, tol 1e-8 (incommensurate sides, simple spectrum)"""
    nx, ny, nz = 17, 19, 23
    A = be64.laplace3d(nx, ny, nz)
    ev, res, nconv, niter, Q, R = be64.subspacejada(A, 8, "SR", 1e-8, block_size=4, min_bas=16, max_bas=32,
                                                    max_iters=300, inner_max_iters=10, inner_fp32=fp32, inner=inner)
    print(f"lap {nx}x{ny}x{nz} inner {inner} fp32 {fp32}: {nconv} converged in {niter} iterations")
    assert nconv >= 8, (nconv, niter, res)
    assert np.allclose(np.sort(ev[:8]), laplace_eigs(nx, ny, nz, 8), rtol=0, atol=1e-7)
    check_schur(be64, A, Q, R, 8, 1e-8)
    for o in (Q, R, A):
        o.delete()


def test_jada_xxz24_full_size_both_precisions(be64):
    """the mixed-precision eigensolver workload at full size: Heisenberg chain L=24, S_z = 0
    (2,704,156 rows), 16 smallest eigenpairs to 1e-8 with blocked-GMRES correction solves (at most
    10 iterations, the reference's default solver), FP64 throughout and with FP32 correction
    solves.  Both runs: 16 converged, residuals <= tol, the partial Schur form's invariants
    (||A Q - Q R|| <= 10 tol per column, Q'Q = I) checked with the oracle's product; the mixed run's
    eigenvalues equal the FP64 run's to within the tolerance (a residual of tol moves a symmetric
    eigenvalue by at most tol * ||A||-scale, far less here)."""
    A = be64.spinchain_sz(24, 12)
    assert A.nrows == 2704156
    tol, k = 1e-8, 16
    evs = {}
    for fp32 in (False, True):
        ev, res, nconv, niter, Q, R = be64.subspacejada(A, k, "SR", tol, block_size=4, min_bas=24, max_bas=48,
                                                        max_iters=300, inner="GMRES", inner_max_iters=10,
                                                        inner_fp32=fp32)
        print(f"L=24 fp32={fp32}: {niter} iterations, nconv {nconv}, ev[:3] {ev[:3]}", flush=True)
        assert nconv >= k, (nconv, niter, res)
        assert np.all(res[:k] <= tol), res[:k]
        check_schur(be64, A, Q, R, k, tol)
        evs[fp32] = np.sort(ev[:k])
        for o in (Q, R):
            o.delete()
    scale = np.maximum(1.0, np.abs(evs[False]))
    assert np.all(np.abs(evs[True] - evs[False]) <= 10 * tol * scale), (evs[True], evs[False])
    A.delete()
