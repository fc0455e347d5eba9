"""GPU: restarted blocked GMRES on shifted systems of general matrices (phist_amd_DblockedGMRES_shifted).

Pinned by the true residual of the returned x computed on the host with the oracle's CRS product
(not by the solver's own estimate), by agreement with blocked MINRES on a symmetric problem, and by
the reference's stopping rule (src/krylov/phist_blockedgmres_def.hpp:600-611): ||r||/||r0|| < tol."""
import os

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def true_relres(A, shifts, B, X, X0=None):
    ro, ci, va = A.crs()
    crs = ol.Crs(A.nrows, A.ncols, ro, ci, va.astype(float))
    r = B - ol.spmv(crs, 1.0, X, 0.0, np.zeros_like(X)) - shifts * X
    r0 = B if X0 is None else B - ol.spmv(crs, 1.0, X0, 0.0, np.zeros_like(X0)) - shifts * X0
    return np.linalg.norm(r, axis=0) / np.linalg.norm(r0, axis=0)


# restart <= 32 with <= 64 systems: per-system Gram-Schmidt sweeps (4, 64 and 1 lanes per row);
# restart 40: Grams and row mixes of the whole basis
@pytest.mark.parametrize("restart,nsys", [(30, 4), (8, 4), (40, 4), (12, 40), (20, 1)])
def test_gmres_nonsymmetric_banded(be64, restart, nsys):
    A = be64.banded_random(20000, 16, 300, 20.0, 5)  # nonsymmetric, diagonally dominant
    n = A.nrows
    shifts = np.array([0.0, 1.5, -3.0, 6.0]) if nsys == 4 else np.linspace(-3.0, 6.0, nsys)
    Bh = np.random.default_rng(1).uniform(-1, 1, (n, nsys))
    B, X = be64.mvec(A.domain_map, nsys, Bh), be64.mvec(A.domain_map, nsys)
    it, rr, unconv = be64.gmres_shifted(A, shifts, B, X, 1e-10, 500, restart=restart)
    assert unconv == 0, (it, rr)
    got = true_relres(A, shifts, Bh, X.get())
    assert np.all(got < 1e-9), got
    assert np.all(it > 0)
    for o in (B, X, A):
        o.delete()


def test_gmres_mm_matrix_with_restarts(be64):
    A = be64.read_mm(os.path.join(ol.MATRICES, "Dsprandn388.mm"))
    n = A.nrows
    shifts = np.array([20.0, 30.0, 45.0])  # the spectrum has real parts in [-13.2, 41.9]: A + sI is
    # then far from singular (GMRES(10) stagnates for shifts inside the spectrum, as scipy's does)
    Bh = np.random.default_rng(2).uniform(-1, 1, (n, 3))
    B, X = be64.mvec(A.domain_map, 3, Bh), be64.mvec(A.domain_map, 3)
    it, rr, unconv = be64.gmres_shifted(A, shifts, B, X, 1e-11, 2000, restart=10)
    got = true_relres(A, shifts, Bh, X.get())
    assert unconv == 0 and np.all(got < 1e-10), (it, rr, got)
    for o in (B, X, A):
        o.delete()


def test_gmres_matches_minres_on_symmetric(be64):
    A = be64.laplace3d(14, 13, 12)
    n = A.nrows
    shifts = np.array([0.3, 1.0, 2.5, 0.05])
    Bh = np.random.default_rng(3).uniform(-1, 1, (n, 4))
    B = be64.mvec(A.domain_map, 4, Bh)
    Xg, Xm = be64.mvec(A.domain_map, 4), be64.mvec(A.domain_map, 4)
    itg, rrg, ug = be64.gmres_shifted(A, shifts, B, Xg, 1e-11, 2000, restart=40)
    itm, rrm, um = be64.minres_shifted(A, shifts, B, Xm, 1e-11, 2000)
    assert ug == 0 and um == 0
    xg, xm = Xg.get(), Xm.get()
    assert np.all(np.linalg.norm(xg - xm, axis=0) <= 1e-8 * np.linalg.norm(xm, axis=0))
    for o in (B, Xg, Xm, A):
        o.delete()


def test_gmres_fp32(be64):
    A = be64.banded_random(30000, 12, 500, 16.0, 7)
    n = A.nrows
    shifts = np.array([0.0, 2.0])
    Bh = np.random.default_rng(4).uniform(-1, 1, (n, 2))
    B, X = be64.mvec(A.domain_map, 2, Bh), be64.mvec(A.domain_map, 2)
    it, rr, unconv = be64.gmres_shifted(A, shifts, B, X, 1e-5, 300, restart=20, fp32=True)
    got = true_relres(A, shifts, Bh, X.get())
    assert unconv == 0 and np.all(got < 5e-5), (it, got)
    for o in (B, X, A):
        o.delete()


def test_gmres_initial_guess_and_zero_rhs(be64):
    A = be64.banded_random(5000, 8, 100, 12.0, 9)
    n = A.nrows
    shifts = np.array([1.0, 1.0, -2.0])
    rng = np.random.default_rng(5)
    Bh = rng.uniform(-1, 1, (n, 3))
    Bh[:, 2] = 0.0
    B, X = be64.mvec(A.domain_map, 3, Bh), be64.mvec(A.domain_map, 3)
    it, rr, unconv = be64.gmres_shifted(A, shifts, B, X, 1e-12, 500)
    x1 = X.get()
    assert unconv == 0 and np.all(x1[:, 2] == 0.0) and it[2] == 0
    # start again from the converged solution (columns 0, 1) and from a perturbed one (column 1)
    X0 = x1.copy()
    X0[:, 1] += 1e-3 * rng.uniform(-1, 1, n)
    X.set(X0)
    it2, rr2, unconv2 = be64.gmres_shifted(A, shifts, B, X, 1e-6, 500)
    # the tolerance is relative to the initial residual of this solve (phist_blockedgmres_def.hpp:604);
    # column 0 starts at the attainable accuracy already, so only its residual w.r.t. b is checked
    x2 = X.get()
    rel1 = true_relres(A, shifts[1:2], Bh[:, 1:2], x2[:, 1:2], X0[:, 1:2])
    rel0 = true_relres(A, shifts[:1], Bh[:, :1], x2[:, :1])
    assert unconv2 == 0 and rel1[0] < 1e-5 and rel0[0] < 1e-11, (it2, rel1, rel0)
    assert it2[1] < it[1]  # a good start needs fewer steps
    for o in (B, X, A):
        o.delete()


def test_gmres_rejects_bad_input(be64):
    import phist_amd
    A = be64.laplace3d(5, 5, 5)
    B, X = be64.mvec(A.domain_map, 2), be64.mvec(A.domain_map, 3)
    with pytest.raises(phist_amd.PhistError):
        be64.gmres_shifted(A, [0.0, 1.0], B, X, 1e-8, 10)
    for o in (B, X, A):
        o.delete()


@pytest.mark.parametrize("restart,nsys,tol", [(30, 4, 1e-10), (8, 4, 1e-10), (12, 40, 1e-9), (20, 1, 1e-10),
                                              (30, 4, 1e-3)])
def test_gmres_device_givens_matches_host(be64, restart, nsys, tol):
    """the Givens rotations on the device (the host one Arnoldi step ahead) and on the host give the same
    iteration counts (within one) and solutions to rounding, in FP64 and in FP32 on the cached copy, with
    systems converging at different steps and across restarts; both reach the tolerance by the true
    residual"""
    import ctypes
    fl = ctypes.c_int(0)
    A = be64.banded_random(20000, 16, 300, 20.0, 5)
    n = A.nrows
    shifts = np.array([0.0, 1.5, -3.0, 6.0]) if nsys == 4 else np.linspace(-3.0, 6.0, nsys)
    Bh = np.random.default_rng(7).uniform(-1, 1, (n, nsys))
    for fp32 in (False, True):
        t = max(tol, 1e-5) if fp32 else tol
        res = {}
        try:
            for dev in (0, 1):
                be64.L.phist_amd_set_gmres_device(dev, ctypes.byref(fl))
                assert fl.value == 0
                B, X = be64.mvec(A.domain_map, nsys, Bh), be64.mvec(A.domain_map, nsys)
                it, rr, unconv = be64.gmres_shifted(A, shifts, B, X, t, 500, restart=restart, fp32=fp32)
                res[dev] = (np.array(it), unconv, X.get())
                B.delete()
                X.delete()
        finally:
            be64.L.phist_amd_set_gmres_device(1, ctypes.byref(fl))
        (it0, u0, x0), (it1, u1, x1) = res[0], res[1]
        assert u0 == 0 and u1 == 0
        assert np.all(np.abs(it1 - it0) <= 1), (it0, it1)
        got = true_relres(A, shifts, Bh, x1)
        assert np.all(got < 2 * t), got
        # the same steps give the same solution to rounding; a convergence decided one step apart
        # (a last-bit difference of the rotations at the threshold) leaves both within the tolerance
        scale = np.linalg.norm(x0, axis=0)
        bound = np.where(it1 == it0, 1e-4 if fp32 else 1e-8, 20 * t)
        assert np.all(np.linalg.norm(x1 - x0, axis=0) <= bound * scale), (it0, it1)
    A.delete()
