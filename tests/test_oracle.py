"""CPU: pins the oracle against the reference's golden vectors, a published
known answer, the reference's own compiled leaf kernels and closed forms.

How each oracle function is pinned:
  kiss (RNG)           Marsaglia's published 64-bit KISS check: the 10^8-th
                       raw draw from the state of phist_random.c:76-90 is
                       1666297717051644203 ("100 million uses of KISS OK").
  crs_from_mm + spmv   A2_precalc_result golden vectors (Dsprandn25, nvec 1/4)
                       and bit-exact agreement with the flang-built builtin
                       dspmvm_* kernels on every fixture matrix.
  dot, nrm2            flang-built ddot_general / dnrm2_general (1 thread).
  times_sdmat(+inplace) flang-built dgemm_sB_generic(_inplace).
  fused_mvsd_mvTmv     flang-built dgemm_sB_augmented_generic.
  gram                 long-double Gram (MvecSdMatTest_def.hpp:523-586 style)
                       and diag(V'V) == dot(V,V).
  cholesky / subst     R'R == M, R(:,perm) upper triangular, detected rank of
                       rank-deficient Gram matrices, R^-1 by backward subst.
gemm_kernels_sC.F90 / gemm_kernels_fused_sCD.F90 need CMake-generated
headers and cannot be built here, hence the Gram is pinned mathematically.
"""
import ctypes
import json
import os

import numpy as np
import pytest

import oracle_lib as ol

eps = np.finfo(np.float64).eps


def test_kiss_known_answer():
    o = ol.oracle()
    s = ol.Kiss()
    o.oracle_kiss_init(ctypes.byref(s))
    o.oracle_kiss_skip(ctypes.byref(s), ctypes.c_uint64(10**8 - 1))
    assert o.oracle_kiss_raw(ctypes.byref(s)) == 1666297717051644203


def test_kiss_doubles_in_range_and_distinct():
    v = ol.kiss_doubles(100000)
    assert v.min() >= -1.0 and v.max() < 1.0
    assert len(np.unique(v)) == len(v)
    assert abs(v.mean()) < 0.01


def test_mvec_random_skip_accounting():
    # two consecutive mvec_random calls: the second starts (n+1)*k draws later
    a, b = ol.mvec_random_sequence([(10, 3), (10, 3)])
    flat = ol.kiss_doubles(2 * 33)
    assert np.array_equal(a.ravel(), flat[:30])
    assert np.array_equal(b.ravel(), flat[33:63])


def golden():
    with open(os.path.join(ol.GOLDEN, "a2_precalc_sprandn25.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("nvec", [1, 4])
def test_spmv_matches_reference_golden_vectors(nvec):
    g = golden()
    A = ol.crs_from_mm(os.path.join(ol.MATRICES, "Dsprandn25.mm"))
    x = np.array([[i + j * 25 for j in range(nvec)] for i in range(25)], float)
    y = ol.spmv(A, 1.0, x, 0.0, np.zeros((25, nvec)), nt=True)
    ref = np.array(g[f"nvec{nvec}"]).reshape(25, nvec)
    assert np.max(np.abs(y - ref)) <= np.sqrt(eps)


def test_spmv_golden_catches_planted_errors():
    """a dropped entry or a transposed operand must fail the golden check"""
    g = golden()
    A = ol.crs_from_mm(os.path.join(ol.MATRICES, "Dsprandn25.mm"))
    x = np.arange(25.0).reshape(25, 1)
    ref = np.array(g["nvec1"]).reshape(25, 1)
    dense = A.dense()
    assert np.max(np.abs(dense @ x - ref)) <= np.sqrt(eps)
    assert np.max(np.abs(dense.T @ x - ref)) > 1e-3
    A2 = ol.Crs(A.nrows, A.ncols, A.row_offset, A.col_idx, A.val.copy())
    A2.val[7] = 0.0
    assert np.max(np.abs(ol.spmv(A2, 1.0, x, 0.0, np.zeros((25, 1))) - ref)) > 1e-3


def test_crs_sort_is_stable_and_keeps_duplicates():
    rows = [([3, 1, 3, 0], [1.0, 2.0, 3.0, 4.0]), ([], []), ([2], [5.0])]
    A = ol.crs_from_rows(3, 4, rows)
    assert list(A.row_offset) == [1, 5, 5, 6]
    assert list(A.col_idx) == [1, 2, 4, 4, 3]
    assert list(A.val) == [4.0, 2.0, 1.0, 3.0, 5.0]


def test_mm_symmetric_mirror(tmp_path):
    f = tmp_path / "s.mm"
    f.write_text("%%MatrixMarket matrix coordinate real symmetric\n3 3 3\n1 1 2.0\n3 1 -1.0\n2 2 4.0\n")
    A = ol.crs_from_mm(str(f))
    assert np.array_equal(A.dense(), np.array([[2.0, 0, -1.0], [0, 4.0, 0], [-1.0, 0, 0]]))


# ---- against the reference's own leaf kernels (flang build, 1 thread) ----

ref = ol.ref_leaf()
needs_ref = pytest.mark.skipif(ref is None, reason="oracle/_ref not built (make -C oracle ref)")


def fortran_spmv(A, nvec, alpha, shifts, x, beta, y):
    """dspmvm_generic(nvec, nlocal, nhalo, ncols, nnz, alpha, row_ptr,
    halo_ptr, col_idx, val, shifts, x, ldx, halo, beta, y, ldy)"""
    I = ctypes.c_int
    L = ctypes.c_int64
    D = ctypes.c_double
    n = A.nrows
    halo_ptr = A.row_offset[1:].copy()  # no halo entries on one process
    halo = np.zeros(max(1, nvec))
    y = np.array(y, copy=True, order="C")
    ref.dspmvm_generic_(ctypes.byref(I(nvec)), ctypes.byref(I(n)), ctypes.byref(I(0)), ctypes.byref(I(A.ncols)),
                        ctypes.byref(L(A.nnz)), ctypes.byref(D(alpha)), ol.p(A.row_offset), ol.p(halo_ptr),
                        ol.p(A.col_idx), ol.p(A.val), ol.p(np.ascontiguousarray(shifts, float)),
                        ol.p(np.ascontiguousarray(x)), ctypes.byref(I(nvec)), ol.p(halo), ctypes.byref(D(beta)),
                        ol.p(y), ctypes.byref(I(nvec)))
    return y


@needs_ref
@pytest.mark.parametrize("mat", ["Dsprandn25", "Dsprandn388", "Dsprandn_nodiag25", "Dsprandsym163", "DsymmMat20",
                                 "DjadaTestMat512", "Dspshift25", "Dspzero25"])
@pytest.mark.parametrize("nvec", [1, 3, 8])
def test_oracle_spmv_bitexact_vs_builtin(mat, nvec, monkeypatch):
    monkeypatch.setenv("OMP_NUM_THREADS", "1")
    A = ol.crs_from_mm(os.path.join(ol.MATRICES, mat + ".mm"))
    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, (A.ncols, nvec))
    y0 = rng.uniform(-1, 1, (A.nrows, nvec))
    sh = rng.uniform(-1, 1, nvec)
    a = ol.spmv(A, 0.7, x, -0.3, y0, shifts=sh)
    b = fortran_spmv(A, nvec, 0.7, sh, x, -0.3, y0)
    assert np.array_equal(a, b)


@needs_ref
@pytest.mark.parametrize("nvec", [1, 5, 8])
def test_oracle_dot_nrm2_vs_builtin(nvec):
    rng = np.random.default_rng(3)
    n = 1000
    v = rng.uniform(-1, 1, (n, nvec))
    w = rng.uniform(-1, 1, (n, nvec))
    out = np.zeros(nvec)
    I = ctypes.c_int
    ref.ddot_general_(ctypes.byref(I(n)), ctypes.byref(I(nvec)), ol.p(v), ctypes.byref(I(nvec)), ol.p(w),
                      ctypes.byref(I(nvec)), ol.p(out))
    assert np.allclose(out, ol.dot(v, w), rtol=1e-13, atol=1e-13)
    nr = np.zeros(nvec)
    ref.dnrm2_general_(ctypes.byref(I(n)), ctypes.byref(I(nvec)), ol.p(v), ctypes.byref(I(nvec)), ol.p(nr))
    assert np.allclose(nr, ol.nrm2(v), rtol=1e-13)


@needs_ref
@pytest.mark.parametrize("k,m", [(4, 4), (8, 3), (5, 7)])
@pytest.mark.parametrize("beta", [0.0, 0.5])
def test_oracle_times_sdmat_vs_builtin(k, m, beta):
    rng = np.random.default_rng(11)
    n = 500
    V = rng.uniform(-1, 1, (n, k))
    C = np.asfortranarray(rng.uniform(-1, 1, (k, m)))
    W = rng.uniform(-1, 1, (n, m))
    I = ctypes.c_int
    D = ctypes.c_double
    Wr = W.copy()
    # dgemm_sB_generic(m=nrows, n=ncolsW, k, alpha, A=V, lda, B=C(k,n), beta, C=W, ldc)
    ref.dgemm_sb_generic_(ctypes.byref(I(n)), ctypes.byref(I(m)), ctypes.byref(I(k)), ctypes.byref(D(0.9)), ol.p(V),
                          ctypes.byref(I(k)), ol.p(C), ctypes.byref(D(beta)), ol.p(Wr), ctypes.byref(I(m)))
    Wo = ol.times_sdmat(0.9, V, C, beta, W)
    assert np.allclose(Wo, Wr, rtol=1e-13, atol=1e-13)


@needs_ref
@pytest.mark.parametrize("k,m", [(6, 6), (8, 5)])
def test_oracle_inplace_vs_builtin(k, m):
    rng = np.random.default_rng(5)
    n = 300
    V = rng.uniform(-1, 1, (n, k))
    M = np.asfortranarray(rng.uniform(-1, 1, (k, m)))
    Vr = V.copy()
    I = ctypes.c_int
    # dgemm_sB_generic_inplace(m=nrows, n=ncols out, k, A=V, lda, B=M)
    ref.dgemm_sb_generic_inplace_(ctypes.byref(I(n)), ctypes.byref(I(m)), ctypes.byref(I(k)), ol.p(Vr),
                                  ctypes.byref(I(k)), ol.p(M))
    assert np.array_equal(ol.times_sdmat_inplace(V, M), Vr[:, :m])


@needs_ref
@pytest.mark.parametrize("k,m", [(4, 4), (8, 3), (5, 7), (16, 16)])
@pytest.mark.parametrize("beta", [0.0, 0.5])
def test_oracle_fused_mvsd_mvTmv_vs_builtin(k, m, beta):
    """fused_mvsd_mvTmv (W = alpha*V*C + beta*W, D = W'W) as the oracle composes it from times_sdmat and
    gram, against the flang-built dgemm_sB_augmented_generic (gemm_kernels_sB_augmented.F90:381-428),
    one thread"""
    rng = np.random.default_rng(13 + k)
    n = 700
    V = rng.uniform(-1, 1, (n, k))
    C = np.asfortranarray(rng.uniform(-1, 1, (k, m)))
    W = rng.uniform(-1, 1, (n, m))
    I = ctypes.c_int
    D = ctypes.c_double
    Wr = W.copy()
    Dr = np.zeros((m, m), order="F")
    # dgemm_sB_augmented_generic(m=nrows, n=ncols W, k, alpha, A=V, lda, B=C(k,n), beta, C=W, ldc, D(n,n))
    ref.dgemm_sb_augmented_generic_(ctypes.byref(I(n)), ctypes.byref(I(m)), ctypes.byref(I(k)), ctypes.byref(D(0.8)),
                                    ol.p(V), ctypes.byref(I(k)), ol.p(C), ctypes.byref(D(beta)), ol.p(Wr),
                                    ctypes.byref(I(m)), ol.p(Dr))
    Wo = ol.times_sdmat(0.8, V, C, beta, W)
    assert np.allclose(Wo, Wr, rtol=1e-13, atol=1e-13)
    Do = ol.gram(1.0, Wo, Wo, 0.0, np.zeros((m, m)))
    assert np.allclose(Do, Dr, rtol=1e-12, atol=1e-12 * n)
    # the test catches a Gram of the old W or of V*C without beta*W
    if beta != 0.0:
        assert not np.allclose(ol.gram(1.0, W, W, 0.0, np.zeros((m, m))), Dr, rtol=1e-6)
        W0 = ol.times_sdmat(0.8, V, C, 0.0, W)
        assert not np.allclose(ol.gram(1.0, W0, W0, 0.0, np.zeros((m, m))), Dr, rtol=1e-6)


def test_oracle_gram_long_double():
    rng = np.random.default_rng(9)
    n, k, m = 2000, 6, 4
    V = rng.uniform(-1, 1, (n, k))
    W = rng.uniform(-1, 1, (n, m))
    C0 = rng.uniform(-1, 1, (k, m))
    got = ol.gram(0.5, V, W, 2.0, C0)
    exact = 0.5 * (V.astype(np.longdouble).T @ W.astype(np.longdouble)) + 2.0 * C0
    bound = 1000 * eps * (np.abs(V).T @ np.abs(W) + np.abs(C0))
    assert np.all(np.abs(got - exact.astype(float)) <= bound)
    # diag of the self Gram is the column dot
    G = ol.gram(1.0, V, V, 0.0, np.zeros((k, k)))
    assert np.allclose(np.diag(G), ol.dot(V, V), rtol=1e-14)
    # operand order: W'V is the transpose of V'W; a dropped row is caught
    G_vw = ol.gram(1.0, V, W, 0.0, np.zeros((k, m)))
    assert np.allclose(ol.gram(1.0, W, V, 0.0, np.zeros((m, k))), G_vw.T, rtol=1e-14)
    short = 0.5 * ol.gram(1.0, V[1:], W[1:], 0.0, np.zeros((k, m))) + 2.0 * C0
    assert not np.all(np.abs(short - exact.astype(float)) <= bound)


def test_oracle_cholesky_full_rank_and_deficient():
    rng = np.random.default_rng(2)
    X = rng.uniform(-1, 1, (50, 6))
    M = X.T @ X
    R, perm, rank = ol.cholesky(M, 10 * eps)
    assert rank == 6
    assert np.allclose(R.T @ R, M, rtol=1e-12, atol=1e-12)
    Rp = R[:, perm]
    assert np.allclose(np.tril(Rp, -1), 0.0)
    # pivot order: first pivot is the largest diagonal
    assert perm[0] == int(np.argmax(np.diag(M)))
    # rank-deficient: two repeated columns
    X2 = np.hstack([X[:, :4], X[:, :2]])
    R2, p2, r2 = ol.cholesky(X2.T @ X2, 10 * eps)
    assert r2 == 4
