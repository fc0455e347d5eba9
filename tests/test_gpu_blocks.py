"""GPU parity: block-vector BLAS-1, random fills, tall-skinny GEMMs and the
CholQR / SVQB / orthog paths, through the C ABI, against the oracle.

Bars: KISS draws bit-exact (they are integer work); FP64 results within
1e-12 relative to the oracle; FP32 within 1e-5; factorisations by their
invariants (Q'Q = I, Q*R = V, detected rank) as PHIST's own QR tests do
(test/kernels/QR_Test_def.hpp:104-222, test/core/OrthogTest_def.hpp).
"""
import ctypes

import numpy as np
import pytest

import oracle_lib as ol
import phist_amd

pytestmark = pytest.mark.gpu
RTOL = {np.float64: 1e-12, np.float32: 1e-5}


def rel(got, ref):
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300))


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@pytest.fixture(params=[np.float64, np.float32], ids=["D", "S"])
def be(request, be64, be32):
    return be64 if request.param == np.float64 else be32


@pytest.mark.parametrize("n,nvec", [(1, 1), (100, 1), (1000, 3), (4097, 8), (777, 12), (513, 32), (300, 65)])
def test_dot_norm(be, n, nvec):
    rng = np.random.default_rng(n + nvec)
    V = rng.uniform(-1, 1, (n, nvec))
    W = rng.uniform(-1, 1, (n, nvec))
    m = be.map(n)
    v, w = be.mvec(m, nvec, V), be.mvec(m, nvec, W)
    Vt, Wt = V.astype(be.dtype).astype(float), W.astype(be.dtype).astype(float)
    d = be.dot(v, w)
    scale = ol.dot(np.abs(Vt), np.abs(Wt))
    assert np.all(np.abs(d - ol.dot(Vt, Wt)) <= 4 * RTOL[be.dtype.type] * scale + 1e-300)
    assert np.allclose(be.norm2(v), ol.nrm2(Vt), rtol=4 * RTOL[be.dtype.type])
    for o in (v, w):
        o.delete()


def test_dot_closed_form(be64):
    n = 123457
    m = be64.map(n)
    v = be64.mvec(m, 2)
    v.put_value(1.0)
    assert np.array_equal(be64.dot(v, v), [float(n), float(n)])
    v.delete()


@pytest.mark.parametrize("nvec", [1, 3, 4, 9])
def test_axpby_family(be, nvec):
    n = 1001
    rng = np.random.default_rng(nvec)
    X, Y = rng.uniform(-1, 1, (n, nvec)), rng.uniform(-1, 1, (n, nvec))
    a = rng.uniform(-1, 1, nvec)
    m = be.map(n)
    x, y = be.mvec(m, nvec, X), be.mvec(m, nvec, Y)
    Xt, Yt = X.astype(be.dtype).astype(float), Y.astype(be.dtype).astype(float)
    at = a.astype(be.dtype).astype(float)
    tol = RTOL[be.dtype.type]
    be.call("mvec_add_mvec", be.s(0.3), x.h, be.s(-2.0), y.h)
    Yt = 0.3 * Xt - 2.0 * Yt
    assert rel(y.get().astype(float), Yt) <= tol
    A = np.ascontiguousarray(a, be.dtype)
    be.call("mvec_vadd_mvec", ptr(A), x.h, be.s(0.5), y.h)
    Yt = at * Xt + 0.5 * Yt
    assert rel(y.get().astype(float), Yt) <= tol
    be.call("mvec_scale", y.h, be.s(3.0))
    Yt = 3.0 * Yt
    assert rel(y.get().astype(float), Yt) <= tol
    be.call("mvec_vscale", y.h, ptr(A))
    Yt = at * Yt
    assert rel(y.get().astype(float), Yt) <= tol
    be.call("mvec_times_mvec_elemwise", be.s(2.0), x.h, y.h)
    Yt = 2.0 * Xt * Yt
    assert rel(y.get().astype(float), Yt) <= 4 * tol
    be.call("mvec_add_mvec", be.s(1.0), x.h, be.s(0.0), y.h)  # copy
    assert np.array_equal(y.get(), X.astype(be.dtype))
    y.put_value(-7.5)
    assert np.all(y.get() == -7.5)
    nrm = np.zeros(nvec, be.dtype)
    be.call("mvec_normalize", x.h, ptr(nrm))
    assert np.allclose(nrm, np.linalg.norm(Xt, axis=0), rtol=4 * tol)
    assert np.allclose(np.linalg.norm(x.get().astype(float), axis=0), 1.0, rtol=4 * tol)
    for o in (x, y):
        o.delete()


def test_random_is_kiss_bitexact(be64):
    """mvec_random / sdMat_random draw PHIST's KISS stream (same values,
    same skip accounting as kernels_common_impl_def.h:256-405)"""
    iflag = ctypes.c_int(0)
    be64.L.phist_amd_kiss_reset(ctypes.byref(iflag))
    shapes = [(1000, 3), (257, 1), (64, 8)]
    m = [be64.map(n) for n, _ in shapes]
    got = []
    for (n, k), mp in zip(shapes, m):
        v = be64.mvec(mp, k)
        v.random()
        got.append(v.get())
        v.delete()
    want = ol.mvec_random_sequence(shapes)
    for g, w in zip(got, want):
        assert np.array_equal(g, w)
    # an sdMat draw continues the stream column-major
    s = be64.sdmat(3, 2)
    be64.call("sdMat_random", s.h)
    flat = ol.kiss_doubles(sum(n * k + k for n, k in shapes) + 6)
    tail = flat[-6:]
    assert np.array_equal(s.get(), tail.reshape(2, 3).T)
    s.delete()


@pytest.mark.parametrize("m,k", [(1, 1), (4, 4), (8, 8), (3, 5), (16, 16), (32, 32), (32, 4), (2, 40), (64, 8),
                                 (7, 7), (130, 3), (16, 24), (17, 33), (40, 40), (64, 48), (20, 16), (19, 4),
                                 (31, 7), (33, 5), (9, 8), (1, 3), (15, 2)])
@pytest.mark.parametrize("beta", [0.0, -0.5])
def test_gram(be, m, k, beta):
    n = 5003
    rng = np.random.default_rng(m * 100 + k)
    V, W = rng.uniform(-1, 1, (n, m)), rng.uniform(-1, 1, (n, k))
    C0 = rng.uniform(-1, 1, (m, k))
    mp = be.map(n)
    v, w = be.mvec(mp, m, V), be.mvec(mp, k, W)
    C = be.sdmat(m, k, C0)
    be.gram(0.8, v, w, beta, C)
    Vt, Wt, Ct = (a.astype(be.dtype).astype(float) for a in (V, W, C0))
    want = ol.gram(0.8, Vt, Wt, beta, Ct)
    bound = 8 * RTOL[be.dtype.type] * (0.8 * np.abs(Vt).T @ np.abs(Wt) + abs(beta) * np.abs(Ct))
    assert np.all(np.abs(C.get().astype(float) - want) <= bound)
    for o in (v, w, C):
        o.delete()


@pytest.mark.parametrize("k", [2, 4, 8, 16, 24, 32, 40, 64])
def test_gram_self_mfma(be, k):
    """V'V on the matrix cores (upper tiles computed, mirrored; 2, 4 and 8 columns by row sets, every
    lane loading a column pair) vs the oracle; 4099 rows leave a partial last step"""
    n = 4099
    V = np.random.default_rng(k).uniform(-1, 1, (n, k))
    mp = be.map(n)
    v = be.mvec(mp, k, V)
    C0 = np.random.default_rng(1).uniform(-1, 1, (k, k))
    C = be.sdmat(k, k, C0)
    be.gram(1.5, v, v, 0.5, C)
    Vt, Ct = V.astype(be.dtype).astype(float), C0.astype(be.dtype).astype(float)
    want = ol.gram(1.5, Vt, Vt, 0.5, Ct)
    bound = 8 * RTOL[be.dtype.type] * (1.5 * np.abs(Vt).T @ np.abs(Vt) + 0.5 * np.abs(Ct))
    got = C.get().astype(float)
    assert np.all(np.abs(got - want) <= bound)
    for o in (v, C):
        o.delete()


@pytest.mark.parametrize("j0,j1,k", [(2, 20, 4), (0, 18, 4), (1, 5, 3), (3, 22, 8)])
def test_gram_narrow_views(be, j0, j1, k):
    """Q(:, j0..j1)' W with Q a column view of an odd-width block (rows not 16-byte aligned), as in the
    projected correction equations"""
    n = 6007
    rng = np.random.default_rng(j0 * 31 + j1)
    X, Wh = rng.uniform(-1, 1, (n, 23)), rng.uniform(-1, 1, (n, k))
    mp = be.map(n)
    x, w = be.mvec(mp, 23, X), be.mvec(mp, k, Wh)
    xv = x.view(j0, j1)
    m = j1 - j0 + 1
    C = be.sdmat(m, k)
    be.gram(1.0, xv, w, 0.0, C)
    Vt, Wt = X[:, j0:j1 + 1].astype(be.dtype).astype(float), Wh.astype(be.dtype).astype(float)
    want = ol.gram(1.0, Vt, Wt, 0.0, np.zeros((m, k)))
    bound = 8 * RTOL[be.dtype.type] * (np.abs(Vt).T @ np.abs(Wt))
    assert np.all(np.abs(C.get().astype(float) - want) <= bound)
    for o in (xv, x, w, C):
        o.delete()


@pytest.mark.parametrize("j0,k", [(2, 4), (4, 8), (6, 2), (3, 4)])
def test_gram_self_narrow_views(be, j0, k):
    """V'V of a column view of a 12-column block (row stride 12): aligned views take the row-set kernel,
    the view at an odd offset the general path; both vs the oracle"""
    n = 3333
    X = np.random.default_rng(j0 + 10 * k).uniform(-1, 1, (n, 12))
    mp = be.map(n)
    x = be.mvec(mp, 12, X)
    xv = x.view(j0, j0 + k - 1)
    C = be.sdmat(k, k)
    be.gram(1.0, xv, xv, 0.0, C)
    Vt = X[:, j0:j0 + k].astype(be.dtype).astype(float)
    want = ol.gram(1.0, Vt, Vt, 0.0, np.zeros((k, k)))
    bound = 8 * RTOL[be.dtype.type] * (np.abs(Vt).T @ np.abs(Vt))
    assert np.all(np.abs(C.get().astype(float) - want) <= bound)
    for o in (xv, x, C):
        o.delete()


def test_gram_self_and_views(be64):
    n = 2000
    X = np.random.default_rng(4).uniform(-1, 1, (n, 12))
    mp = be64.map(n)
    x = be64.mvec(mp, 12, X)
    xv = x.view(3, 8)
    C = be64.sdmat(6, 6)
    be64.gram(1.0, xv, xv, 0.0, C)
    assert rel(C.get(), ol.gram(1.0, X[:, 3:9], X[:, 3:9], 0.0, np.zeros((6, 6)))) <= 1e-12
    for o in (xv, x, C):
        o.delete()


@pytest.mark.parametrize("k,m", [(1, 1), (4, 4), (8, 8), (5, 3), (16, 16), (32, 32), (40, 6), (3, 33), (8, 33),
                                 (12, 20), (64, 64), (64, 17), (100, 40), (9, 70)])
@pytest.mark.parametrize("beta", [0.0, 0.5])
def test_times_sdmat(be, k, m, beta):
    n = 3001
    rng = np.random.default_rng(k * 7 + m)
    V, C0, W0 = rng.uniform(-1, 1, (n, k)), rng.uniform(-1, 1, (k, m)), rng.uniform(-1, 1, (n, m))
    mp = be.map(n)
    v, w = be.mvec(mp, k, V), be.mvec(mp, m, W0)
    C = be.sdmat(k, m, C0)
    be.times_sdmat(-1.2, v, C, beta, w)
    Vt, Ct, Wt = (a.astype(be.dtype).astype(float) for a in (V, C0, W0))
    want = ol.times_sdmat(-1.2, Vt, Ct, beta, Wt)
    assert rel(w.get().astype(float), want) <= 4 * RTOL[be.dtype.type] * max(1, k) ** 0.5
    for o in (v, w, C):
        o.delete()


@pytest.mark.parametrize("k,m", [(1, 1), (4, 4), (8, 8), (6, 3), (16, 16), (32, 32), (33, 17), (64, 64), (24, 9),
                                 (40, 40), (70, 20)])
def test_times_sdmat_inplace(be, k, m):
    n = 2049
    rng = np.random.default_rng(k + 3 * m)
    V, M0 = rng.uniform(-1, 1, (n, k)), rng.uniform(-1, 1, (k, m))
    mp = be.map(n)
    v = be.mvec(mp, k, V)
    M = be.sdmat(k, m, M0)
    be.times_sdmat_inplace(v, M)
    Vt, Mt = V.astype(be.dtype).astype(float), M0.astype(be.dtype).astype(float)
    got = v.get().astype(float)
    assert rel(got[:, :m], ol.times_sdmat_inplace(Vt, Mt)) <= 4 * RTOL[be.dtype.type] * max(1, k) ** 0.5
    assert np.array_equal(got[:, m:], Vt[:, m:])  # columns past m untouched
    for o in (v, M):
        o.delete()


@pytest.mark.parametrize("j0,k,m", [(1, 17, 9), (2, 16, 16), (3, 12, 12), (0, 8, 8)])
def test_times_sdmat_views(be, j0, k, m):
    """row-mixing on column views (unaligned rows take the scalar-load path)"""
    n, ncol = 1537, 20
    rng = np.random.default_rng(j0 + k)
    X, C0, M0 = rng.uniform(-1, 1, (n, ncol)), rng.uniform(-1, 1, (k, m)), rng.uniform(-1, 1, (k, m))
    mp = be.map(n)
    x = be.mvec(mp, ncol, X)
    w = be.mvec(mp, m)
    xv = x.view(j0, j0 + k - 1)
    C, M = be.sdmat(k, m, C0), be.sdmat(k, m, M0)
    Xt, Ct, Mt = (a.astype(be.dtype).astype(float) for a in (X, C0, M0))
    tol = 4 * RTOL[be.dtype.type] * k ** 0.5
    be.times_sdmat(1.0, xv, C, 0.0, w)
    assert rel(w.get().astype(float), Xt[:, j0:j0 + k] @ Ct) <= tol
    be.times_sdmat_inplace(xv, M)
    got = x.get().astype(float)
    want = Xt.copy()
    want[:, j0:j0 + m] = Xt[:, j0:j0 + k] @ Mt
    assert rel(got, want) <= tol
    assert np.array_equal(got[:, :j0], Xt[:, :j0]) and np.array_equal(got[:, j0 + m:], Xt[:, j0 + m:])
    for o in (xv, x, w, C, M):
        o.delete()


def test_times_sdmat_add_times_sdmat(be64):
    n, k, m = 1000, 5, 4
    rng = np.random.default_rng(8)
    V, C0, W0, D0 = (rng.uniform(-1, 1, s) for s in [(n, k), (k, m), (n, m), (m, m)])
    mp = be64.map(n)
    v, w = be64.mvec(mp, k, V), be64.mvec(mp, m, W0)
    C, D = be64.sdmat(k, m, C0), be64.sdmat(m, m, D0)
    be64.call("mvec_times_sdMat_add_mvec_times_sdMat", v.h, C.h, w.h, D.h)
    assert rel(w.get(), W0 @ D0 + V @ C0) <= 1e-12
    for o in (v, w, C, D):
        o.delete()


def test_fused_mvsd_and_mvsdi(be64):
    n = 4000
    rng = np.random.default_rng(12)
    V, W0, C0 = rng.uniform(-1, 1, (n, 6)), rng.uniform(-1, 1, (n, 4)), rng.uniform(-1, 1, (6, 4))
    mp = be64.map(n)
    v, w = be64.mvec(mp, 6, V), be64.mvec(mp, 4, W0)
    C = be64.sdmat(6, 4, C0)
    WtW = be64.sdmat(4, 4)
    be64.call("fused_mvsd_mvTmv", be64.s(-1.0), v.h, C.h, be64.s(1.0), w.h, WtW.h)
    Wn = ol.times_sdmat(-1.0, V, C0, 1.0, W0)
    assert rel(w.get(), Wn) <= 1e-12
    assert rel(WtW.get(), ol.gram(1.0, Wn, Wn, 0.0, np.zeros((4, 4)))) <= 1e-12
    N0 = rng.uniform(-1, 1, (4, 4))
    Nm = be64.sdmat(4, 4, N0)
    D = be64.sdmat(6, 4, np.ones((6, 4)))
    be64.call("fused_mvsdi_mvTmv", be64.s(2.0), v.h, w.h, Nm.h, be64.s(0.5), D.h)
    Wi = ol.times_sdmat_inplace(Wn, N0)
    assert rel(w.get(), Wi) <= 1e-12
    assert rel(D.get(), ol.gram(2.0, V, Wi, 0.5, np.ones((6, 4)))) <= 1e-12
    for o in (v, w, C, WtW, Nm, D):
        o.delete()


def gram_bound(dtype, alpha, A, B, beta, C0):
    return 8 * RTOL[dtype] * (abs(alpha) * np.abs(A).T @ np.abs(B) + abs(beta) * np.abs(C0)) + 1e-300


@pytest.mark.parametrize("k,m", [(8, 8), (16, 16), (12, 20), (32, 32), (64, 32), (9, 31), (40, 40), (4, 8)])
def test_fused_mvsd_mvTmv_shapes(be, k, m):
    """W = alpha*V*C + beta*W and W'W in one sweep (8 <= k <= 64, 8 <= m <= 32), else two passes"""
    n = 3001
    rng = np.random.default_rng(k * 13 + m)
    V, W0, C0 = rng.uniform(-1, 1, (n, k)), rng.uniform(-1, 1, (n, m)), rng.uniform(-1, 1, (k, m))
    mp = be.map(n)
    v, w = be.mvec(mp, k, V), be.mvec(mp, m, W0)
    C = be.sdmat(k, m, C0)
    WtW = be.sdmat(m, m, np.full((m, m), 7.0))
    be.call("fused_mvsd_mvTmv", be.s(-0.5), v.h, C.h, be.s(1.5), w.h, WtW.h)
    Vt, Wt, Ct = (a.astype(be.dtype).astype(float) for a in (V, W0, C0))
    Wn = ol.times_sdmat(-0.5, Vt, Ct, 1.5, Wt)
    got_w = w.get().astype(float)
    assert rel(got_w, Wn) <= 4 * RTOL[be.dtype.type] * k ** 0.5
    want = ol.gram(1.0, got_w, got_w, 0.0, np.zeros((m, m)))
    assert np.all(np.abs(WtW.get().astype(float) - want) <= gram_bound(be.dtype.type, 1.0, got_w, got_w, 0.0, want))
    for o in (v, w, C, WtW):
        o.delete()


@pytest.mark.parametrize("nv,k,m", [(8, 8, 8), (16, 16, 16), (5, 24, 12), (32, 32, 32), (40, 16, 16), (0, 16, 16),
                                    (0, 24, 24), (0, 32, 32), (0, 24, 16), (0, 8, 8)])
def test_fused_mvsdi_mvTmv_shapes(be, nv, k, m):
    """W(:,1:m) = W*C in place and D = alpha*V'W + beta*D in one sweep; nv = 0 is the
    self product V == W (one sweep when m == k)"""
    n = 2777
    rng = np.random.default_rng(nv * 100 + k * 7 + m)
    W0, C0 = rng.uniform(-1, 1, (n, k)), rng.uniform(-1, 1, (k, m))
    mp = be.map(n)
    w = be.mvec(mp, k, W0)
    C = be.sdmat(k, m, C0)
    Wt, Ct = W0.astype(be.dtype).astype(float), C0.astype(be.dtype).astype(float)
    self_ = nv == 0
    if self_:
        v, Vt, nvv = w, None, k
    else:
        V = rng.uniform(-1, 1, (n, nv))
        v, Vt, nvv = be.mvec(mp, nv, V), V.astype(be.dtype).astype(float), nv
    D0 = rng.uniform(-1, 1, (nvv, m))
    D = be.sdmat(nvv, m, D0)
    Dt = D0.astype(be.dtype).astype(float)
    be.call("fused_mvsdi_mvTmv", be.s(2.0), v.h, w.h, C.h, be.s(0.5), D.h)
    got_w = w.get().astype(float)
    want_w = Wt.copy()
    want_w[:, :m] = Wt @ Ct
    assert rel(got_w, want_w) <= 4 * RTOL[be.dtype.type] * k ** 0.5
    left = got_w if self_ else Vt
    want = ol.gram(2.0, left, got_w[:, :m], 0.5, Dt)
    bound = gram_bound(be.dtype.type, 2.0, left, got_w[:, :m], 0.5, Dt)
    assert np.all(np.abs(D.get().astype(float) - want) <= bound)
    for o in ((w, C, D) if self_ else (v, w, C, D)):
        o.delete()


@pytest.mark.parametrize("k", [1, 2, 4, 8, 16, 32])
def test_chol_qr(be64, k):
    n = 3000
    V0 = np.random.default_rng(k).uniform(-1, 1, (n, k))
    mp = be64.map(n)
    v = be64.mvec(mp, k, V0)
    R = be64.sdmat(k, k)
    rc = be64.call("chol_QR", v.h, R.h, allow=(0,))
    assert rc == 0
    Q, Rm = v.get(), R.get()
    assert np.max(np.abs(Q.T @ Q - np.eye(k))) <= 1e-12 * max(1, k)
    assert np.max(np.abs(Q @ Rm - V0)) <= 1e-12 * np.max(np.abs(V0)) * k
    for o in (v, R):
        o.delete()


def test_chol_qr_rank_deficient(be64):
    n, k = 2000, 6
    X = np.random.default_rng(1).uniform(-1, 1, (n, 4))
    V0 = np.hstack([X, X[:, :1] + X[:, 1:2], 2 * X[:, 3:4]])
    mp = be64.map(n)
    v = be64.mvec(mp, k, V0)
    R = be64.sdmat(k, k)
    iflag = ctypes.c_int(0)
    be64.fn("chol_QR")(v.h, R.h, ctypes.byref(iflag))
    assert iflag.value == 2  # ncols - rank
    Q = v.get()
    assert np.max(np.abs(Q.T @ Q - np.eye(k))) <= 1e-10
    for o in (v, R):
        o.delete()


@pytest.mark.parametrize("k", [3, 8])
def test_svqb(be64, k):
    n = 2500
    V0 = np.random.default_rng(k + 50).uniform(-1, 1, (n, k)) * np.logspace(0, 3, k)
    mp = be64.map(n)
    v = be64.mvec(mp, k, V0)
    B = be64.sdmat(k, k)
    e = np.zeros(k)
    iflag = ctypes.c_int(0)
    be64.fn("svqb")(v.h, B.h, ptr(e), ctypes.byref(iflag))
    assert iflag.value == 0
    Q = v.get()
    assert np.max(np.abs(Q.T @ Q - np.eye(k))) <= 1e-11
    assert np.allclose(V0 @ B.get(), Q, atol=1e-11)
    assert np.allclose(e, np.linalg.norm(V0, axis=0), rtol=1e-12)
    for o in (v, B):
        o.delete()


@pytest.mark.parametrize("m,k", [(0, 4), (4, 4), (8, 2), (16, 8)])
def test_orthog(be64, m, k):
    n = 3000
    rng = np.random.default_rng(m * 10 + k)
    mp = be64.map(n)
    V = None
    Vq = np.zeros((n, 0))
    if m:
        Vq, _ = np.linalg.qr(rng.uniform(-1, 1, (n, m)))
        V = be64.mvec(mp, m, Vq)
    W0 = rng.uniform(-1, 1, (n, k))
    w = be64.mvec(mp, k, W0)
    R1 = be64.sdmat(k, k)
    R2 = be64.sdmat(max(m, 1), k)
    rank = ctypes.c_int(0)
    iflag = ctypes.c_int(0)
    be64.fn("orthog")(V.h if V else None, w.h, None, R1.h, R2.h if m else None, ctypes.c_int(3), ctypes.byref(rank),
                      ctypes.byref(iflag))
    assert iflag.value == 0 and rank.value == m + k
    Q = w.get()
    assert np.max(np.abs(Q.T @ Q - np.eye(k))) <= 1e-12 * 10
    if m:
        assert np.max(np.abs(Vq.T @ Q)) <= 1e-12 * 10
        recon = Q @ R1.get() + Vq @ R2.get()
    else:
        recon = Q @ R1.get()
    assert np.max(np.abs(recon - W0)) <= 1e-11
    for o in [w, R1, R2] + ([V] if V else []):
        o.delete()


def test_views_and_blocks(be64):
    n = 500
    X = np.random.default_rng(0).uniform(-1, 1, (n, 10))
    mp = be64.map(n)
    x = be64.mvec(mp, 10, X)
    b = be64.mvec(mp, 3)
    be64.call("mvec_get_block", x.h, b.h, ctypes.c_int(2), ctypes.c_int(4))
    assert np.array_equal(b.get(), X[:, 2:5])
    b.put_value(1.0)
    be64.call("mvec_set_block", x.h, b.h, ctypes.c_int(7), ctypes.c_int(9))
    Y = X.copy()
    Y[:, 7:10] = 1.0
    assert np.array_equal(x.get(), Y)
    # view of a view and re-pointing an existing view handle
    v1 = x.view(2, 8)
    v2 = v1.view(1, 3)
    assert np.array_equal(v2.get(), Y[:, 3:6])
    h = v2.h
    be64.call("mvec_view_block", x.h, ctypes.byref(h), ctypes.c_int(0), ctypes.c_int(0))
    assert np.array_equal(be64_mvec_get(be64, h, n, 1), Y[:, 0:1])
    # a non-view passed as *Vblock is rejected (builtin: PHIST_BAD_CAST)
    hb = b.h
    iflag = ctypes.c_int(0)
    be64.fn("mvec_view_block")(x.h, ctypes.byref(hb), ctypes.c_int(0), ctypes.c_int(1), ctypes.byref(iflag))
    assert iflag.value == -88
    for o in (v2, v1, b, x):
        o.delete()


def be64_mvec_get(be, h, n, k):
    a = np.empty((n, k))
    be.call("mvec_get_data", h, ptr(a), ctypes.c_int32(k), ctypes.c_int(1))
    return a


def test_host_mirror_roundtrip(be64):
    """extract_view + from_device/to_device semantics (PHIST's GPU contract)"""
    n, k = 64, 3
    mp = be64.map(n)
    v = be64.mvec(mp, k, np.zeros((n, k)))
    raw = ctypes.POINTER(ctypes.c_double)()
    lda = ctypes.c_int32()
    be64.call("mvec_extract_view", v.h, ctypes.byref(raw), ctypes.byref(lda))
    assert lda.value == k
    for i in range(n):
        for j in range(k):
            raw[i * lda.value + j] = i + 0.5 * j
    be64.call("mvec_to_device", v.h)
    got = v.get()
    assert got[10, 2] == 11.0
    v.put_value(2.0)
    be64.call("mvec_from_device", v.h)
    assert raw[5] == 2.0
    v.delete()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_fast_sync_same_results(be64, be32, dtype):
    """host reductions and small downloads give the same values whether the device publishes them
    into pinned memory (phist_amd_set_fast_sync(1), the default) or they are copied after a stream
    synchronisation (0): column dots and norms of 1..40 columns (one or several reduction
    workgroups), the fused SpMV's dots, and sdMat downloads of whole and viewed blocks"""
    be = be64 if dtype == np.float64 else be32
    fl = ctypes.c_int(0)
    n = 50000
    m = be.map(n)
    A = be.laplace3d(50, 40, 25)
    rng = np.random.default_rng(21)
    res = {}
    try:
        for fast in (0, 1):
            be.L.phist_amd_set_fast_sync(fast, ctypes.byref(fl))
            assert fl.value == 0
            out = []
            for k in (1, 4, 8, 20, 40):
                X = rng.uniform(-1, 1, (n, k)).astype(dtype) if fast == 0 else res[0]["X"][k]
                Y = rng.uniform(-1, 1, (n, k)).astype(dtype) if fast == 0 else res[0]["Y"][k]
                x, y = be.mvec(m, k, X), be.mvec(m, k, Y)
                out.append(be.dot(x, y))
                out.append(be.norm2(x))
                w = be.mvec(m, k)
                out.extend(be.fused_spmv_mvdot(1.0, A, x, 0.0, w))
                C = be.sdmat(k, k)
                be.gram(1.0, x, y, 0.0, C)
                out.append(C.get())
                if k >= 4:
                    Cv = ctypes.c_void_p()
                    be.call("sdMat_view_block", C.h, ctypes.byref(Cv), 1, k - 2, 1, k - 1)
                    V = phist_amd.SdMat(be, k - 2, k - 1, Cv)
                    out.append(V.get())
                    V.delete()
                for o in (x, y, w, C):
                    o.delete()
                if fast == 0:
                    res.setdefault(0, {"X": {}, "Y": {}})
                    res[0]["X"][k], res[0]["Y"][k] = X, Y
            res.setdefault(fast, {})["out"] = out
    finally:
        be.L.phist_amd_set_fast_sync(1, ctypes.byref(fl))
    assert len(res[0]["out"]) == len(res[1]["out"])
    for a, b in zip(res[0]["out"], res[1]["out"]):
        assert np.array_equal(a, b)
    A.delete()
    m.delete()
