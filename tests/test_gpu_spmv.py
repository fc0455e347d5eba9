"""GPU parity: SELL-C-sigma sparse matrix times block vector and its fused
variants, through the C ABI, against the oracle (the builtin restatement
pinned in test_oracle.py).

Bars: index work (CRS order, SELL layout) bit-exact; FP64 results within
1e-12 relative to the oracle (normwise per column); FP32 results within
1e-5 relative to the FP64 oracle.  The reference's own golden vectors
(A2_precalc_result) are checked to sqrt(eps) as in its test.
"""
import json
import os

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

MATS = ["Dsprandn25", "Dsprandn388", "Dsprandn_nodiag25", "Dsprandsym163", "DsymmMat20", "DjadaTestMat512",
        "Dspshift25", "Dspzero25", "Dspeye25"]
RTOL = {np.float64: 1e-12, np.float32: 1e-5}


def colwise_rel(got, ref):
    """max over columns of ||got-ref|| / max(||ref||, tiny)"""
    num = np.linalg.norm(got - ref, axis=0)
    den = np.maximum(np.linalg.norm(ref, axis=0), 1e-300)
    return float(np.max(num / den)) if got.size else 0.0


def mm(name):
    return os.path.join(ol.MATRICES, name + ".mm")


@pytest.mark.parametrize("mat", MATS)
def test_crs_index_bitexact(be64, mat):
    A = be64.read_mm(mm(mat))
    ref = ol.crs_from_mm(mm(mat))
    ro, ci, va = A.crs()
    assert np.array_equal(ro, ref.row_offset)
    assert np.array_equal(ci, ref.col_idx)
    assert np.array_equal(va, ref.val)
    A.delete()


@pytest.mark.parametrize("mat", MATS)
def test_sell_invariants(be64, mat):
    A = be64.read_mm(mm(mat))
    ref = ol.crs_from_mm(mm(mat))
    s = A.sell()
    C, nch = s["C"], s["nchunks"]
    perm = s["perm"]
    n = ref.nrows
    # bijection on the real rows, padding rows marked -1
    assert sorted(perm[perm >= 0].tolist()) == list(range(n))
    assert np.all(perm[n:] == -1)
    lens = np.diff(ref.row_offset)
    # descending length inside sigma windows (stable)
    sig = s["sigma"]
    for w0 in range(0, n, max(sig, 1)):
        win = perm[w0:min(n, w0 + max(sig, 1))]
        l = lens[win]
        assert np.all(l[:-1] >= l[1:]) or sig == 1
    # chunk widths and slots reproduce every row in CRS order, padding is zero
    for c in range(nch):
        w = s["chunk_w"][c]
        assert s["chunk_ptr"][c + 1] - s["chunk_ptr"][c] == w * C
        for r in range(C):
            row = perm[c * C + r]
            slots = s["chunk_ptr"][c] + np.arange(w) * C + r
            if row < 0:
                assert np.all(s["val"][slots] == 0)
                continue
            a, b = ref.row_offset[row] - 1, ref.row_offset[row + 1] - 1
            L = b - a
            assert w >= L
            assert np.array_equal(s["col"][slots[:L]], ref.col_idx[a:b] - 1)
            assert np.array_equal(s["val"][slots[:L]], ref.val[a:b])
            assert np.all(s["val"][slots[L:]] == 0)
    A.delete()


@pytest.mark.parametrize("nvec", [1, 4])
def test_reference_golden_vectors(be64, nvec):
    with open(os.path.join(ol.GOLDEN, "a2_precalc_sprandn25.json")) as f:
        g = json.load(f)
    A = be64.read_mm(mm("Dsprandn25"))
    m = be64.map(25)
    x = be64.mvec(m, nvec, np.array([[i + j * 25 for j in range(nvec)] for i in range(25)], float))
    y = be64.mvec(m, nvec)
    be64.spmv(1.0, A, x, 0.0, y)
    ref = np.array(g[f"nvec{nvec}"]).reshape(25, nvec)
    assert np.max(np.abs(y.get() - ref)) <= np.sqrt(np.finfo(float).eps)
    for o in (x, y, A):
        o.delete()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("mat", MATS)
@pytest.mark.parametrize("nvec", [1, 2, 3, 4, 7, 8, 12, 16, 32])
@pytest.mark.parametrize("alpha,beta,shifted", [(1.0, 0.0, False), (0.7, -0.3, False), (-1.3, 0.5, True)])
def test_spmv_vs_oracle(be64, be32, dtype, mat, nvec, alpha, beta, shifted):
    be = be64 if dtype == np.float64 else be32
    A = be.read_mm(mm(mat))
    ref = ol.crs_from_mm(mm(mat))
    rng = np.random.default_rng(nvec * 31 + len(mat))
    X = rng.uniform(-1, 1, (ref.ncols, nvec)).astype(dtype)
    Y0 = rng.uniform(-1, 1, (ref.nrows, nvec)).astype(dtype)
    sh = rng.uniform(-1, 1, nvec).astype(dtype)
    m = be.map(ref.nrows)
    x = be.mvec(m, nvec, X)
    y = be.mvec(m, nvec, Y0)
    if shifted:
        be.spmv_shifted(alpha, A, sh, x, beta, y)
    else:
        be.spmv(alpha, A, x, beta, y)
    want = ol.spmv(ref, alpha, X.astype(float), beta, Y0.astype(float), shifts=sh.astype(float) if shifted else None)
    assert colwise_rel(y.get().astype(float), want) <= RTOL[dtype]
    for o in (x, y, A):
        o.delete()
    m.delete()


@pytest.mark.parametrize("nvec,jmin,jmax", [(8, 2, 5), (12, 4, 11), (5, 1, 1), (16, 0, 7)])
def test_spmv_on_column_views(be64, nvec, jmin, jmax):
    """strided views (lda > nvec): x and y are column blocks of wider mvecs"""
    A = be64.read_mm(mm("Dsprandn388"))
    ref = ol.crs_from_mm(mm("Dsprandn388"))
    rng = np.random.default_rng(1)
    X = rng.uniform(-1, 1, (388, nvec))
    Y = rng.uniform(-1, 1, (388, nvec))
    m = be64.map(388)
    x = be64.mvec(m, nvec, X)
    y = be64.mvec(m, nvec, Y)
    xv = x.view(jmin, jmax)
    yv = y.view(jmin, jmax)
    be64.spmv(0.9, A, xv, 0.4, yv)
    want = Y.copy()
    want[:, jmin:jmax + 1] = ol.spmv(ref, 0.9, X[:, jmin:jmax + 1], 0.4, Y[:, jmin:jmax + 1])
    got = y.get()
    assert colwise_rel(got, want) <= 1e-12
    # columns outside the view are untouched (bit-exact)
    mask = np.ones(nvec, bool)
    mask[jmin:jmax + 1] = False
    assert np.array_equal(got[:, mask], Y[:, mask])
    for o in (xv, yv, x, y, A):
        o.delete()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nvec", [1, 4, 8, 16])
@pytest.mark.parametrize("beta", [0.0, 0.25])
@pytest.mark.parametrize("which", ["both", "ww", "vw", "none"])
def test_fused_spmv_mvdot(be64, be32, dtype, nvec, beta, which):
    be = be64 if dtype == np.float64 else be32
    A = be.read_mm(mm("Dsprandsym163"))
    ref = ol.crs_from_mm(mm("Dsprandsym163"))
    rng = np.random.default_rng(nvec)
    V = rng.uniform(-1, 1, (163, nvec)).astype(dtype)
    W0 = rng.uniform(-1, 1, (163, nvec)).astype(dtype)
    m = be.map(163)
    v = be.mvec(m, nvec, V)
    w = be.mvec(m, nvec, W0)
    ww, vw = be.fused_spmv_mvdot(1.1, A, v, beta, w, want_ww=which in ("both", "ww"), want_vw=which in ("both", "vw"))
    Wref = ol.spmv(ref, 1.1, V.astype(float), beta, W0.astype(float))
    tol = RTOL[dtype]
    assert colwise_rel(w.get().astype(float), Wref) <= tol
    if ww is not None:
        assert np.allclose(ww, ol.dot(Wref, Wref), rtol=10 * tol, atol=0)
    if vw is not None:
        scale = ol.dot(np.abs(V.astype(float)), np.abs(Wref))
        assert np.all(np.abs(vw - ol.dot(V.astype(float), Wref)) <= 10 * tol * scale)
    for o in (v, w, A):
        o.delete()


def test_fused_spmv_mvdot_mvadd(be64):
    A = be64.read_mm(mm("Dsprandn388"))
    ref = ol.crs_from_mm(mm("Dsprandn388"))
    rng = np.random.default_rng(5)
    V, W0, U0 = (rng.uniform(-1, 1, (388, 4)) for _ in range(3))
    m = be64.map(388)
    v, w, u = be64.mvec(m, 4, V), be64.mvec(m, 4, W0), be64.mvec(m, 4, U0)
    import ctypes
    ww = np.zeros(4)
    vw = np.zeros(4)
    be64.call("fused_spmv_mvdot_mvadd", be64.s(0.5), A.h, v.h, be64.s(2.0), w.h, be64.s(-1.5), be64.s(0.75), u.h,
              ww.ctypes.data_as(ctypes.c_void_p), vw.ctypes.data_as(ctypes.c_void_p))
    Wref = ol.spmv(ref, 0.5, V, 2.0, W0)
    Uref = -1.5 * Wref + 0.75 * U0
    assert colwise_rel(w.get(), Wref) <= 1e-12
    assert colwise_rel(u.get(), Uref) <= 1e-12
    assert np.allclose(ww, ol.dot(Wref, Wref), rtol=1e-12)
    assert np.allclose(vw, ol.dot(V, Wref), rtol=1e-11, atol=1e-12)
    for o in (v, w, u, A):
        o.delete()


def test_spmv_alpha_zero_only_scales(be64):
    A = be64.read_mm(mm("Dsprandn25"))
    m = be64.map(25)
    Y = np.random.default_rng(0).uniform(-1, 1, (25, 3))
    x = be64.mvec(m, 3, np.full((25, 3), np.nan))
    y = be64.mvec(m, 3, Y)
    be64.spmv(0.0, A, x, 0.5, y)
    assert np.array_equal(y.get(), 0.5 * Y)
    for o in (x, y, A):
        o.delete()


def test_empty_rows_and_zero_matrix(be64):
    # Dspzero25 holds explicit zeros; a fully empty matrix only scales y
    import ctypes
    import phist_amd
    rows = ol.crs_from_rows(6, 6, [([], []), ([0, 5], [1.0, 2.0]), ([], []), ([3], [4.0]), ([], []), ([], [])])

    @phist_amd.ROWFUNC
    def rf(i, nnz, cols, vals, arg):
        a, b = rows.row_offset[i] - 1, rows.row_offset[i + 1] - 1
        nnz[0] = b - a
        dv = ctypes.cast(vals, ctypes.POINTER(ctypes.c_double))
        for k in range(a, b):
            cols[k - a] = rows.col_idx[k] - 1
            dv[k - a] = rows.val[k]
        return 0

    A = be64.from_rowfunc(6, 2, rf, None)
    m = be64.map(6)
    X = np.arange(12.0).reshape(6, 2)
    x = be64.mvec(m, 2, X)
    y = be64.mvec(m, 2, np.full((6, 2), 3.0))
    be64.spmv(2.0, A, x, 0.0, y)
    assert np.array_equal(y.get(), ol.spmv(rows, 2.0, X, 0.0, np.zeros((6, 2))))
    ro, ci, va = A.crs()
    assert np.array_equal(ro, rows.row_offset) and np.array_equal(ci, rows.col_idx)
    for o in (x, y, A):
        o.delete()


@pytest.mark.parametrize("C,sigma", [(1, 1), (4, 1), (8, 16), (32, 256), (64, 64)])
def test_sell_parameters(be64, C, sigma):
    """any power-of-two C and any sigma give the same product"""
    import ctypes
    iflag = ctypes.c_int(0)
    be64.L.phist_amd_sparseMat_set_format(C, sigma, ctypes.byref(iflag))
    try:
        A = be64.read_mm(mm("DjadaTestMat512"))
    finally:
        be64.L.phist_amd_sparseMat_set_format(0, 0, ctypes.byref(iflag))
    ref = ol.crs_from_mm(mm("DjadaTestMat512"))
    s = A.sell()
    assert s["C"] == C
    X = np.random.default_rng(C).uniform(-1, 1, (512, 8))
    m = be64.map(512)
    x = be64.mvec(m, 8, X)
    y = be64.mvec(m, 8)
    be64.spmv(1.0, A, x, 0.0, y)
    assert colwise_rel(y.get(), ol.spmv(ref, 1.0, X, 0.0, np.zeros((512, 8)))) <= 1e-12
    for o in (x, y, A):
        o.delete()


@pytest.mark.parametrize("n,nnz,hb", [(20000, 32, 500), (4096, 9, 100)])
def test_banded_generator_rows_and_product(be64, n, nnz, hb):
    """the bench's generator through create_fromRowFuncWithConstructor:
    distinct columns inside the band, diagonal 40, product vs the oracle"""
    A = be64.banded_random(n, nnz, hb, 40.0, 1234)
    ro, ci, va = A.crs()
    ref = ol.Crs(n, n, ro, ci, va)
    lens = np.diff(ro)
    assert np.all(lens == nnz)
    for i in range(0, n, 997):
        cols = ci[ro[i] - 1:ro[i + 1] - 1] - 1
        assert len(set(cols)) == len(cols) and np.all(np.diff(cols) > 0)
        assert np.all(np.abs(cols - i) <= hb)
        d = va[ro[i] - 1:ro[i + 1] - 1][cols == i]
        assert d.tolist() == [40.0]
    X = np.random.default_rng(2).uniform(-1, 1, (n, 8))
    m = be64.map(n)
    x = be64.mvec(m, 8, X)
    y = be64.mvec(m, 8)
    ww, vw = be64.fused_spmv_mvdot(1.0, A, x, 0.0, y)
    Yref = ol.spmv(ref, 1.0, X, 0.0, np.zeros((n, 8)))
    assert colwise_rel(y.get(), Yref) <= 1e-12
    assert np.allclose(ww, ol.dot(Yref, Yref), rtol=1e-12)
    for o in (x, y, A):
        o.delete()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nvec", [12, 16, 24])
def test_spmv_column_blocks_large(be64, be32, dtype, nvec):
    """300000 rows of 9 nonzeros and 12-24 columns (a persistent grid of
    many tiles per workgroup): shifts, beta and both dot products"""
    be = be64 if dtype == np.float64 else be32
    n = 300000
    A = be.banded_random(n, 9, 100, 40.0, 77)
    ro, ci, va = A.crs()
    ref = ol.Crs(n, n, ro, ci, va.astype(float))
    rng = np.random.default_rng(nvec)
    X = rng.uniform(-1, 1, (n, nvec)).astype(dtype)
    Y0 = rng.uniform(-1, 1, (n, nvec)).astype(dtype)
    shifts = rng.uniform(-2, 2, nvec).astype(dtype)
    m = be.map(n)
    x, y = be.mvec(m, nvec, X), be.mvec(m, nvec, Y0)
    be.spmv_shifted(0.7, A, shifts, x, 0.3, y)
    Xf = X.astype(float)
    want = ol.spmv(ref, 0.7, Xf, 0.3, Y0.astype(float), shifts=shifts.astype(float))
    tol = RTOL[dtype]
    assert colwise_rel(y.get().astype(float), want) <= tol
    w = be.mvec(m, nvec, Y0)
    ww, vw = be.fused_spmv_mvdot(1.1, A, x, 0.5, w)
    Wref = ol.spmv(ref, 1.1, Xf, 0.5, Y0.astype(float))
    assert colwise_rel(w.get().astype(float), Wref) <= tol
    assert np.allclose(ww, ol.dot(Wref, Wref), rtol=10 * tol, atol=0)
    scale = ol.dot(np.abs(Xf), np.abs(Wref))
    assert np.all(np.abs(vw - ol.dot(Xf, Wref)) <= 10 * tol * scale)
    for o in (x, y, w, A):
        o.delete()


def _p(a):
    import ctypes
    return a.ctypes.data_as(ctypes.c_void_p)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nvec", [1, 4, 8])
def test_spmv_add_mvec_scalar_shift(be64, be32, dtype, nvec):
    """sparseMat_times_mvec_add_mvec: y = alpha*(A + shift I)x + beta*y with one scalar shift
    (kernels_common_impl_def.h:55-64)"""
    be = be64 if dtype == np.float64 else be32
    A = be.read_mm(mm("Dsprandn388"))
    ref = ol.crs_from_mm(mm("Dsprandn388"))
    rng = np.random.default_rng(nvec + 40)
    X = rng.uniform(-1, 1, (388, nvec)).astype(dtype)
    Y0 = rng.uniform(-1, 1, (388, nvec)).astype(dtype)
    m = be.map(388)
    x, y = be.mvec(m, nvec, X), be.mvec(m, nvec, Y0)
    be.call("sparseMat_times_mvec_add_mvec", be.s(0.6), A.h, be.s(-1.7), x.h, be.s(0.4), y.h)
    want = ol.spmv(ref, 0.6, X.astype(float), 0.4, Y0.astype(float), shifts=np.full(nvec, -1.7))
    assert colwise_rel(y.get().astype(float), want) <= RTOL[dtype]
    for o in (x, y, A):
        o.delete()


def _gram_bound(tol, a, b):
    return 8 * tol * (np.abs(a).T @ np.abs(b))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nvec", [1, 3, 4, 8, 16, 17, 24, 32, 40])
@pytest.mark.parametrize("case", ["W", "W_beta", "W_null", "ww_only", "vw_only", "W_no_gram"])
def test_fused_spmv_mvTmv(be64, be32, dtype, nvec, case):
    """fused_spmv_mvTmv (phist_kernels_fused_decl.h:68, kernels_no_fused_spmv_mvTmv.cpp:9-57):
    W = alpha*A*V + beta*W, WtW = W'W, VtW = V'W; W = NULL means alpha = 1, beta = 0 and W is not
    returned; WtW / VtW may be NULL.  Up to 32 columns this is one pass (Gram epilogue on the matrix
    cores), 40 columns take the three-pass path."""
    be = be64 if dtype == np.float64 else be32
    A = be.read_mm(mm("DjadaTestMat512"))
    ref = ol.crs_from_mm(mm("DjadaTestMat512"))
    n = ref.nrows
    rng = np.random.default_rng(nvec * 7 + len(case))
    V = rng.uniform(-1, 1, (n, nvec)).astype(dtype)
    W0 = rng.uniform(-1, 1, (n, nvec)).astype(dtype)
    alpha, beta = 0.9, (0.35 if case == "W_beta" else 0.0)
    m = be.map(n)
    v = be.mvec(m, nvec, V)
    w = None if case == "W_null" else be.mvec(m, nvec, W0)
    C0 = rng.uniform(-1, 1, (nvec, nvec))
    wtw = None if case in ("vw_only", "W_no_gram") else be.sdmat(nvec, nvec, C0)
    vtw = None if case in ("ww_only", "W_no_gram") else be.sdmat(nvec, nvec, C0)
    be.call("fused_spmv_mvTmv", be.s(alpha), A.h, v.h, be.s(beta), w.h if w else None, wtw.h if wtw else None,
            vtw.h if vtw else None)
    Vf = V.astype(float)
    a_eff, b_eff = (1.0, 0.0) if w is None else (alpha, beta)
    Wref = ol.spmv(ref, a_eff, Vf, b_eff, W0.astype(float))
    tol = RTOL[dtype]
    if w is not None:
        Wgot = w.get().astype(float)
        assert colwise_rel(Wgot, Wref) <= tol
        if case == "W_no_gram":
            assert np.array_equal(Wgot, Wref) or colwise_rel(Wgot, Wref) <= tol
    if wtw is not None:
        G = wtw.get().astype(float)
        assert np.all(np.abs(G - ol.gram(1.0, Wref, Wref, 0.0, np.zeros((nvec, nvec)))) <=
                      _gram_bound(tol, Wref, Wref))
        assert np.array_equal(G, G.T)  # the mirrored tiles are exact copies
    if vtw is not None:
        G = vtw.get().astype(float)
        assert np.all(np.abs(G - ol.gram(1.0, Vf, Wref, 0.0, np.zeros((nvec, nvec)))) <= _gram_bound(tol, Vf, Wref))
    for o in (v, w, wtw, vtw, A):
        if o is not None:
            o.delete()


@pytest.mark.parametrize("nvec", [8, 32])
def test_fused_spmv_mvTmv_large(be64, nvec):
    """one-pass Gram epilogue over a persistent grid (300000 rows): both Grams against the oracle"""
    n = 300000
    A = be64.banded_random(n, 9, 100, 40.0, 31)
    ro, ci, va = A.crs()
    ref = ol.Crs(n, n, ro, ci, va)
    rng = np.random.default_rng(nvec)
    V = rng.uniform(-1, 1, (n, nvec))
    m = be64.map(n)
    v, w = be64.mvec(m, nvec, V), be64.mvec(m, nvec)
    wtw, vtw = be64.sdmat(nvec, nvec), be64.sdmat(nvec, nvec)
    be64.call("fused_spmv_mvTmv", be64.s(1.3), A.h, v.h, be64.s(0.0), w.h, wtw.h, vtw.h)
    Wref = ol.spmv(ref, 1.3, V, 0.0, np.zeros((n, nvec)))
    assert colwise_rel(w.get(), Wref) <= 1e-12
    assert np.all(np.abs(wtw.get() - ol.gram(1.0, Wref, Wref, 0.0, np.zeros((nvec, nvec)))) <=
                  _gram_bound(1e-12, Wref, Wref))
    assert np.all(np.abs(vtw.get() - ol.gram(1.0, V, Wref, 0.0, np.zeros((nvec, nvec)))) <=
                  _gram_bound(1e-12, V, Wref))
    for o in (v, w, wtw, vtw, A):
        o.delete()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nvec", [1, 4])
def test_fused_spmv_pair(be64, be32, dtype, nvec):
    """fused_spmv_pair: Y = alpha*(shifts1[j]*A1 + shifts2[j]*A2) X_j + beta*Y
    (phist_kernels_fused_decl.h:45, kernels_no_fused_spmv_pair.cpp:9-49)"""
    be = be64 if dtype == np.float64 else be32
    A1, A2 = be.read_mm(mm("Dsprandn25")), be.read_mm(mm("Dspshift25"))
    r1, r2 = ol.crs_from_mm(mm("Dsprandn25")), ol.crs_from_mm(mm("Dspshift25"))
    rng = np.random.default_rng(nvec + 3)
    X = rng.uniform(-1, 1, (25, nvec)).astype(dtype)
    Y0 = rng.uniform(-1, 1, (25, nvec)).astype(dtype)
    s1 = rng.uniform(-2, 2, nvec).astype(dtype)
    s2 = rng.uniform(-2, 2, nvec).astype(dtype)
    m = be.map(25)
    x, y = be.mvec(m, nvec, X), be.mvec(m, nvec, Y0)
    be.call("fused_spmv_pair", be.s(0.7), _p(s1), A1.h, _p(s2), A2.h, x.h, be.s(-0.5), y.h)
    Xf = X.astype(float)
    t1 = ol.spmv(r1, 0.7, Xf, 0.0, np.zeros((25, nvec)))
    t2 = ol.spmv(r2, 0.7, Xf, 0.0, np.zeros((25, nvec)))
    want = s1.astype(float) * t1 + s2.astype(float) * t2 - 0.5 * Y0.astype(float)
    assert colwise_rel(y.get().astype(float), want) <= RTOL[dtype]
    for o in (x, y, A1, A2):
        o.delete()


def test_mvec_QR_not_implemented(be64):
    """builtin's mvec_QR returns -99 so that callers fall back to CholQR (kernels_def.hpp:939-947)"""
    m = be64.map(100)
    v = be64.mvec(m, 4, np.random.default_rng(0).uniform(-1, 1, (100, 4)))
    r = be64.sdmat(4, 4)
    assert be64.call("mvec_QR", v.h, r.h, allow=(0, -99)) == -99
    for o in (v, r):
        o.delete()
    m.delete()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nvec", [8, 16, 32])
def test_spmv_lane_widths_bitexact(be64, be32, dtype, nvec):
    """every gather width per lane (one scalar, one to four 16-byte vectors)
    gives the default's SpMV result bit for bit, shifts and beta included (a
    row's sum is formed by the lane owning its columns, in entry order), and
    the default agrees with the oracle.  The column dots are reduced over the
    rows of a wave, whose number changes with the lane width, so they agree
    to rounding only."""
    import ctypes
    be = be64 if dtype == np.float64 else be32
    n = 40000
    A = be.banded_random(n, 17, 3000, 40.0, 99)
    ro, ci, va = A.crs()
    ref = ol.Crs(n, n, ro, ci, va.astype(float))
    rng = np.random.default_rng(nvec + 5)
    X = rng.uniform(-1, 1, (n, nvec)).astype(dtype)
    Y0 = rng.uniform(-1, 1, (n, nvec)).astype(dtype)
    shifts = rng.uniform(-2, 2, nvec).astype(dtype)
    m = be.map(n)
    x = be.mvec(m, nvec, X)
    fl = ctypes.c_int(0)
    out = {}
    try:
        for lb in (0, np.dtype(dtype).itemsize, 16, 32, 64):
            be.L.phist_amd_set_spmv_lane_bytes(lb, ctypes.byref(fl))
            assert fl.value == 0
            y = be.mvec(m, nvec, Y0)
            be.spmv_shifted(0.7, A, shifts, x, 0.3, y)
            w = be.mvec(m, nvec, Y0)
            ww, vw = be.fused_spmv_mvdot(1.1, A, x, 0.5, w)
            out[lb] = (y.get(), w.get(), ww, vw)
            y.delete()
            w.delete()
    finally:
        be.L.phist_amd_set_spmv_lane_bytes(0, ctypes.byref(fl))
    want = ol.spmv(ref, 0.7, X.astype(float), 0.3, Y0.astype(float), shifts=shifts.astype(float))
    assert colwise_rel(out[0][0].astype(float), want) <= RTOL[dtype]
    dtol = 1e-13 if dtype == np.float64 else 1e-5
    for lb, got in out.items():
        for g, d in zip(got[:2], out[0][:2]):
            assert np.array_equal(g, d), f"lane width {lb} bytes: SpMV result differs from the default"
        for g, d in zip(got[2:], out[0][2:]):
            assert np.max(np.abs(g - d) / np.abs(d)) <= dtol, f"lane width {lb} bytes: dots differ"
    for o in (x, A):
        o.delete()
    m.delete()
