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
