"""GPU: the SpMV autotuner (phist_[DS]amd_sparseMat_tune) through the C ABI.

Checked: the settings it reports are valid choices; the product after tuning
(new SELL layout and launch settings) equals the product before bit for bit
and the oracle to the FP64 bar; the host CRS is untouched and the rebuilt
SELL layout keeps its invariants; a second call with the cache returns the
same settings; the FP32 copy can be tuned with the FP64 matrix.  Timings are
reported, not asserted beyond a loose sanity bound (the untuned choice is
one of the candidates)."""
import ctypes

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _tune(be, A, nvecs, flags):
    n = len(nvecs)
    nv = np.array(nvecs, dtype=np.int32)
    dms, tms = np.zeros(n), np.zeros(n)
    st = np.zeros(7 * n, dtype=np.int32)
    fl = ctypes.c_int(0)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    getattr(be.L, f"phist_{be.p}amd_sparseMat_tune")(A.h, ctypes.c_int(n), p(nv), ctypes.c_int(flags), p(dms),
                                                      p(tms), p(st), ctypes.byref(fl))
    assert fl.value == 0
    return dms, tms, st.reshape(n, 7)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_tune_layout_and_launch(be64, be32, dtype):
    be = be64 if dtype == np.float64 else be32
    n = 150000
    A = be.banded_random(n, 13, 4000, 40.0, 321)
    ro, ci, va = A.crs()
    ref = ol.Crs(n, n, ro, ci, va.astype(float))
    m = be.map(n)
    rng = np.random.default_rng(4)
    before = {}
    for nvec in (1, 8):
        X = rng.uniform(-1, 1, (n, nvec)).astype(dtype)
        x, y = be.mvec(m, nvec, X), be.mvec(m, nvec)
        be.spmv(1.0, A, x, 0.0, y)
        before[nvec] = (X, y.get())
        x.delete()
        y.delete()
    dms, tms, st = _tune(be, A, [1, 8], 1 | 2)
    item = np.dtype(dtype).itemsize
    for (C, sigma, lb, bpc, depth, cblk, stg), d, t in zip(st, dms, tms):
        assert C in (16, 32, 64) and sigma in (1, 256)
        assert lb in (0, item, 32, 64) and bpc in (1, 2, 3, 4, 6, 8) and depth in (8, 16, 32)
        assert cblk == 0  # widths 1 and 8: one launch
        assert 0 < t <= 1.5 * d + 0.02, (t, d)
    s = A.sell()
    assert s["C"] == st[0][0]
    perm = s["perm"]
    assert sorted(perm[perm >= 0].tolist()) == list(range(n))
    ro2, ci2, va2 = A.crs()
    assert np.array_equal(ro, ro2) and np.array_equal(ci, ci2) and np.array_equal(va, va2)
    for nvec, (X, Y0) in before.items():
        x, y = be.mvec(m, nvec, X), be.mvec(m, nvec)
        be.spmv(1.0, A, x, 0.0, y)
        Y = y.get()
        assert np.array_equal(Y, Y0), f"nvec {nvec}: tuned product differs"
        want = ol.spmv(ref, 1.0, X.astype(float), 0.0, np.zeros((n, nvec)))
        num = np.linalg.norm(Y.astype(float) - want, axis=0)
        assert np.all(num <= (1e-12 if dtype == np.float64 else 1e-5) * np.linalg.norm(want, axis=0))
        x.delete()
        y.delete()
    # the cache answers a second call with the same choice
    _, _, st2 = _tune(be, A, [1, 8], 0)
    assert np.array_equal(st2, st)
    A.delete()
    m.delete()


def test_tune_fp32_twin(be64):
    """flags & 4: the FP32 copy used by mixed-precision solvers is tuned too;
    the FP64 matrix's products are unchanged"""
    n = 60000
    A = be64.banded_random(n, 9, 500, 40.0, 8)
    m = be64.map(n)
    X = np.random.default_rng(6).uniform(-1, 1, (n, 4))
    x, y = be64.mvec(m, 4, X), be64.mvec(m, 4)
    be64.spmv(1.0, A, x, 0.0, y)
    Y0 = y.get()
    _tune(be64, A, [4], 4 | 2)
    be64.spmv(1.0, A, x, 0.0, y)
    assert np.array_equal(y.get(), Y0)
    for o in (x, y, A):
        o.delete()
    m.delete()


@pytest.mark.parametrize("lane_bytes", [0, 32])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nvec", [1, 4, 8, 16])
def test_spmv_depth_bitexact(be64, be32, dtype, nvec, lane_bytes):
    """depths 16 and 32 (the next batch's indices prefetched; a whole 32-entry
    row's gathers in flight) and depth 8 give the same product bit for bit, on rows of 32 entries, of 40 entries (the
    second batch of a row mostly empty) and of at most 7 (a 3D Laplacian:
    every batch mostly empty), and agree with the oracle"""
    be = be64 if dtype == np.float64 else be32
    fl = ctypes.c_int(0)
    for n, nnz, hb in ((50000, 32, 3000), (30000, 40, 500), (8000, 0, 0)):
        A = be.banded_random(n, nnz, hb, 40.0, 17 + nnz) if nnz else be.laplace3d(20, 20, 20)
        ro, ci, va = A.crs()
        ref = ol.Crs(n, n, ro, ci, va.astype(float))
        m = be.map(n)
        X = np.random.default_rng(nvec).uniform(-1, 1, (n, nvec)).astype(dtype)
        x = be.mvec(m, nvec, X)
        out = {}
        try:
            be.L.phist_amd_set_spmv_lane_bytes(lane_bytes, ctypes.byref(fl))
            assert fl.value == 0
            for depth in (8, 16, 32):
                be.L.phist_amd_set_spmv_depth(depth, ctypes.byref(fl))
                assert fl.value == 0
                y = be.mvec(m, nvec)
                be.spmv(1.0, A, x, 0.0, y)
                out[depth] = y.get()
                y.delete()
        finally:
            be.L.phist_amd_set_spmv_depth(0, ctypes.byref(fl))
            be.L.phist_amd_set_spmv_lane_bytes(0, ctypes.byref(fl))
        assert np.array_equal(out[8], out[16]) and np.array_equal(out[8], out[32])
        want = ol.spmv(ref, 1.0, X.astype(float), 0.0, np.zeros((n, nvec)))
        num = np.linalg.norm(out[32].astype(float) - want, axis=0)
        assert np.all(num <= (1e-12 if dtype == np.float64 else 1e-5) * np.linalg.norm(want, axis=0))
        x.delete()
        A.delete()
        m.delete()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nvec", [32, 48])
def test_spmv_column_blocks_bitexact(be64, be32, dtype, nvec):
    """launches of 16 or 32 block columns (phist_amd_set_spmv_col_block) give the one-launch SpMV
    result bit for bit, shifts and beta included; fused column dots (to the host and, through the
    device MINRES, to the device) agree to rounding; the tuner may choose a column block at 48
    columns and its choice keeps the product bit for bit"""
    be = be64 if dtype == np.float64 else be32
    n = 30000
    A = be.banded_random(n, 20, 2000, 40.0, 55)
    ro, ci, va = A.crs()
    ref = ol.Crs(n, n, ro, ci, va.astype(float))
    rng = np.random.default_rng(nvec)
    X = rng.uniform(-1, 1, (n, nvec)).astype(dtype)
    Y0 = rng.uniform(-1, 1, (n, nvec)).astype(dtype)
    shifts = rng.uniform(-2, 2, nvec).astype(dtype)
    m = be.map(n)
    x = be.mvec(m, nvec, X)
    fl = ctypes.c_int(0)
    out = {}
    try:
        for cb in (0, 16, 32):
            be.L.phist_amd_set_spmv_col_block(cb, ctypes.byref(fl))
            assert fl.value == 0
            y = be.mvec(m, nvec, Y0)
            be.spmv_shifted(0.7, A, shifts, x, 0.3, y)
            w = be.mvec(m, nvec, Y0)
            ww, vw = be.fused_spmv_mvdot(1.1, A, x, 0.0, w)
            out[cb] = (y.get(), w.get(), ww, vw)
            y.delete()
            w.delete()
    finally:
        be.L.phist_amd_set_spmv_col_block(-1, ctypes.byref(fl))
    want = ol.spmv(ref, 0.7, X.astype(float), 0.3, Y0.astype(float), shifts=shifts.astype(float))
    tol = 1e-12 if dtype == np.float64 else 1e-5
    num = np.linalg.norm(out[0][0].astype(float) - want, axis=0)
    assert np.all(num <= tol * np.linalg.norm(want, axis=0))
    for cb in (16, 32):
        assert np.array_equal(out[cb][0], out[0][0]) and np.array_equal(out[cb][1], out[0][1]), cb
        for g, d in zip(out[cb][2:], out[0][2:]):
            assert np.max(np.abs(g - d) / np.abs(d)) <= (1e-13 if dtype == np.float64 else 1e-5)
    # the device dots of the MINRES iteration with column blocks: the same solutions to the tolerance
    if dtype == np.float64:  # (MINRES needs a symmetric matrix: a 3D Laplacian)
        Al = be.laplace3d(30, 30, 30)
        ml = be.map(27000)
        B = be.mvec(ml, nvec, rng.uniform(-1, 1, (27000, nvec)))
        sol = {}
        try:
            for cb in (0, 16):
                be.L.phist_amd_set_spmv_col_block(cb, ctypes.byref(fl))
                Xs = be.mvec(ml, nvec)
                it, rr, unconv = be.minres_shifted(Al, np.full(nvec, 0.5), B, Xs, 1e-10, 500)
                assert unconv == 0
                sol[cb] = Xs.get()
                Xs.delete()
        finally:
            be.L.phist_amd_set_spmv_col_block(-1, ctypes.byref(fl))
        assert np.max(np.abs(sol[16] - sol[0])) <= 1e-8 * np.max(np.abs(sol[0]))
        for o in (B, Al, ml):
            o.delete()
    # the tuner's choice at this width
    before = out[0][1]
    _tune(be, A, [nvec], 2)
    w = be.mvec(m, nvec, Y0)
    be.fused_spmv_mvdot(1.1, A, x, 0.0, w)
    assert np.array_equal(w.get(), before)
    for o in (x, w, A):
        o.delete()
    m.delete()



def _staged(be):
    c, fl = ctypes.c_int64(0), ctypes.c_int(0)
    be.L.phist_amd_spmv_staged_launches(ctypes.byref(c), ctypes.byref(fl))
    assert fl.value == 0
    return c.value


def _stage_expected(nvec, item, lane_bytes):
    """whether a one-launch SpMV of nvec contiguous columns can stage x (the library's rule:
    lanes of whole 16-byte vectors, at most 8 lanes per row), the LDS window fitting"""
    v16 = 16 // item
    vpl = v16
    if lane_bytes == 32 and nvec % (32 // item) == 0:
        vpl = 32 // item
    if nvec % v16:
        return False
    groups, lpr = -(-nvec // vpl), 1
    while lpr < groups:
        lpr *= 2
    return lpr <= 8


@pytest.mark.parametrize("depth", [8, 16])
@pytest.mark.parametrize("lane_bytes", [0, 32])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("nvec", [2, 4, 8, 16, 32])
def test_spmv_stage_bitexact(be64, be32, dtype, nvec, lane_bytes, depth):
    """x staged in LDS per workgroup tile (phist_amd_set_spmv_stage) gives the unstaged product bit for
    bit, shifts and beta included, and the fused column dots to rounding, on a 3D Laplacian (tiles
    gather from three short intervals: staged wherever the block rows are whole 16-byte vectors in
    at most 8 lanes), a narrow band (one interval per tile) and a honeycomb sheet; a wide band whose
    windows do not fit falls back to the unstaged kernel; all agree with the oracle"""
    be = be64 if dtype == np.float64 else be32
    item = np.dtype(dtype).itemsize
    fl = ctypes.c_int(0)
    cases = [("laplace3d", lambda: be.laplace3d(20, 21, 22), True),
             ("band300", lambda: be.banded_random(30000, 20, 300, 40.0, 5), None),
             ("honeycomb", lambda: be.honeycomb(60, 80, 0, 0, 0, 1.0, 7), None),
             ("band20k", lambda: be.banded_random(60000, 24, 20000, 40.0, 6), False)]
    for name, make, expect in cases:
        A = make()
        n = A.nrows
        ro, ci, va = A.crs()
        ref = ol.Crs(n, n, ro, ci, va.astype(float))
        rng = np.random.default_rng(nvec + 100 * depth)
        X = rng.uniform(-1, 1, (n, nvec)).astype(dtype)
        Y0 = rng.uniform(-1, 1, (n, nvec)).astype(dtype)
        shifts = rng.uniform(-2, 2, nvec).astype(dtype)
        m = be.map(n)
        x = be.mvec(m, nvec, X)
        out = {}
        try:
            be.L.phist_amd_set_spmv_lane_bytes(lane_bytes, ctypes.byref(fl))
            be.L.phist_amd_set_spmv_depth(depth, ctypes.byref(fl))
            for stage in (0, 1):
                be.L.phist_amd_set_spmv_stage(stage, ctypes.byref(fl))
                assert fl.value == 0
                c0 = _staged(be)
                y = be.mvec(m, nvec, Y0)
                be.spmv_shifted(0.7, A, shifts, x, 0.3, y)
                w = be.mvec(m, nvec, Y0)
                ww, vw = be.fused_spmv_mvdot(1.1, A, x, 0.0, w)
                out[stage] = (y.get(), w.get(), ww, vw, _staged(be) - c0)
                y.delete()
                w.delete()
        finally:
            be.L.phist_amd_set_spmv_stage(-1, ctypes.byref(fl))
            be.L.phist_amd_set_spmv_depth(0, ctypes.byref(fl))
            be.L.phist_amd_set_spmv_lane_bytes(0, ctypes.byref(fl))
        assert out[0][4] == 0
        if expect is True:
            assert (out[1][4] == 2) == _stage_expected(nvec, item, lane_bytes), (name, out[1][4])
        elif expect is False:
            assert out[1][4] == 0, name
        assert np.array_equal(out[1][0], out[0][0]), f"{name}: staged shifted product differs"
        assert np.array_equal(out[1][1], out[0][1]), f"{name}: staged fused product differs"
        for g, d in zip(out[1][2:4], out[0][2:4]):
            assert np.max(np.abs(g - d) / np.abs(d)) <= (1e-13 if dtype == np.float64 else 1e-5)
        want = ol.spmv(ref, 0.7, X.astype(float), 0.3, Y0.astype(float), shifts=shifts.astype(float))
        num = np.linalg.norm(out[1][0].astype(float) - want, axis=0)
        assert np.all(num <= (1e-12 if dtype == np.float64 else 1e-5) * np.linalg.norm(want, axis=0))
        x.delete()
        A.delete()
        m.delete()


def test_spmv_stage_tuned_laplacian(be64):
    """the tuner weighs staging on a 3D Laplacian at 4 and 8 columns (reported as the 7th setting); the
    product under its choice equals the unstaged one bit for bit"""
    n = 32 ** 3
    A = be64.laplace3d(32, 32, 32)
    m = be64.map(n)
    rng = np.random.default_rng(11)
    res = {}
    for nvec in (4, 8):
        X = rng.uniform(-1, 1, (n, nvec))
        x, y = be64.mvec(m, nvec, X), be64.mvec(m, nvec)
        fl = ctypes.c_int(0)
        be64.L.phist_amd_set_spmv_stage(0, ctypes.byref(fl))
        try:
            be64.spmv(1.0, A, x, 0.0, y)
        finally:
            be64.L.phist_amd_set_spmv_stage(-1, ctypes.byref(fl))
        res[nvec] = (X, y.get())
        x.delete()
        y.delete()
    _, _, st = _tune(be64, A, [4, 8], 2)
    assert set(st[:, 6].tolist()) <= {0, 1}
    for nvec, (X, Y0) in res.items():
        x, y = be64.mvec(m, nvec, X), be64.mvec(m, nvec)
        be64.spmv(1.0, A, x, 0.0, y)
        assert np.array_equal(y.get(), Y0), nvec
        x.delete()
        y.delete()
    A.delete()
    m.delete()
