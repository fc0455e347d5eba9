"""GPU: phist_Dfused_spmv_mvdot_noalloc (the entry point that allocates
nothing; dots left in device arrays, partials in the library's fixed device
scratch) against phist_Dfused_spmv_mvdot on the same inputs, through the C
ABI: W and both dot arrays equal bit for bit (same kernels, same summation
order), and W agrees with the oracle.  Two calls in a row, the second on a
smaller matrix and block, so stale contents of the shared scratch would
show; plus beta != 0, NULL outputs, alpha = 0 and the rejected sizes."""
import ctypes

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

MVEC_RUN_ON_DEVICE = 2048


def colwise_rel(got, ref):
    num = np.linalg.norm(got - ref, axis=0)
    den = np.maximum(np.linalg.norm(ref, axis=0), 1e-300)
    return float(np.max(num / den))


class DevArray:
    """nvec doubles in device memory: the single row of a 1-row block"""

    def __init__(self, be, nvec):
        self.m = be.map(1)
        self.v = be.mvec(self.m, nvec, np.full((1, nvec), np.nan))
        raw = ctypes.POINTER(ctypes.c_double)()
        lda = ctypes.c_int32(0)
        fl = ctypes.c_int(MVEC_RUN_ON_DEVICE)
        be.L.phist_Dmvec_extract_view(self.v.h, ctypes.byref(raw), ctypes.byref(lda), ctypes.byref(fl))
        assert fl.value == 0
        self.ptr = ctypes.cast(raw, ctypes.c_void_p)

    def get(self):
        return self.v.get()[0]

    def delete(self):
        self.v.delete()
        self.m.delete()


def noalloc(be, alpha, A, V, beta, W, ww, vw):
    fl = ctypes.c_int(0)
    be.L.phist_Dfused_spmv_mvdot_noalloc(ctypes.c_double(alpha), A.h, V.h, ctypes.c_double(beta), W.h,
                                         ww.ptr if ww is not None else None, vw.ptr if vw is not None else None,
                                         ctypes.byref(fl))
    return fl.value


def one_case(be, n, nnz, hb, nvec, alpha, beta, seed, want=(True, True)):
    A = be.banded_random(n, nnz, hb, 40.0, seed)
    ro, ci, va = A.crs()
    ref = ol.Crs(n, n, ro, ci, va)
    rng = np.random.default_rng(seed)
    X = rng.uniform(-1, 1, (n, nvec))
    Y0 = rng.uniform(-1, 1, (n, nvec))
    m = be.map(n)
    x = be.mvec(m, nvec, X)
    w1 = be.mvec(m, nvec, Y0)
    w2 = be.mvec(m, nvec, Y0)
    ww1, vw1 = be.fused_spmv_mvdot(alpha, A, x, beta, w1, *want)
    dww = DevArray(be, nvec) if want[0] else None
    dvw = DevArray(be, nvec) if want[1] else None
    assert noalloc(be, alpha, A, x, beta, w2, dww, dvw) == 0
    W1, W2 = w1.get(), w2.get()
    assert np.array_equal(W1, W2)
    if want[0]:
        assert np.array_equal(dww.get(), ww1)
    if want[1]:
        assert np.array_equal(dvw.get(), vw1)
    Wref = ol.spmv(ref, alpha, X, beta, Y0)
    assert colwise_rel(W2, Wref) <= 1e-12
    if want[0]:
        assert np.max(np.abs(dww.get() - ol.dot(Wref, Wref)) / ol.dot(np.abs(Wref), np.abs(Wref))) <= 1e-12
    for o in (dww, dvw):
        if o is not None:
            o.delete()
    for o in (x, w1, w2, A):
        o.delete()
    m.delete()


def test_noalloc_matches_fused_spmv_mvdot_twice(be64):
    # a larger call, then a smaller one: the second must not see the first's partials
    one_case(be64, 60000, 24, 4000, 8, 1.0, 0.0, 11)
    one_case(be64, 9000, 9, 700, 4, 0.8, 0.0, 12)


def test_noalloc_beta_and_null_outputs(be64):
    one_case(be64, 30000, 17, 2000, 16, 1.3, 0.6, 13)
    one_case(be64, 30000, 17, 2000, 8, 1.0, 0.0, 14, want=(True, False))
    one_case(be64, 30000, 17, 2000, 3, 1.0, 0.2, 15, want=(False, True))
    one_case(be64, 20000, 5, 100, 1, 1.0, 0.0, 16, want=(False, False))


def test_noalloc_alpha_zero_scales_only(be64):
    one_case(be64, 5000, 7, 50, 8, 0.0, 0.5, 17)


def test_noalloc_rejects_oversized_blocks(be64):
    n = 1000
    A = be64.banded_random(n, 5, 10, 40.0, 3)
    m = be64.map(n)
    x = be64.mvec(m, 65)
    w = be64.mvec(m, 65)
    d = DevArray(be64, 65)
    assert noalloc(be64, 1.0, A, x, 0.0, w, d, None) == -55
    assert np.all(np.isnan(d.get()))  # nothing was launched
    for o in (d, x, w, A):
        o.delete()
    m.delete()


@pytest.mark.parametrize("nvec", [4, 8])
def test_noalloc_with_staging_forced(be64, nvec):
    """with x staging forced on (phist_amd_set_spmv_stage(1)) the entry point still allocates nothing:
    before the matrix's staging tables exist it runs unstaged (same W bit for bit, dots to rounding);
    once a regular call has built them it stages too and matches that call bit for bit"""
    fl = ctypes.c_int(0)
    A = be64.laplace3d(20, 21, 22)
    n = A.nrows
    rng = np.random.default_rng(nvec)
    X = rng.uniform(-1, 1, (n, nvec))
    m = be64.map(n)
    x = be64.mvec(m, nvec, X)
    staged = lambda: (be64.L.phist_amd_spmv_staged_launches(ctypes.byref(c), ctypes.byref(fl)), c.value)[1]
    c = ctypes.c_int64(0)
    try:
        be64.L.phist_amd_set_spmv_stage(1, ctypes.byref(fl))
        assert fl.value == 0
        wn1, wr, wn2 = be64.mvec(m, nvec), be64.mvec(m, nvec), be64.mvec(m, nvec)
        d1, v1 = DevArray(be64, nvec), DevArray(be64, nvec)
        s0 = staged()
        assert noalloc(be64, 1.0, A, x, 0.0, wn1, d1, v1) == 0
        assert staged() == s0  # no tables yet: not staged
        wwr, vwr = be64.fused_spmv_mvdot(1.0, A, x, 0.0, wr)
        assert staged() == s0 + 1  # the regular call built the tables and staged
        d2, v2 = DevArray(be64, nvec), DevArray(be64, nvec)
        assert noalloc(be64, 1.0, A, x, 0.0, wn2, d2, v2) == 0
        assert staged() == s0 + 2
        Wr = wr.get()
        assert np.array_equal(wn1.get(), Wr) and np.array_equal(wn2.get(), Wr)
        assert np.array_equal(d2.get(), wwr) and np.array_equal(v2.get(), vwr)
        assert np.max(np.abs(d1.get() - wwr) / np.abs(wwr)) <= 1e-13
        assert np.max(np.abs(v1.get() - vwr) / np.abs(vwr)) <= 1e-13
        for o in (wn1, wr, wn2, d1, v1, d2, v2):
            o.delete()
    finally:
        be64.L.phist_amd_set_spmv_stage(-1, ctypes.byref(fl))
    x.delete()
    A.delete()
    m.delete()
