"""GPU: matrices built from the spin-chain and honeycomb generators through
create_fromRowFuncWithConstructor.  The stored CRS must equal the oracle's rows
after builtin's stable column sort (bit for bit), and products with the
device SELL copy must match the oracle SpMV."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def crs_of_rows(rows):
    """builtin CRS of unsorted generator rows: stable ascending column sort"""
    ro = [1]
    ci, va = [], []
    for c, v in rows:
        o = np.argsort(c, kind="stable")
        ci += (c[o] + 1).tolist()
        va += v[o].tolist()
        ro.append(ro[-1] + len(c))
    return np.array(ro, np.int64), np.array(ci, np.int32), np.array(va)


def check_matrix(be, A, rows):
    ro, ci, va = A.crs()
    wro, wci, wva = crs_of_rows(rows)
    assert np.array_equal(ro, wro) and np.array_equal(ci, wci)
    assert np.array_equal(va, wva.astype(be.dtype))
    n = len(rows)
    ref = ol.Crs(n, n, wro, wci, wva.astype(be.dtype).astype(float))
    X = np.random.default_rng(n).uniform(-1, 1, (n, 4))
    m = be.map(n)
    x, y = be.mvec(m, 4, X), be.mvec(m, 4)
    be.spmv(1.0, A, x, 0.0, y)
    want = ol.spmv(ref, 1.0, X.astype(be.dtype).astype(float), 0.0, np.zeros((n, 4)))
    tol = 1e-12 if be.dtype == np.float64 else 1e-5
    assert np.max(np.abs(y.get() - want)) <= tol * max(1.0, np.max(np.abs(want))) * 8
    for o in (x, y):
        o.delete()


@pytest.mark.parametrize("L,obc", [(8, 0), (12, 0), (14, 1)])
def test_spin_chain_matrix(be64, L, obc):
    A = be64.spinchain_sz(L, L // 2, obc)
    check_matrix(be64, A, ol.spin_sz_rows(L, L // 2, obc))
    A.delete()


def test_spin_chain_matrix_float(be32):
    A = be32.spinchain_sz(10, 5, 0)
    check_matrix(be32, A, ol.spin_sz_rows(10, 5, 0))
    A.delete()


@pytest.mark.parametrize("W,L,pbc,lr", [(8, 6, 0, 0), (10, 8, 1, 0), (12, 10, 1, 1), (9, 7, 0, 1)])
def test_honeycomb_matrix(be64, W, L, pbc, lr):
    A = be64.honeycomb(W, L, pbc, pbc, lr, 0.0, 0)
    check_matrix(be64, A, ol.honeycomb_rows(W, L, pbc, pbc, lr))
    A.delete()
