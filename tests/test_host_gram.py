"""CPU: the host Gram used as the CPU reference of the Gram and CholQR benches (oracle/host_gram.c, the
row loop of the reference's gemm_kernels_sC.F90): equal to numpy's V'W to rounding for self and mixed
Grams of several widths and ragged row counts, exact on integer-valued data, and bad sizes rejected."""
import ctypes

import numpy as np
import pytest

import oracle_lib as ol


@pytest.mark.parametrize("n,k,m", [(1000, 8, 8), (12345, 16, 16), (777, 32, 32), (5000, 4, 12), (3, 1, 1), (0, 2, 2)])
def test_host_gram_matches_numpy(n, k, m):
    rng = np.random.default_rng(n + k)
    V = rng.uniform(-1, 1, (n, k))
    W = rng.uniform(-1, 1, (n, m))
    G = ol.host_gram(V, W)
    if G is None:
        pytest.skip("oracle/libhost_gram.so is not built (make -C oracle)")
    assert G.shape == (k, m)
    assert np.allclose(G, V.T @ W, rtol=1e-12, atol=1e-12 * max(1, n))
    if k == m:
        S = ol.host_gram(V)
        assert np.allclose(S, V.T @ V, rtol=1e-12, atol=1e-12 * max(1, n))


def test_host_gram_exact_on_integers():
    rng = np.random.default_rng(5)
    V = rng.integers(-8, 9, (40000, 8)).astype(float)
    G = ol.host_gram(V)
    if G is None:
        pytest.skip("oracle/libhost_gram.so is not built (make -C oracle)")
    assert np.array_equal(G, (V.T @ V))


def test_host_gram_rejects_bad_sizes():
    if ol.host_gram(np.zeros((4, 2))) is None:
        pytest.skip("oracle/libhost_gram.so is not built (make -C oracle)")
    lib = ctypes.CDLL(ol.GRAM_SO)
    buf = np.zeros(4)
    for n, k, m in ((-1, 1, 1), (4, 0, 1), (4, 65, 1), (4, 1, 0)):
        assert lib.host_gram(ol.p(buf), ol.p(buf), ctypes.c_int64(n), ctypes.c_int(k), ctypes.c_int(m), ol.p(buf)) == -1
