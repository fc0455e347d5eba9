"""CPU: the row-blocked Gram of oracle/blocks_port.py (the CPU reference of the
Gram and CholQR benches) equals V'V to rounding for narrow and wide blocks,
ragged row counts and one thread, and a planted error (a dropped row block)
is caught."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import blocks_port as bp  # noqa: E402


@pytest.mark.parametrize("k", [1, 3, 8, 16, 32, 40])
@pytest.mark.parametrize("n,threads", [(100003, 4), (4096, 1), (65, 8)])
def test_gram_rows(k, n, threads):
    V = np.random.default_rng(k + n).uniform(-1, 1, (n, k))
    G = bp.gram(V, threads)
    ref = V.T @ V
    assert np.max(np.abs(G - ref)) <= 1e-12 * n
    assert np.allclose(G, G.T, rtol=0, atol=1e-12 * n)
    if n > 1000:  # the check above sees a lost block of rows
        assert np.max(np.abs(G - V[n // 2:].T @ V[n // 2:])) > 1e-3
