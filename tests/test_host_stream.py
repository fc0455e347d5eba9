"""CPU: the host STREAM triad reported beside the CPU baselines
(oracle/host_stream.c, after src/kernels/common/stream_bench.c's triad):
it runs, checks its own result (z = x + a*y on every element) and reports a
positive bandwidth; bad arguments are rejected."""
import ctypes

import pytest

import oracle_lib as ol


def test_host_stream_triad_runs():
    r = ol.host_stream_triad(1 << 20, 3)
    if r is None:
        pytest.skip("oracle/libhost_stream.so is not built (make -C oracle)")
    best, mean = r
    assert best > 0 and 0 < mean <= best


def test_host_stream_triad_rejects_bad_sizes():
    r = ol.host_stream_triad(1 << 10, 1)
    if r is None:
        pytest.skip("oracle/libhost_stream.so is not built (make -C oracle)")
    lib = ctypes.CDLL(ol.STREAM_SO)
    b, m = ctypes.c_double(), ctypes.c_double()
    for n, reps in ((1, 1), (3, 1), (0, 1), (8, 0)):
        assert lib.host_stream_triad(ctypes.c_int64(n), ctypes.c_int(reps), ctypes.byref(b), ctypes.byref(m)) == -1
