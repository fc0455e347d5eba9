"""CPU: the C-ABI library loads and exports every symbol include/*.h
declares, and its host-only pieces (KISS jump-ahead) agree with the oracle.
No GPU calls."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle_lib as ol

ROOT = ol.ROOT


def declared_symbols():
    inc = os.path.join(ROOT, "include")
    names = set()
    for h in ("phist_amd.h", "phist_amd_ext.h"):
        src = open(os.path.join(inc, h)).read()
        names.update(re.findall(r"^\s*(?:void|int)\s+(phist_\w+)\s*\(", src, re.M))
    decl = open(os.path.join(inc, "phist_amd_decl.h")).read()
    for n in re.findall(r"^\s*void\s+SUBR_\((\w+)\)\s*\(", decl, re.M):
        names.add("phist_D" + n)
        names.add("phist_S" + n)
    return sorted(names)


def test_header_declares_the_hot_path():
    names = set(declared_symbols())
    for t in "DS":
        for n in ["sparseMat_times_mvec", "sparseMat_times_mvec_vadd_mvec", "fused_spmv_mvdot", "mvecT_times_mvec",
                  "mvec_times_sdMat", "mvec_times_sdMat_inplace", "mvec_dot_mvec", "chol_QR", "svqb", "orthog",
                  "fused_mvsd_mvTmv", "fused_mvsdi_mvTmv", "sparseMat_create_fromRowFunc", "mvec_random"]:
            assert f"phist_{t}{n}" in names


def test_library_exports_every_declared_symbol():
    import phist_amd
    lib = phist_amd.lib()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("skip", [0, 1, 4095, 4096, 10**6 + 3, 2**33 + 17])
def test_kiss_jump_ahead_matches_stepping(skip):
    import phist_amd
    lib = phist_amd.lib()
    iflag = ctypes.c_int(0)
    lib.phist_amd_kiss_reset(ctypes.byref(iflag))
    lib.phist_amd_kiss_skip(ctypes.c_uint64(skip), ctypes.byref(iflag))
    got = np.empty(5)
    lib.phist_amd_kiss_draw(ctypes.c_int64(5), got.ctypes.data_as(ctypes.c_void_p), ctypes.byref(iflag))
    assert iflag.value == 0
    if skip < 10**7:
        want = ol.kiss_doubles(5, skip_before=skip)
    else:
        # 2^33 steps are too many to walk; compose two jumps instead
        lib.phist_amd_kiss_reset(ctypes.byref(iflag))
        lib.phist_amd_kiss_skip(ctypes.c_uint64(2**33), ctypes.byref(iflag))
        lib.phist_amd_kiss_skip(ctypes.c_uint64(17), ctypes.byref(iflag))
        want = np.empty(5)
        lib.phist_amd_kiss_draw(ctypes.c_int64(5), want.ctypes.data_as(ctypes.c_void_p), ctypes.byref(iflag))
    assert np.array_equal(got, want)


def test_kiss_parallel_draw_equals_sequential():
    """segments generated in parallel (jump per segment) equal one stream"""
    import phist_amd
    lib = phist_amd.lib()
    iflag = ctypes.c_int(0)
    n = 3 * (1 << 20) + 12345
    lib.phist_amd_kiss_reset(ctypes.byref(iflag))
    got = np.empty(n)
    lib.phist_amd_kiss_draw(ctypes.c_int64(n), got.ctypes.data_as(ctypes.c_void_p), ctypes.byref(iflag))
    assert np.array_equal(got, ol.kiss_doubles(n))
    # the stream continues after the draw
    nxt = np.empty(3)
    lib.phist_amd_kiss_draw(ctypes.c_int64(3), nxt.ctypes.data_as(ctypes.c_void_p), ctypes.byref(iflag))
    assert np.array_equal(nxt, ol.kiss_doubles(n + 3)[n:])


def test_kiss_known_answer_via_library_jump():
    """Marsaglia's 10^8 check through the library's jump-ahead: the raw
    value is recovered from the double (exact for this draw's magnitude)"""
    import phist_amd
    lib = phist_amd.lib()
    iflag = ctypes.c_int(0)
    lib.phist_amd_kiss_reset(ctypes.byref(iflag))
    lib.phist_amd_kiss_skip(ctypes.c_uint64(10**8 - 1), ctypes.byref(iflag))
    d = np.empty(1)
    lib.phist_amd_kiss_draw(ctypes.c_int64(1), d.ctypes.data_as(ctypes.c_void_p), ctypes.byref(iflag))
    expect = 1.0 / 9223372036854775808.0 * float(1666297717051644203) - 1.0
    assert d[0] == expect


def test_no_gpu_call_needs_a_device_or_fails_cleanly():
    """without a device, init reports an error code instead of crashing"""
    import phist_amd
    lib = phist_amd.lib()
    iflag = ctypes.c_int(0)
    lib.phist_kernels_init(None, None, ctypes.byref(iflag))
    assert iflag.value in (0, -1)
