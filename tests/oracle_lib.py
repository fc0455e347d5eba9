"""ctypes access to the oracle (oracle/liboracle.so, the C restatement of the
builtin kernels) and to the reference's own leaf kernels compiled by
oracle/Makefile into oracle/_ref/libbuiltin_leaf.so.  Test code only."""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")
REF_SO = os.path.join(ORACLE_DIR, "_ref", "libbuiltin_leaf.so")
STREAM_SO = os.path.join(ORACLE_DIR, "libhost_stream.so")
GRAM_SO = os.path.join(ORACLE_DIR, "libhost_gram.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")
MATRICES = os.path.join(GOLDEN, "matrices")

c_i64p = ctypes.POINTER(ctypes.c_int64)
c_i32p = ctypes.POINTER(ctypes.c_int32)
c_dp = ctypes.POINTER(ctypes.c_double)


def p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Kiss(ctypes.Structure):
    _fields_ = [("x", ctypes.c_uint64), ("y", ctypes.c_uint64), ("z", ctypes.c_uint64), ("c", ctypes.c_uint64)]


_oracle = None


def oracle():
    global _oracle
    if _oracle is None:
        srcs = [os.path.join(ORACLE_DIR, f) for f in ("builtin_oracle.c", "matfuncs_oracle.c", "carp_oracle.c")]
        if not os.path.exists(ORACLE_SO) or os.path.getmtime(ORACLE_SO) < max(map(os.path.getmtime, srcs)):
            subprocess.check_call(["make", "-s", "-C", ORACLE_DIR, "all"])
        _oracle = ctypes.CDLL(ORACLE_SO)
        _oracle.oracle_kiss_raw.restype = ctypes.c_uint64
        _oracle.oracle_kiss_double.restype = ctypes.c_double
    return _oracle


def ref_leaf():
    """the flang-built builtin leaf kernels, or None if not built"""
    if not os.path.exists(REF_SO):
        return None
    return ctypes.CDLL(REF_SO)


def _process_cpus():
    """the CPUs of this process's cgroup (a thread's own affinity can be narrower)"""
    try:
        with open("/sys/fs/cgroup/cpuset.cpus.effective") as f:
            cpus = set()
            for part in f.read().strip().split(","):
                if part:
                    a, _, b = part.partition("-")
                    cpus.update(range(int(a), int(b or a) + 1))
            if cpus:
                return cpus
    except OSError:
        pass
    return set(range(os.cpu_count() or 1))


def host_stream_triad(n=1 << 27, reps=10):
    """the host's STREAM triad z = x + a*y over n doubles on the OpenMP threads
    (oracle/host_stream.c): (best, mean) GB/s of reps runs at 24 bytes per
    element, or None if that library is not built.  An OpenMP runtime that ran
    before in this process (the reference's flang-built kernels, with
    OMP_PROC_BIND) may have bound the calling thread to one core, and the
    triad's runtime takes its thread places from that thread's affinity: the
    affinity is widened to the process's CPUs first."""
    if not os.path.exists(STREAM_SO):
        return None
    try:
        os.sched_setaffinity(0, _process_cpus())
    except OSError:
        pass
    lib = ctypes.CDLL(STREAM_SO)
    best, mean = ctypes.c_double(), ctypes.c_double()
    rc = lib.host_stream_triad(ctypes.c_int64(n), ctypes.c_int(reps), ctypes.byref(best), ctypes.byref(mean))
    if rc != 0:
        raise RuntimeError(f"host_stream_triad returned {rc}")
    return best.value, mean.value


def host_gram(V, W=None):
    """V'W (k x m) of row-major FP64 blocks V (n x k) and W (n x m; None: V) on the host's OpenMP threads
    (oracle/host_gram.c, the row loop of the reference's gemm_kernels_sC.F90), or None if that library
    is not built; the calling thread's affinity is widened first as for host_stream_triad"""
    if not os.path.exists(GRAM_SO):
        return None
    try:
        os.sched_setaffinity(0, _process_cpus())
    except OSError:
        pass
    lib = ctypes.CDLL(GRAM_SO)
    V = np.ascontiguousarray(V, np.float64)
    W = V if W is None else np.ascontiguousarray(W, np.float64)
    n, k = V.shape
    m = W.shape[1]
    M = np.zeros((m, k))  # column-major k x m = row-major m x k
    rc = lib.host_gram(p(V), p(W), ctypes.c_int64(n), ctypes.c_int(k), ctypes.c_int(m), p(M))
    if rc != 0:
        raise RuntimeError(f"host_gram returned {rc}")
    return M.T.copy()


class Crs:
    """builtin CRS: 1-based int64 row_offset, 1-based int32 col_idx, values"""

    def __init__(self, nrows, ncols, row_offset, col_idx, val):
        self.nrows, self.ncols = nrows, ncols
        self.row_offset, self.col_idx, self.val = row_offset, col_idx, val

    @property
    def nnz(self):
        return int(self.row_offset[-1] - 1)

    def dense(self):
        A = np.zeros((self.nrows, self.ncols))
        for i in range(self.nrows):
            for k in range(self.row_offset[i] - 1, self.row_offset[i + 1] - 1):
                A[i, self.col_idx[k] - 1] += self.val[k]
        return A


def crs_from_mm(path):
    o = oracle()
    nr, nc = ctypes.c_int64(), ctypes.c_int64()
    ro, ci, va = c_i64p(), c_i32p(), c_dp()
    rc = o.oracle_crs_from_mm(path.encode(), ctypes.byref(nr), ctypes.byref(nc), ctypes.byref(ro), ctypes.byref(ci),
                              ctypes.byref(va))
    assert rc == 0, rc
    n = nr.value
    row_offset = np.ctypeslib.as_array(ro, (n + 1,)).copy()
    nnz = int(row_offset[-1] - 1)
    col = np.ctypeslib.as_array(ci, (max(1, nnz),))[:nnz].copy()
    val = np.ctypeslib.as_array(va, (max(1, nnz),))[:nnz].copy()
    for q in (ro, ci, va):
        o.oracle_free(q)
    return Crs(n, nc.value, row_offset, col, val)


def crs_from_rows(nrows, ncols, rows):
    """rows: list of (cols0based, vals) in fill order -> sorted builtin CRS"""
    o = oracle()
    ptr = np.zeros(nrows + 1, np.int64)
    for i, (c, _) in enumerate(rows):
        ptr[i + 1] = ptr[i] + len(c)
    gcol = np.concatenate([np.asarray(c, np.int64) for c, _ in rows]) if ptr[-1] else np.zeros(1, np.int64)
    gval = np.concatenate([np.asarray(v, np.float64) for _, v in rows]) if ptr[-1] else np.zeros(1)
    ro = np.empty(nrows + 1, np.int64)
    ci = np.empty(max(1, ptr[-1]), np.int32)
    va = np.empty(max(1, ptr[-1]))
    o.oracle_crs_sort_rows(ctypes.c_int64(nrows), p(ptr), p(gcol), p(gval), p(ro), p(ci), p(va))
    return Crs(nrows, ncols, ro, ci[:ptr[-1]], va[:ptr[-1]])


def spmv(A, alpha, x, beta, y, shifts=None, nt=False):
    """returns alpha*(A + diag(shifts))x + beta*y (builtin FP order)"""
    o = oracle()
    x = np.ascontiguousarray(x, np.float64)
    y = np.array(y, np.float64, copy=True, order="C")
    nvec = x.shape[1]
    sh = None if shifts is None else np.ascontiguousarray(shifts, np.float64)
    o.oracle_spmv(ctypes.c_int64(A.nrows), p(A.row_offset), p(A.col_idx), p(A.val), ctypes.c_int(nvec),
                  ctypes.c_double(alpha), None if sh is None else p(sh), p(x), ctypes.c_int64(nvec),
                  ctypes.c_double(beta), p(y), ctypes.c_int64(nvec), ctypes.c_int(1 if nt else 0))
    return y


def dot(v, w):
    o = oracle()
    v = np.ascontiguousarray(v, np.float64)
    w = np.ascontiguousarray(w, np.float64)
    out = np.empty(v.shape[1])
    o.oracle_dot(ctypes.c_int64(v.shape[0]), ctypes.c_int(v.shape[1]), p(v), ctypes.c_int64(v.shape[1]), p(w),
                 ctypes.c_int64(w.shape[1]), p(out))
    return out


def nrm2(v):
    o = oracle()
    v = np.ascontiguousarray(v, np.float64)
    out = np.empty(v.shape[1])
    o.oracle_nrm2(ctypes.c_int64(v.shape[0]), ctypes.c_int(v.shape[1]), p(v), ctypes.c_int64(v.shape[1]), p(out))
    return out


def gram(alpha, v, w, beta, c):
    o = oracle()
    v = np.ascontiguousarray(v, np.float64)
    w = np.ascontiguousarray(w, np.float64)
    cf = np.asfortranarray(np.array(c, np.float64))
    o.oracle_gram(ctypes.c_int64(v.shape[0]), ctypes.c_int(v.shape[1]), ctypes.c_int(w.shape[1]), ctypes.c_double(alpha),
                  p(v), ctypes.c_int64(v.shape[1]), p(w), ctypes.c_int64(w.shape[1]), ctypes.c_double(beta), p(cf),
                  ctypes.c_int(v.shape[1]))
    return np.ascontiguousarray(cf)


def times_sdmat(alpha, v, c, beta, w):
    o = oracle()
    v = np.ascontiguousarray(v, np.float64)
    cf = np.asfortranarray(np.asarray(c, np.float64))
    w = np.array(w, np.float64, copy=True, order="C")
    k, m = cf.shape
    o.oracle_times_sdmat(ctypes.c_int64(v.shape[0]), ctypes.c_int(k), ctypes.c_int(m), ctypes.c_double(alpha), p(v),
                         ctypes.c_int64(v.shape[1]), p(cf), ctypes.c_int(k), ctypes.c_double(beta), p(w),
                         ctypes.c_int64(w.shape[1]))
    return w


def times_sdmat_inplace(v, m):
    o = oracle()
    v = np.array(v, np.float64, copy=True, order="C")
    mf = np.asfortranarray(np.asarray(m, np.float64))
    k, mm = mf.shape
    o.oracle_times_sdmat_inplace(ctypes.c_int64(v.shape[0]), ctypes.c_int(k), ctypes.c_int(mm), p(v),
                                 ctypes.c_int64(v.shape[1]), p(mf), ctypes.c_int(k))
    return v[:, :mm]


def cholesky(m, rank_tol):
    o = oracle()
    a = np.asfortranarray(np.array(m, np.float64))
    n = a.shape[0]
    perm = np.zeros(n, np.int32)
    rank = ctypes.c_int()
    o.oracle_cholesky(p(a), ctypes.c_int(n), ctypes.c_int(n), p(perm), ctypes.byref(rank), ctypes.c_double(rank_tol))
    return np.ascontiguousarray(a), perm, rank.value


def kiss_doubles(n, skip_before=0):
    o = oracle()
    s = Kiss()
    o.oracle_kiss_init(ctypes.byref(s))
    if skip_before:
        o.oracle_kiss_skip(ctypes.byref(s), ctypes.c_uint64(skip_before))
    out = np.empty(n)
    o.oracle_kiss_fill(ctypes.byref(s), ctypes.c_int64(n), p(out))
    return out


def mvec_random_sequence(shapes):
    """values of consecutive mvec_random calls on one process (row-major)"""
    o = oracle()
    s = Kiss()
    o.oracle_kiss_init(ctypes.byref(s))
    outs = []
    for (n, k) in shapes:
        v = np.empty((n, k))
        o.oracle_mvec_random(ctypes.byref(s), ctypes.c_int64(n), ctypes.c_int(k), p(v), ctypes.c_int64(k))
        outs.append(v)
    return outs


# ---- problem generators (drivers/matfuncs restatements) ----

def spin_sz_rows(L, nup, obc=0):
    """rows of the S_z-sector spin chain from the oracle: list of (cols, vals)"""
    o = oracle()
    o.oracle_spin_sz_init.restype = ctypes.c_void_p
    n = ctypes.c_int64(0)
    h = o.oracle_spin_sz_init(ctypes.c_int(L), ctypes.c_int(nup), ctypes.c_int(obc), ctypes.byref(n))
    cols = np.zeros(L + 1, np.int64)
    vals = np.zeros(L + 1)
    k = ctypes.c_int(0)
    rows = []
    for r in range(n.value):
        rc = o.oracle_spin_sz_row(ctypes.c_void_p(h), ctypes.c_int64(r), ctypes.byref(k), p(cols), p(vals))
        assert rc == 0
        rows.append((cols[:k.value].copy(), vals[:k.value].copy()))
    o.oracle_spin_sz_free(ctypes.c_void_p(h))
    return rows


def honeycomb_rows(W, L, pbc_w=0, pbc_l=0, long_range=0, onsite=None):
    """rows of the honeycomb sheet from the oracle; onsite[r] is the disorder part of row r"""
    o = oracle()
    cols = np.zeros(16, np.int64)
    vals = np.zeros(16)
    k = ctypes.c_int(0)
    rows = []
    for r in range(W * L):
        v = 0.0 if onsite is None else float(onsite[r])
        o.oracle_honeycomb_row(ctypes.c_int64(W), ctypes.c_int64(L), ctypes.c_int(pbc_w), ctypes.c_int(pbc_l),
                               ctypes.c_int(long_range), ctypes.c_double(v), ctypes.c_int64(r), ctypes.byref(k),
                               p(cols), p(vals))
        rows.append((cols[:k.value].copy(), vals[:k.value].copy()))
    return rows


def rows_to_dense(rows, n):
    A = np.zeros((n, n))
    for r, (c, v) in enumerate(rows):
        for cc, vv in zip(c, v):
            A[r, cc] += vv
    return A


def kacz_rows(ro, ci, va, sr, omega, xr, seq, b=None, si=None, xi=None):
    """oracle_kacz_rows_{d,f}: Kaczmarz row updates over rows seq (0-based), in place on xr (and xi);
    ro/ci 1-based builtin CRS, mvecs row-major (n, nvec)"""
    dt = xr.dtype
    f = oracle().oracle_kacz_rows_d if dt == np.float64 else oracle().oracle_kacz_rows_f
    nvec = xr.shape[1]
    ro = np.ascontiguousarray(ro, np.int64)
    ci = np.ascontiguousarray(ci, np.int32)
    va = np.ascontiguousarray(va, dt)
    sr = np.ascontiguousarray(np.broadcast_to(np.asarray(sr, dt), (nvec,)))
    om = np.ascontiguousarray(np.broadcast_to(np.asarray(omega, dt), (nvec,)))
    si_a = None if si is None else np.ascontiguousarray(np.broadcast_to(np.asarray(si, dt), (nvec,)))
    seq = np.ascontiguousarray(seq, np.int32)
    assert xr.flags.c_contiguous and (xi is None or (xi.flags.c_contiguous and xi.shape == xr.shape))
    if b is not None:
        b = np.ascontiguousarray(b, dt)
        assert b.shape == xr.shape
    f(p(ro), p(ci), p(va), ctypes.c_int(nvec), p(sr), p(si_a) if si_a is not None else None, p(om),
      p(b) if b is not None else None, ctypes.c_int64(nvec), p(xr), p(xi) if xi is not None else None,
      ctypes.c_int64(nvec), p(seq), ctypes.c_int64(len(seq)))
