"""ctypes binding of libphist_kernels_amd.so, the MI355X kernel library for PHIST.

This plays the role of PHIST's own ctypes loader (python/phist_kernels.py in
the PHIST sources, which loads libphist_kernels_<lib>.so): no arithmetic
happens in Python, every call goes through the C ABI declared in
include/phist_amd.h.  The thin object wrappers (Map, Mvec, SdMat, SparseMat)
only hold handles and convert numpy arrays at the boundary.

The library needs a HIP device; importing this module works without one (the
shared object loads), but the first call that touches the GPU fails with the
library's own error code.  There is no CPU fallback.
"""
import ctypes
import os

import numpy as np

__all__ = ["lib", "Backend", "PhistError", "LIB_PATH", "build"]

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libphist_kernels_amd.so")

# return codes (include/phist_amd.h)
SUCCESS = 0
NOT_IMPLEMENTED = -99
INVALID_INPUT = -55
BAD_CAST = -88
# input flags
ROBUST_REDUCTIONS = 1
NO_GLOBAL_REDUCTION = 2
MVEC_RUN_ON_DEVICE = 2048
ORTHOG_RANDOMIZE_NULLSPACE = 131072

ROWFUNC = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_int64, ctypes.POINTER(ctypes.c_int32),
                           ctypes.POINTER(ctypes.c_int64), ctypes.c_void_p, ctypes.c_void_p)


class PhistError(RuntimeError):
    def __init__(self, name, code):
        super().__init__(f"{name} returned iflag={code}")
        self.code = code


class BandedParams(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("nnz_per_row", ctypes.c_int32), ("halfband", ctypes.c_int64),
                ("diag", ctypes.c_double), ("seed", ctypes.c_uint64)]


class Laplace3dParams(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int64), ("ny", ctypes.c_int64), ("nz", ctypes.c_int64)]


class SpinSzParams(ctypes.Structure):
    _fields_ = [("L", ctypes.c_int32), ("nup", ctypes.c_int32), ("obc", ctypes.c_int32), ("pad_", ctypes.c_int32),
                ("jz", ctypes.c_double), ("jxy", ctypes.c_double)]


class HoneycombParams(ctypes.Structure):
    _fields_ = [("W", ctypes.c_int64), ("L", ctypes.c_int64), ("pbc_w", ctypes.c_int32), ("pbc_l", ctypes.c_int32),
                ("long_range", ctypes.c_int32), ("pad_", ctypes.c_int32), ("gamma", ctypes.c_double),
                ("seed", ctypes.c_uint64)]


def generator_dims(kind, params):
    """(nrows, max entries per row) of a library generator ('spin_sz' or 'honeycomb')"""
    n, mx, fl = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int(0)
    getattr(lib(), f"phist_amd_{'spinchain_sz' if kind == 'spin_sz' else kind}_dims")(
        ctypes.byref(params), ctypes.byref(n), ctypes.byref(mx), ctypes.byref(fl))
    if fl.value:
        raise PhistError(f"{kind}_dims", fl.value)
    return n.value, mx.value


def generator_row(kind, params, row, dtype=np.float64):
    """one row (cols, vals) of a library generator, computed on the host"""
    n, mx = generator_dims(kind, params)
    cols = np.zeros(mx, np.int64)
    vals = np.zeros(mx, dtype)
    k = ctypes.c_int32(0)
    name = {"spin_sz": "spinchain_sz", "honeycomb": "honeycomb"}[kind]
    f = getattr(lib(), f"phist_amd_{name}_rowfunc" + ("" if dtype == np.float64 else "_f"))
    rc = f(ctypes.c_int64(row), ctypes.byref(k), cols.ctypes.data_as(ctypes.c_void_p),
           vals.ctypes.data_as(ctypes.c_void_p), ctypes.byref(params))
    if rc != 0:
        raise PhistError(f"{name}_rowfunc", rc)
    return cols[:k.value].copy(), vals[:k.value].copy()


class JadaOpts(ctypes.Structure):
    """phist_jadaOpts (src/jada/phist_jadaOpts.h:33-199)"""
    _fields_ = [("numEigs", ctypes.c_int), ("which", ctypes.c_int), ("convTol", ctypes.c_double),
                ("relConvTol", ctypes.c_int), ("symmetry", ctypes.c_int), ("how", ctypes.c_int),
                ("maxIters", ctypes.c_int), ("blockSize", ctypes.c_int), ("lookAhead", ctypes.c_int),
                ("minBas", ctypes.c_int), ("maxBas", ctypes.c_int), ("v0", ctypes.c_void_p), ("arno", ctypes.c_int),
                ("initialShift_r", ctypes.c_double), ("initialShift_i", ctypes.c_double),
                ("initialShiftIters", ctypes.c_int), ("innerSolvType", ctypes.c_int),
                ("innerSolvBlockSize", ctypes.c_int), ("innerSolvMaxBas", ctypes.c_int),
                ("innerSolvMaxIters", ctypes.c_int), ("innerSolvBaseTol", ctypes.c_double),
                ("innerSolvRobust", ctypes.c_int), ("innerSolvStopAfterFirstConverged", ctypes.c_int),
                ("innerSolvMaxProjectionSpace", ctypes.c_int), ("preconOp", ctypes.c_void_p),
                ("preconType", ctypes.c_int), ("preconOpts", ctypes.c_char * 1024),
                ("preconSkewProject", ctypes.c_int), ("preconUpdate", ctypes.c_int),
                ("customSolver", ctypes.c_void_p), ("customSolver_run", ctypes.c_void_p),
                ("customSolver_run1", ctypes.c_void_p)]


class LinearOp(ctypes.Structure):
    """phist_[DS]linearOp (src/core/phist_operator_decl.h:14-74); function members as raw pointers"""
    _fields_ = [("A", ctypes.c_void_p), ("aux", ctypes.c_void_p), ("range_map", ctypes.c_void_p),
                ("domain_map", ctypes.c_void_p), ("use_transpose", ctypes.c_int), ("shifts", ctypes.c_void_p),
                ("apply", ctypes.c_void_p), ("applyT", ctypes.c_void_p), ("apply_shifted", ctypes.c_void_p),
                ("fused_apply_mvTmv", ctypes.c_void_p), ("update", ctypes.c_void_p), ("destroy", ctypes.c_void_p)]


EIGSORT = {"LM": 1, "SM": 2, "LR": 3, "SR": 4}
LINSOLV = {"none": 0, "GMRES": 1, "MINRES": 2}
HERMITIAN = 1

_lib = None


def build(jobs=8):
    """Compile the library in-tree (hipcc for gfx950)."""
    import subprocess
    subprocess.check_call(["make", "-C", os.path.join(HERE, "csrc"), f"-j{jobs}"])


def lib():
    """The loaded shared library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C phist_amd/csrc` "
                               "(or __graft_entry__.build())")
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Backend:
    """Typed front end: dtype float64 -> phist_D*, float32 -> phist_S*."""

    def __init__(self, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        if self.dtype == np.float64:
            self.p, self.ct = "D", ctypes.c_double
        elif self.dtype == np.float32:
            self.p, self.ct = "S", ctypes.c_float
        else:
            raise ValueError("dtype must be float64 or float32")
        self.L = lib()

    # -- raw calls -----------------------------------------------------------
    def fn(self, name):
        return getattr(self.L, f"phist_{self.p}{name}")

    def call(self, name, *args, flags=0, allow=(0,)):
        iflag = ctypes.c_int(flags)
        self.fn(name)(*args, ctypes.byref(iflag))
        if iflag.value not in allow and iflag.value < 0:
            raise PhistError(f"phist_{self.p}{name}", iflag.value)
        return iflag.value

    def gcall(self, name, *args, flags=0):
        """type-independent call phist_<name>"""
        iflag = ctypes.c_int(flags)
        getattr(self.L, f"phist_{name}")(*args, ctypes.byref(iflag))
        if iflag.value < 0:
            raise PhistError(f"phist_{name}", iflag.value)
        return iflag.value

    def s(self, x):
        return self.ct(x)

    # -- objects -------------------------------------------------------------
    def init(self):
        self.gcall("kernels_init", None, None)

    def map(self, n):
        return Map(self, n)

    def mvec(self, map_, nvec, data=None):
        v = Mvec(self, map_, nvec)
        if data is not None:
            v.set(data)
        return v

    def sdmat(self, nrows, ncols, data=None):
        m = SdMat(self, nrows, ncols)
        if data is not None:
            m.set(data)
        return m

    def read_mm(self, path):
        h = ctypes.c_void_p()
        self.call("sparseMat_read_mm", ctypes.byref(h), None, path.encode())
        return SparseMat(self, h)

    def from_rowfunc(self, n, maxnne, rowfunc, arg, ncols=None, parallel=False):
        """rowfunc: a ctypes function pointer (ROWFUNC or a library symbol)."""
        h = ctypes.c_void_p()
        nc = n if ncols is None else ncols
        if parallel:
            ctor = self.L.phist_amd_rowfunc_shared_ctor
            self.call("sparseMat_create_fromRowFuncWithConstructor", ctypes.byref(h), None, ctypes.c_int64(n),
                      ctypes.c_int64(nc), ctypes.c_int32(maxnne), rowfunc, ctor, arg)
        else:
            self.call("sparseMat_create_fromRowFunc", ctypes.byref(h), None, ctypes.c_int64(n), ctypes.c_int64(nc),
                      ctypes.c_int32(maxnne), rowfunc, arg)
        return SparseMat(self, h)

    def banded_random(self, n, nnz_per_row, halfband, diag, seed, parallel=True):
        self._banded = BandedParams(n, nnz_per_row, halfband, diag, seed)
        f = self.L.phist_amd_banded_random_rowfunc if self.p == "D" else self.L.phist_amd_banded_random_rowfunc_f
        return self.from_rowfunc(n, nnz_per_row, f, ctypes.byref(self._banded), parallel=parallel)

    def laplace3d(self, nx, ny, nz):
        self._lap = Laplace3dParams(nx, ny, nz)
        f = self.L.phist_amd_laplace3d_rowfunc if self.p == "D" else self.L.phist_amd_laplace3d_rowfunc_f
        return self.from_rowfunc(nx * ny * nz, 7, f, ctypes.byref(self._lap), parallel=True)

    def spinchain_sz(self, L, nup=None, obc=0, jz=1.0, jxy=1.0):
        """XXZ/Heisenberg chain of L spins in the sector of nup up spins (default L/2)"""
        self._spin = SpinSzParams(L, L // 2 if nup is None else nup, obc, 0, jz, jxy)
        n, mx = generator_dims("spin_sz", self._spin)
        f = self.L.phist_amd_spinchain_sz_rowfunc if self.p == "D" else self.L.phist_amd_spinchain_sz_rowfunc_f
        return self.from_rowfunc(n, mx, f, ctypes.byref(self._spin), parallel=True)

    def honeycomb(self, W, L, pbc_w=0, pbc_l=0, long_range=0, gamma=0.0, seed=0):
        """tight-binding honeycomb sheet of W x L sites"""
        self._honey = HoneycombParams(W, L, pbc_w, pbc_l, long_range, 0, gamma, seed)
        n, mx = generator_dims("honeycomb", self._honey)
        f = self.L.phist_amd_honeycomb_rowfunc if self.p == "D" else self.L.phist_amd_honeycomb_rowfunc_f
        return self.from_rowfunc(n, mx, f, ctypes.byref(self._honey), parallel=True)

    # -- solvers -------------------------------------------------------------
    def subspacejada(self, A, num_eigs, which="SR", tol=1e-8, block_size=1, min_bas=None, max_bas=None,
                     max_iters=300, inner="MINRES", inner_max_iters=20, inner_base_tol=0.1, inner_fp32=False,
                     rel_tol=0, v0=None, lookahead=-1):
        """phist_Xsubspacejada on a wrapped sparse matrix (symmetric, standard).
        Returns (ev, resNorm, nConv, nIter, Q, R) with Q an Mvec and R an SdMat."""
        opts = JadaOpts()
        self.L.phist_jadaOpts_setDefaults(ctypes.byref(opts))
        opts.numEigs = num_eigs
        opts.which = EIGSORT[which]
        opts.convTol = tol
        opts.relConvTol = rel_tol
        opts.symmetry = HERMITIAN
        opts.blockSize = block_size
        opts.minBas = min_bas if min_bas is not None else max(10, num_eigs + block_size)
        opts.maxBas = max_bas if max_bas is not None else opts.minBas + 2 * block_size + 10
        opts.maxIters = max_iters
        opts.lookAhead = lookahead
        opts.innerSolvType = LINSOLV[inner]
        opts.innerSolvMaxIters = inner_max_iters
        opts.innerSolvBaseTol = inner_base_tol
        if v0 is not None:
            opts.v0 = v0.h
        op = LinearOp()
        fl = ctypes.c_int(0)
        getattr(self.L, f"phist_{self.p}linearOp_wrap_sparseMat")(ctypes.byref(op), A.h, ctypes.byref(fl))
        if fl.value:
            raise PhistError("linearOp_wrap_sparseMat", fl.value)
        self.L.phist_amd_jada_set_inner_fp32(1 if inner_fp32 else 0, ctypes.byref(fl))
        nq = num_eigs + block_size - 1
        Q = self.mvec(A.domain_map, nq)
        R = self.sdmat(nq, nq)
        ev = np.zeros(nq, self.dtype)
        res = np.zeros(num_eigs, self.dtype)
        nconv, niter = ctypes.c_int(0), ctypes.c_int(0)
        fl = ctypes.c_int(0)
        getattr(self.L, f"phist_{self.p}subspacejada")(ctypes.byref(op), None, opts, Q.h, R.h, _ptr(ev), _ptr(res),
                                                       ctypes.byref(nconv), ctypes.byref(niter), ctypes.byref(fl))
        self.L.phist_amd_jada_set_inner_fp32(0, ctypes.byref(ctypes.c_int(0)))
        if fl.value < 0:
            Q.delete()
            R.delete()
            raise PhistError("subspacejada", fl.value)
        return ev, res, nconv.value, niter.value, Q, R

    def minres_shifted(self, A, shifts, B, X, tol, max_iter, fp32=False):
        """phist_amd_DblockedMINRES_shifted: (A + shifts[j] I) x_j = b_j; returns (iters, relres, unconverged)"""
        k = len(shifts)
        sh = np.ascontiguousarray(shifts, np.float64)
        tl = np.ascontiguousarray(np.broadcast_to(np.asarray(tol, np.float64), (k,)))
        it = np.zeros(k, np.int32)
        rr = np.zeros(k)
        fl = ctypes.c_int(0)
        self.L.phist_amd_DblockedMINRES_shifted(A.h, ctypes.c_int(k), _ptr(sh), B.h, X.h, _ptr(tl),
                                                ctypes.c_int(max_iter), ctypes.c_int(1 if fp32 else 0), _ptr(it),
                                                _ptr(rr), ctypes.byref(fl))
        if fl.value < 0:
            raise PhistError("blockedMINRES_shifted", fl.value)
        return it, rr, fl.value

    def minres_refill(self, A, shifts, B, X, tol, max_iter, slots, fp32=False):
        """phist_amd_DblockedMINRES_shifted_refill: len(shifts) systems through `slots` columns, converged
        columns replaced by the next system; returns (iters, relres, unconverged)"""
        k = len(shifts)
        sh = np.ascontiguousarray(shifts, np.float64)
        tl = np.ascontiguousarray(np.broadcast_to(np.asarray(tol, np.float64), (k,)))
        it = np.zeros(k, np.int32)
        rr = np.zeros(k)
        fl = ctypes.c_int(0)
        self.L.phist_amd_DblockedMINRES_shifted_refill(A.h, ctypes.c_int(k), _ptr(sh), B.h, X.h, _ptr(tl),
                                                       ctypes.c_int(max_iter), ctypes.c_int(1 if fp32 else 0),
                                                       ctypes.c_int(slots), _ptr(it), _ptr(rr), ctypes.byref(fl))
        if fl.value < 0:
            raise PhistError("blockedMINRES_shifted_refill", fl.value)
        return it, rr, fl.value

    def set_minres_graphs(self, enable, iters_per_call=0, graph_len=4):
        """device MINRES: queue up to iters_per_call iterations per step; with enable, replay runs of
        graph_len iterations from captured HIP graphs"""
        self.gcall("amd_set_minres_graphs", ctypes.c_int(1 if enable else 0), ctypes.c_int(iters_per_call),
                   ctypes.c_int(graph_len))

    def minres_graph_stats(self):
        """(graphs captured, graph launches) since initialisation"""
        c, n = ctypes.c_int64(0), ctypes.c_int64(0)
        self.gcall("amd_minres_graph_stats", ctypes.byref(c), ctypes.byref(n))
        return c.value, n.value

    def gmres_shifted(self, A, shifts, B, X, tol, max_iter, restart=30, fp32=False):
        """phist_amd_DblockedGMRES_shifted: (A + shifts[j] I) x_j = b_j, X the initial guess;
        returns (iters, relres, unconverged)"""
        k = len(shifts)
        sh = np.ascontiguousarray(shifts, np.float64)
        tl = np.ascontiguousarray(np.broadcast_to(np.asarray(tol, np.float64), (k,)))
        it = np.zeros(k, np.int32)
        rr = np.zeros(k)
        fl = ctypes.c_int(0)
        self.L.phist_amd_DblockedGMRES_shifted(A.h, ctypes.c_int(k), _ptr(sh), B.h, X.h, _ptr(tl),
                                               ctypes.c_int(max_iter), ctypes.c_int(restart),
                                               ctypes.c_int(1 if fp32 else 0), _ptr(it), _ptr(rr),
                                               ctypes.byref(fl))
        if fl.value < 0:
            raise PhistError("blockedGMRES_shifted", fl.value)
        return it, rr, fl.value

    def server(self, budget_bytes=200 * 2**30, slots=16):
        return Server(self, budget_bytes, slots)

    def carp_setup(self, A, sigma_r, sigma_i=None):
        """phist_XcarpSetup(_rc): a work object for Kaczmarz sweeps over A - sigma_j I"""
        return CarpWork(self, A, sigma_r, sigma_i)

    # -- kernels (thin) ------------------------------------------------------
    def spmv(self, alpha, A, x, beta, y):
        self.call("sparseMat_times_mvec", self.s(alpha), A.h, x.h, self.s(beta), y.h)

    def spmv_shifted(self, alpha, A, shifts, x, beta, y):
        sh = np.ascontiguousarray(shifts, dtype=self.dtype)
        self.call("sparseMat_times_mvec_vadd_mvec", self.s(alpha), A.h, _ptr(sh), x.h, self.s(beta), y.h)

    def fused_spmv_mvdot(self, alpha, A, V, beta, W, want_ww=True, want_vw=True):
        ww = np.zeros(V.nvec, self.dtype)
        vw = np.zeros(V.nvec, self.dtype)
        self.call("fused_spmv_mvdot", self.s(alpha), A.h, V.h, self.s(beta), W.h,
                  _ptr(ww) if want_ww else None, _ptr(vw) if want_vw else None)
        return (ww if want_ww else None), (vw if want_vw else None)

    def dot(self, V, W):
        out = np.zeros(V.nvec, self.dtype)
        self.call("mvec_dot_mvec", V.h, W.h, _ptr(out))
        return out

    def norm2(self, V):
        out = np.zeros(V.nvec, self.dtype)
        self.call("mvec_norm2", V.h, _ptr(out))
        return out

    def gram(self, alpha, V, W, beta, C):
        self.call("mvecT_times_mvec", self.s(alpha), V.h, W.h, self.s(beta), C.h)

    def times_sdmat(self, alpha, V, C, beta, W):
        self.call("mvec_times_sdMat", self.s(alpha), V.h, C.h, self.s(beta), W.h)

    def times_sdmat_inplace(self, V, M):
        self.call("mvec_times_sdMat_inplace", V.h, M.h)

    def synchronize(self):
        self.gcall("amd_synchronize")


class Map:
    def __init__(self, be, n):
        self.be, self.n = be, n
        self.h = ctypes.c_void_p()
        be.gcall("map_create", ctypes.byref(self.h), None, ctypes.c_int64(n))

    def delete(self):
        if self.h:
            self.be.gcall("map_delete", self.h)
            self.h = ctypes.c_void_p()


class Mvec:
    def __init__(self, be, map_, nvec, handle=None):
        self.be, self.map, self.nvec = be, map_, nvec
        self.n = map_.n if map_ is not None else None
        if handle is None:
            self.h = ctypes.c_void_p()
            be.call("mvec_create", ctypes.byref(self.h), map_.h, ctypes.c_int(nvec))
        else:
            self.h = handle

    def set(self, a):
        a = np.ascontiguousarray(np.asarray(a, dtype=self.be.dtype).reshape(self.n, self.nvec))
        self.be.call("mvec_set_data", self.h, _ptr(a), ctypes.c_int32(self.nvec), ctypes.c_int(1))

    def get(self):
        a = np.empty((self.n, self.nvec), dtype=self.be.dtype)
        self.be.call("mvec_get_data", self.h, _ptr(a), ctypes.c_int32(self.nvec), ctypes.c_int(1))
        return a

    def view(self, jmin, jmax):
        vh = ctypes.c_void_p()
        self.be.call("mvec_view_block", self.h, ctypes.byref(vh), ctypes.c_int(jmin), ctypes.c_int(jmax))
        return Mvec(self.be, self.map, jmax - jmin + 1, handle=vh)

    def random(self):
        self.be.call("mvec_random", self.h)

    def put_value(self, v):
        self.be.call("mvec_put_value", self.h, self.be.s(v))

    def delete(self):
        if self.h:
            self.be.call("mvec_delete", self.h)
            self.h = ctypes.c_void_p()


class SdMat:
    def __init__(self, be, nrows, ncols, handle=None):
        self.be, self.nrows, self.ncols = be, nrows, ncols
        if handle is None:
            self.h = ctypes.c_void_p()
            be.call("sdMat_create", ctypes.byref(self.h), ctypes.c_int(nrows), ctypes.c_int(ncols), None)
        else:
            self.h = handle

    def set(self, a):
        a = np.asfortranarray(np.asarray(a, dtype=self.be.dtype).reshape(self.nrows, self.ncols))
        self.be.call("sdMat_set_data", self.h, _ptr(a), ctypes.c_int32(max(1, self.nrows)), ctypes.c_int(0))

    def get(self):
        a = np.zeros((self.nrows, self.ncols), dtype=self.be.dtype, order="F")
        self.be.call("sdMat_from_device", self.h)
        self.be.call("sdMat_get_data", self.h, _ptr(a), ctypes.c_int32(max(1, self.nrows)), ctypes.c_int(0))
        return np.ascontiguousarray(a)

    def delete(self):
        if self.h:
            self.be.call("sdMat_delete", self.h)
            self.h = ctypes.c_void_p()


class SparseMat:
    def __init__(self, be, h):
        self.be, self.h = be, h
        n = ctypes.c_int64()
        be.call("sparseMat_global_nrows", h, ctypes.byref(n))
        self.nrows = n.value
        be.call("sparseMat_global_ncols", h, ctypes.byref(n))
        self.ncols = n.value
        be.call("sparseMat_global_nnz", h, ctypes.byref(n))
        self.nnz = n.value
        m = ctypes.c_void_p()
        be.call("sparseMat_get_domain_map", h, ctypes.byref(m))
        self.domain_map = _BorrowedMap(self.ncols, m)
        be.call("sparseMat_get_range_map", h, ctypes.byref(m))
        self.range_map = _BorrowedMap(self.nrows, m)

    def crs(self):
        """builtin-order CRS (1-based row_offset, col_idx; values)"""
        ro = np.empty(self.nrows + 1, np.int64)
        ci = np.empty(max(1, self.nnz), np.int32)
        va = np.empty(max(1, self.nnz), self.be.dtype)
        iflag = ctypes.c_int(0)
        getattr(self.be.L, f"phist_{self.be.p}amd_sparseMat_get_crs")(self.h, _ptr(ro), _ptr(ci), _ptr(va),
                                                                       ctypes.byref(iflag))
        if iflag.value:
            raise PhistError("get_crs", iflag.value)
        return ro, ci[:self.nnz], va[:self.nnz]

    def sell(self):
        C, sig = ctypes.c_int(), ctypes.c_int()
        nch, nsl = ctypes.c_int64(), ctypes.c_int64()
        f = getattr(self.be.L, f"phist_{self.be.p}amd_sparseMat_get_sell")
        iflag = ctypes.c_int(0)
        f(self.h, ctypes.byref(C), ctypes.byref(sig), ctypes.byref(nch), ctypes.byref(nsl), None, None, None, None,
          None, ctypes.byref(iflag))
        if iflag.value:
            raise PhistError("get_sell", iflag.value)
        cp = np.empty(nch.value + 1, np.int64)
        cw = np.empty(max(1, nch.value), np.int32)
        pm = np.empty(max(1, nch.value * C.value), np.int32)
        co = np.empty(max(1, nsl.value), np.int32)
        va = np.empty(max(1, nsl.value), self.be.dtype)
        f(self.h, None, None, None, None, _ptr(cp), _ptr(cw), _ptr(pm), _ptr(co), _ptr(va), ctypes.byref(iflag))
        if iflag.value:
            raise PhistError("get_sell", iflag.value)
        return dict(C=C.value, sigma=sig.value, nchunks=nch.value, nslots=nsl.value, chunk_ptr=cp,
                    chunk_w=cw[:nch.value], perm=pm[:nch.value * C.value], col=co[:nsl.value], val=va[:nsl.value])

    def delete(self):
        if self.h:
            self.be.call("sparseMat_delete", self.h)
            self.h = ctypes.c_void_p()


class _BorrowedMap:
    """a map owned by a matrix (not deleted by the caller)"""

    def __init__(self, n, h):
        self.n, self.h = n, h


class CarpWork:
    """work object of phist_XcarpSetup / _rc; sweep() = phist_XcarpSweep / _rc
    (kernels in phist_kernels_carp_decl.h:35-102)"""

    def __init__(self, be, A, sigma_r, sigma_i=None):
        self.be, self.A = be, A
        sr = np.ascontiguousarray(sigma_r, be.dtype)
        self.h = ctypes.c_void_p()
        if sigma_i is None:
            be.call("carp_setup", A.h, ctypes.c_int(len(sr)), _ptr(sr), ctypes.byref(self.h))
        else:
            si = np.ascontiguousarray(sigma_i, be.dtype)
            be.call("carp_setup_rc", A.h, ctypes.c_int(len(sr)), _ptr(sr), _ptr(si), ctypes.byref(self.h))

    def sweep(self, sigma_r, B, X, omega, sigma_i=None, Xi=None):
        """forward + backward sweep in place on X (and Xi for complex shifts); B None means b = 0"""
        be = self.be
        sr = np.ascontiguousarray(np.broadcast_to(np.asarray(sigma_r, be.dtype), (X.nvec,)))
        om = np.ascontiguousarray(np.broadcast_to(np.asarray(omega, be.dtype), (X.nvec,)))
        bh = B.h if B is not None else None
        if Xi is None:
            be.call("carp_sweep", self.A.h, _ptr(sr), bh, X.h, self.h, _ptr(om))
        else:
            si = np.ascontiguousarray(np.broadcast_to(np.asarray(sigma_i, be.dtype), (X.nvec,)))
            be.call("carp_sweep_rc", self.A.h, _ptr(sr), _ptr(si), bh, X.h, Xi.h, self.h, _ptr(om))

    def colors(self):
        """(color_ptr, rows): rows[color_ptr[c]:color_ptr[c+1]] have colour c"""
        f = getattr(self.be.L, f"phist_amd_{self.be.p}carp_colors")
        nc, fl = ctypes.c_int(0), ctypes.c_int(0)
        f(self.h, ctypes.byref(nc), None, None, ctypes.byref(fl))
        if fl.value:
            raise PhistError("carp_colors", fl.value)
        cp = np.zeros(nc.value + 1, np.int64)
        rows = np.zeros(max(1, self.A.nrows), np.int32)
        f(self.h, ctypes.byref(nc), _ptr(cp), _ptr(rows), ctypes.byref(fl))
        if fl.value:
            raise PhistError("carp_colors", fl.value)
        return cp, rows[:self.A.nrows]

    def delete(self):
        if self.h:
            self.be.call("carp_destroy", self.A.h, self.h)
            self.h = None


class Server:
    """phist_amd_server_*: continuous-batching loop for shifted solves and JD requests (FP64 interface)"""
    QUEUED, RUNNING, DONE, DONE_UNCONVERGED, CANCELLED = 0, 1, 2, 3, 4

    def __init__(self, be, budget_bytes, slots):
        assert be.p == "D"
        self.be, self.L = be, be.L
        self.h = ctypes.c_void_p()
        fl = ctypes.c_int(0)
        self.L.phist_amd_server_create(ctypes.byref(self.h), ctypes.c_int64(int(budget_bytes)), ctypes.c_int(slots),
                                       ctypes.byref(fl))
        if fl.value:
            raise PhistError("server_create", fl.value)
        self._keep = {}

    def _chk(self, name, fl):
        if fl.value < 0:
            raise PhistError(name, fl.value)

    def submit_solve(self, A, shifts, B, X, tol, max_iter, fp32=False):
        sh = np.ascontiguousarray(shifts, np.float64)
        rid, fl = ctypes.c_int64(0), ctypes.c_int(0)
        self.L.phist_amd_server_submit_solve(self.h, A.h, ctypes.c_int(len(sh)), _ptr(sh), B.h, X.h,
                                             ctypes.c_double(tol), ctypes.c_int(max_iter), ctypes.c_int(int(fp32)),
                                             ctypes.byref(rid), ctypes.byref(fl))
        self._chk("server_submit_solve", fl)
        self._keep[rid.value] = (sh,)
        return rid.value

    def submit_jada(self, A, num_eigs, which="SR", tol=1e-8, block_size=4, min_bas=None, max_bas=None,
                    max_iters=300, inner_max_iters=10, fp32=False, inner="MINRES"):
        """returns (id, Q, R, ev, resNorm); Q/R are Mvec/SdMat owned by the caller.
        `which` is a key of EIGSORT or a raw phist_EigSort value"""
        be = self.be
        opts = JadaOpts()
        self.L.phist_jadaOpts_setDefaults(ctypes.byref(opts))
        opts.numEigs, opts.which, opts.convTol = num_eigs, EIGSORT.get(which, which), tol
        opts.symmetry, opts.blockSize = HERMITIAN, block_size
        opts.minBas = min_bas if min_bas is not None else max(10, num_eigs + block_size)
        opts.maxBas = max_bas if max_bas is not None else opts.minBas + 2 * block_size + 10
        opts.maxIters, opts.innerSolvType, opts.innerSolvMaxIters = max_iters, LINSOLV[inner], inner_max_iters
        nq = num_eigs + block_size - 1
        Q, R = be.mvec(A.domain_map, nq), be.sdmat(nq, nq)
        ev, res = np.zeros(nq), np.zeros(num_eigs)
        rid, fl = ctypes.c_int64(0), ctypes.c_int(0)
        self.L.phist_amd_server_submit_jada(self.h, A.h, ctypes.byref(opts), ctypes.c_int(int(fp32)), Q.h, R.h,
                                            _ptr(ev), _ptr(res), ctypes.byref(rid), ctypes.byref(fl))
        if fl.value < 0:
            Q.delete()
            R.delete()
        self._chk("server_submit_jada", fl)
        self._keep[rid.value] = (opts, ev, res)
        return rid.value, Q, R, ev, res

    def step(self):
        pending, fl = ctypes.c_int(0), ctypes.c_int(0)
        self.L.phist_amd_server_step(self.h, ctypes.byref(pending), ctypes.byref(fl))
        self._chk("server_step", fl)
        return pending.value

    def status(self, rid):
        st, lat, it, fl = ctypes.c_int(0), ctypes.c_double(0), ctypes.c_int(0), ctypes.c_int(0)
        self.L.phist_amd_server_status(self.h, ctypes.c_int64(rid), ctypes.byref(st), ctypes.byref(lat),
                                       ctypes.byref(it), ctypes.byref(fl))
        self._chk("server_status", fl)
        return st.value, lat.value, it.value

    def cancel(self, rid):
        fl = ctypes.c_int(0)
        self.L.phist_amd_server_cancel(self.h, ctypes.c_int64(rid), ctypes.byref(fl))
        self._chk("server_cancel", fl)

    def release(self, rid):
        fl = ctypes.c_int(0)
        self.L.phist_amd_server_release(self.h, ctypes.c_int64(rid), ctypes.byref(fl))
        self._chk("server_release", fl)
        self._keep.pop(rid, None)

    def delete(self):
        if self.h:
            fl = ctypes.c_int(0)
            self.L.phist_amd_server_delete(self.h, ctypes.byref(fl))
            self.h = None
