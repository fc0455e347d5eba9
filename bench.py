#!/usr/bin/env python3
"""Fused block-SpMV benchmark on one MI355X.

Workload (the kernel-bench configuration): random banded sparse matrix with
4,194,304 rows and 32 nonzeros per row (the diagonal 40 plus 31 distinct
columns drawn uniformly within +-50,000 of the diagonal, values U(-1,1)),
stored as SELL-C-sigma with C=32, sigma=256, FP64.  One step is one
phist_DfusedSpmv_mvdot call: W = A*V (alpha=1, beta=0) with both column
dots <W,W> and <V,W>, V an 8-column block of KISS random numbers
(phist_Dmvec_random).  Inputs are resident in HBM when timing starts.

Algorithmic bytes per step (compulsory traffic, x read once, dots free):
    nnz*(8+4) + 4*N + N*nvec*8*2
`value` divides them by the wall-clock time per step; roofline.achieved
divides them by the SpMV kernel's own average duration, taken with HIP
events on the library stream during the timed steps.

cpu_baseline: PHIST's own builtin leaf kernels (dspmvm_8 + two ddot_8,
compiled from the reference sources by oracle/Makefile, OpenMP) on the
first quarter of the same matrix, timed on this host; if that build is
absent, the oracle's single-threaded C restatement.  Beside it the host's
STREAM triad (z = x + a*y, 24 bytes per element, all threads) as the host's
own bandwidth ceiling.
"""
import argparse
import ctypes
import json
import os
import sys
import time

# CPU baseline threads are pinned one per physical core, as the builtin path
# is run (set before any OpenMP runtime starts; a caller's own settings win)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")
# the JD CPU baseline alternates OpenMP kernels (the reference's SpMV) with
# BLAS calls (numpy's OpenBLAS): threads of either runtime go to sleep at once
# after their region instead of spinning on the cores the other one needs
os.environ.setdefault("KMP_BLOCKTIME", "0")
os.environ.setdefault("OPENBLAS_THREAD_TIMEOUT", "4")

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
METRIC = "fused block-SpMV GB/s (% HBM peak); JD time-to-1e-8 mixed vs FP64; p99"


def spmv_bytes(nnz, n, nvec, sval=8, sx=8, beta_nonzero=False):
    return nnz * (sval + 4) + 4 * n + n * nvec * sx * (2 + (1 if beta_nonzero else 0))


def read_traffic(path, key="traffic_bytes_per_launch"):
    """HBM bytes per launch (or per the unit `key` names) from a committed rocprofv3 --pmc summary,
    or None.  The file holds lines `<key>: <float>`."""
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":", 1)[1])
    return None


def host_cpu():
    """(CPU model, physical cores) of this host from /proc/cpuinfo"""
    model, cores = "unknown", set()
    phys = core = None
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and model == "unknown":
                model = v
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
            elif not k and core is not None:
                cores.add((phys, core))
                phys = core = None
        if core is not None:
            cores.add((phys, core))
    except OSError:
        pass
    return model, (len(cores) or os.cpu_count() or 1)


def cpu_info_fields(threads):
    model, phys = host_cpu()
    return {"cpu_model": model, "physical_cores": phys, "threads": threads,
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND"), "omp_places": os.environ.get("OMP_PLACES")}


def cpu_baseline(A, n_sample, nvec, budget_s=12.0):
    """builtin dspmvm_8 + ddot_8 x2 on the first n_sample rows"""
    import oracle_lib as ol
    ro, ci, va = A.crs()
    nnz_s = int(ro[n_sample] - 1)
    ro_s = np.ascontiguousarray(ro[:n_sample + 1])
    ci_s = np.ascontiguousarray(ci[:nnz_s])
    va_s = np.ascontiguousarray(va[:nnz_s])
    del ro, ci, va
    ncols = int(ci_s.max())  # 1-based: x rows 0..ncols-1 are read
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, (ncols, nvec))
    y = np.zeros((n_sample, nvec))
    ref = ol.ref_leaf()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    I, L, D = ctypes.c_int, ctypes.c_int64, ctypes.c_double
    p = ol.p
    if ref is not None and nvec == 8:
        kind = "reference"
        halo_ptr = ro_s[1:].copy()
        halo = np.zeros(8)
        shifts = np.zeros(8)
        ww = np.zeros(8)
        vw = np.zeros(8)

        def step():
            ref.dspmvm_8_(ctypes.byref(I(n_sample)), ctypes.byref(I(0)), ctypes.byref(I(ncols)),
                          ctypes.byref(L(nnz_s)), ctypes.byref(D(1.0)), p(ro_s), p(halo_ptr), p(ci_s), p(va_s),
                          p(shifts), p(x), p(halo), ctypes.byref(D(0.0)), p(y))
            ref.ddot_8_(ctypes.byref(I(n_sample)), p(x), p(y), p(vw))
            ref.ddot_8_(ctypes.byref(I(n_sample)), p(y), p(y), p(ww))
    else:
        kind = "port"
        threads = 1
        crs = ol.Crs(n_sample, ncols, ro_s, ci_s, va_s)

        def step():
            ol.spmv(crs, 1.0, x, 0.0, y, nt=True)
            ol.dot(x[:n_sample], y)
            ol.dot(y, y)
    step()
    times = []
    t_start = time.perf_counter()
    while time.perf_counter() - t_start < budget_s or len(times) < 3:
        t0 = time.perf_counter()
        step()
        times.append(time.perf_counter() - t0)
        if len(times) >= 200:
            break
    t = float(np.median(times))
    gbs = spmv_bytes(nnz_s, n_sample, nvec) / t / 1e9
    sample = (f"first {n_sample} rows ({nnz_s} nnz) of the same matrix, nvec={nvec}, "
              f"spmv (alpha=1, beta=0) + 2 column dots, median of {len(times)} runs")
    del x, y
    # the host's own memory bandwidth beside it: STREAM triad on 3 x 1 GiB
    # arrays, all OpenMP threads (what the reference's stream benchmark reports)
    triad = ol.host_stream_triad(1 << 27, 10)
    return {"value": round(gbs, 2), "unit": "GB/s", "cores": threads, "kind": kind, "sample": sample,
            "ms_per_step": round(t * 1e3, 3),
            **({} if triad is None else {"host_stream_triad_GBs": round(triad[0], 1),
                                         "frac_of_host_triad": round(gbs / triad[0], 3)}),
            **cpu_info_fields(threads)}


def tall_skinny(be, args):
    """Gram (self, V'V), CholQR or fused_mvsd_mvTmv on N = 16,777,216 rows x k columns"""
    L = be.L
    iflag = ctypes.c_int(0)
    N, k = 16777216, args.k
    timed = [False]  # the kernel timer is on (the CholQR step leaves its input refresh out of it)
    m = be.map(N)
    v = be.mvec(m, k)
    v.random()
    C = be.sdmat(k, k)
    if args.op == "gram":
        f = be.fn("mvecT_times_mvec")
        nbytes = N * k * 8

        def step():
            fl = ctypes.c_int(0)
            f(ctypes.c_double(1.0), v.h, v.h, ctypes.c_double(0.0), C.h, ctypes.byref(fl))
            if fl.value != 0:
                raise RuntimeError(f"mvecT_times_mvec iflag={fl.value}")
    elif args.op == "mvsd":
        f = be.fn("fused_mvsd_mvTmv")
        nbytes = 2 * N * k * 8
        w = be.mvec(m, k)
        Cm = be.sdmat(k, k, np.random.default_rng(3).uniform(-1, 1, (k, k)))

        def step():
            fl = ctypes.c_int(0)
            f(ctypes.c_double(1.0), v.h, Cm.h, ctypes.c_double(0.0), w.h, C.h, ctypes.byref(fl))
            if fl.value != 0:
                raise RuntimeError(f"fused_mvsd_mvTmv iflag={fl.value}")
    else:
        f = be.fn("chol_QR")
        nbytes = 3 * N * k * 8
        w0 = be.mvec(m, k)
        be.call("mvec_add_mvec", be.s(1.0), v.h, be.s(0.0), w0.h)

        def step():
            # fresh input, left out of the kernel timer (a no-op while the timer is off)
            if timed[0]:
                L.phist_amd_kernel_timer(2, ctypes.byref(iflag))
            be.call("mvec_add_mvec", be.s(1.0), w0.h, be.s(0.0), v.h)
            if timed[0]:
                L.phist_amd_kernel_timer(3, ctypes.byref(iflag))
            fl = ctypes.c_int(0)
            f(v.h, C.h, ctypes.byref(fl))
            if fl.value != 0:
                raise RuntimeError(f"chol_QR iflag={fl.value}")
    for _ in range(args.warmup):
        step()
    be.synchronize()
    L.phist_amd_kernel_timer(1, ctypes.byref(iflag))
    timed[0] = True
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    be.synchronize()
    t1 = time.perf_counter()
    kms, kcount = ctypes.c_double(0.0), ctypes.c_int64(0)
    L.phist_amd_kernel_timer_read(ctypes.byref(kms), ctypes.byref(kcount), ctypes.byref(iflag))
    L.phist_amd_kernel_timer(0, ctypes.byref(iflag))
    ms_step = (t1 - t0) * 1e3 / args.steps
    if args.op == "cholqr":
        # the step also re-copies the input (2*N*k*8 bytes through k_axpby, not timed by the kernel timer);
        # the kernel timer holds the Gram and the in-place update of each CholQR
        k_ms = kms.value / args.steps
    else:
        k_ms = kms.value / max(1, kcount.value)
    out = {"metric": f"{args.op} GB/s (N=16777216, k={k}, FP64)", "value": round(nbytes / (k_ms * 1e-3) / 1e9, 2),
           "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
           "higher_is_better": True, "dtype": "f64", "data": "synthetic (KISS random block vector)",
           "config": {"workload": f"{args.op} on 16,777,216 x {k} U(-1,1)", "bytes_per_op": nbytes,
                      "kernel_ms_per_op": round(k_ms, 4), "kernel_launches_timed": kcount.value},
           "roofline": {"bound": "hbm", "achieved": round(nbytes / (k_ms * 1e-3) / 1e9, 2), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(nbytes / (k_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
           "cpu_baseline": None if args.no_cpu else tall_skinny_cpu_baseline(args.op, v.get(), nbytes)}
    print(json.dumps(out), flush=True)


def blas1_bench(be, args):
    """column dots <V, W>, norms ||V|| and Y = a X + b Y (mvec_dot_mvec, mvec_norm2, mvec_vadd_mvec with
    per-column a) on N = 16,777,216 rows x k FP64 columns; byte models 2 N k 8, N k 8, 3 N k 8; the
    kernel time is the library's HIP-event timer over the row kernel (the dots' fixed-order second
    stage is not in it)"""
    L = be.L
    iflag = ctypes.c_int(0)
    N, k = 16777216, args.k
    m = be.map(N)
    v, w = be.mvec(m, k), be.mvec(m, k)
    v.random()
    w.random()
    out_h = np.zeros(k)
    alpha = np.linspace(0.5, 1.5, k)
    op = args.op
    if op == "dot":
        nbytes = 2 * N * k * 8

        def step():
            be.call("mvec_dot_mvec", v.h, w.h, out_h.ctypes.data_as(ctypes.c_void_p))
    elif op == "norm2":
        nbytes = N * k * 8

        def step():
            be.call("mvec_norm2", v.h, out_h.ctypes.data_as(ctypes.c_void_p))
    else:
        nbytes = 3 * N * k * 8

        def step():
            be.call("mvec_vadd_mvec", alpha.ctypes.data_as(ctypes.c_void_p), v.h, be.s(0.5), w.h)
    for _ in range(args.warmup):
        step()
    be.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    be.synchronize()
    ms_step = (time.perf_counter() - t0) * 1e3 / args.steps
    L.phist_amd_kernel_timer(1, ctypes.byref(iflag))
    for _ in range(args.steps):
        step()
    be.synchronize()
    kms, kcount = ctypes.c_double(0.0), ctypes.c_int64(0)
    L.phist_amd_kernel_timer_read(ctypes.byref(kms), ctypes.byref(kcount), ctypes.byref(iflag))
    L.phist_amd_kernel_timer(0, ctypes.byref(iflag))
    k_ms = kms.value / max(1, kcount.value)
    gbs = nbytes / (k_ms * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        cpu = blas1_cpu_baseline(op, v.get(), w.get(), alpha, nbytes)
    out = {"metric": f"{op} GB/s (N=16777216, k={k}, FP64)", "value": round(gbs, 2), "unit": "GB/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4), "higher_is_better": True,
           "dtype": "f64", "data": "synthetic (KISS random block vectors)",
           "config": {"workload": f"{op} on 16,777,216 x {k} U(-1,1)", "bytes_per_op": nbytes,
                      "kernel_ms_per_op": round(k_ms, 4), "kernel_launches_timed": kcount.value},
           "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(gbs / HBM_PEAK_GBS, 4)},
           "cpu_baseline": cpu}
    for o in (v, w):
        o.delete()
    m.delete()
    print(json.dumps(out), flush=True)


def blas1_cpu_baseline(op, Vh, Wh, alpha, nbytes, reps=5):
    """the same operation on the host by the reference's flang-built OpenMP leaf kernels (ddot_k,
    dnrm2_k, daxpby_k for k = 1, 2, 4, 8; row-major blocks are their column-major k x N arrays), median
    of `reps` after one untimed run; None for other widths or without that build"""
    import oracle_lib as ol
    ref = ol.ref_leaf()
    k = Vh.shape[1]
    if ref is None or k not in (1, 2, 4, 8):
        return None
    I, D = ctypes.c_int, ctypes.c_double
    p = ol.p
    n = Vh.shape[0]
    V = np.ascontiguousarray(Vh)
    W = np.ascontiguousarray(Wh)
    res = np.zeros(k)
    al = np.ascontiguousarray(alpha, np.float64)
    if op == "dot":
        f = getattr(ref, f"ddot_{k}_")
        run = lambda: f(ctypes.byref(I(n)), p(V), p(W), p(res))
    elif op == "norm2":
        f = getattr(ref, f"dnrm2_{k}_")
        run = lambda: f(ctypes.byref(I(n)), p(V), p(res))
    else:
        f = getattr(ref, f"daxpby_{k}_") if k > 1 else ref.daxpby_1_
        if k == 1:
            run = lambda: f(ctypes.byref(I(n)), ctypes.byref(D(float(al[0]))), p(V), ctypes.byref(D(0.5)), p(W))
        else:
            run = lambda: f(ctypes.byref(I(n)), p(al), p(V), ctypes.byref(D(0.5)), p(W))
    run()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        run()
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(nbytes / t / 1e9, 2), "unit": "GB/s", "cores": threads, "kind": "reference",
            "sample": f"the whole 16,777,216 x {k} operation by the reference's {f.__name__.rstrip('_')} "
                      f"(flang, OpenMP), median of {reps}",
            "ms_per_op": round(t * 1e3, 3), **cpu_info_fields(threads)}


def tall_skinny_cpu_baseline(op, Vh, nbytes, reps=5):
    """the same operation on the host cores, same byte model, median of `reps` after one untimed run:
    gram: V'V by oracle/host_gram.c, the OpenMP row loop of the reference's gemm_kernels_sC.F90 (which
    needs its CMake-generated header); cholqr: that Gram, the k x k Cholesky (LAPACK), then V <- V R^-1
    by the reference's flang-built OpenMP dgemm_sB_generic_inplace; mvsd: W = V C and W'W by the
    reference's dgemm_sB_augmented_generic"""
    import oracle_lib as ol
    ref = ol.ref_leaf()
    n, k = Vh.shape
    I, D = ctypes.c_int, ctypes.c_double
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    if ol.host_gram(Vh[:64]) is None:
        return None
    if op == "gram":
        def run():
            return ol.host_gram(Vh)
        how = (f"V'V by the OpenMP row loop of the reference's gemm_kernels_sC.F90 restated in C "
               f"(oracle/host_gram.c, {threads} threads; the Fortran file is unbuildable here)")
    elif op == "mvsd":
        if ref is None:
            return None
        Cm = np.asfortranarray(np.random.default_rng(3).uniform(-1, 1, (k, k)))
        W = np.zeros((n, k))
        Dm = np.zeros((k, k), order="F")

        def run():
            ref.dgemm_sb_augmented_generic_(ctypes.byref(I(n)), ctypes.byref(I(k)), ctypes.byref(I(k)),
                                            ctypes.byref(D(1.0)), ol.p(Vh), ctypes.byref(I(k)), ol.p(Cm),
                                            ctypes.byref(D(0.0)), ol.p(W), ctypes.byref(I(k)), ol.p(Dm))
        how = "the reference's dgemm_sB_augmented_generic (flang, OpenMP)"
    else:
        if ref is None:
            return None
        V0 = Vh.copy()

        def run():
            np.copyto(Vh, V0)  # fresh input, as on the GPU (outside the byte model there too)
            t = time.perf_counter()
            R = np.linalg.cholesky(ol.host_gram(Vh)).T
            Rinv = np.asfortranarray(np.linalg.inv(R))
            ref.dgemm_sb_generic_inplace_(ctypes.byref(I(n)), ctypes.byref(I(k)), ctypes.byref(I(k)), ol.p(Vh),
                                          ctypes.byref(I(k)), ol.p(Rinv))
            return time.perf_counter() - t
        how = "Gram as for the gram op (oracle/host_gram.c), k x k Cholesky by LAPACK, V R^-1 in place by " \
              "the reference's dgemm_sB_generic_inplace (flang, OpenMP)"
    run()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        r = run()
        times.append(r if isinstance(r, float) else time.perf_counter() - t0)
    sec = float(np.median(times))
    return {"value": round(nbytes / sec / 1e9, 2), "unit": "GB/s", "cores": threads, "kind": "port",
            "ms_per_op": round(1e3 * sec, 3),
            "sample": f"the whole {n} x {k} block (FP64), median of {reps}: {how}", **cpu_info_fields(threads)}


def jada_bench(be, args):
    """Time to 1e-8 of the symmetric block Jacobi-Davidson, FP64 throughout vs
    FP32 correction equations (outer iteration FP64), mean over the same
    --starts start vectors for both."""
    prob = args.problem
    t0 = time.perf_counter()
    if prob.startswith("xxz"):
        L = int(prob[3:])
        A = be.spinchain_sz(L, L // 2)
        cfg = dict(num_eigs=16, block_size=4, min_bas=24, max_bas=48, inner_max_iters=10)
        name = f"XXZ chain L={L} S_z=0 ({A.nrows} rows, {A.nnz} nnz), 16 smallest"
    else:
        n = int(prob[3:])
        A = be.laplace3d(n, n, n)
        cfg = dict(num_eigs=8, block_size=4, min_bas=16, max_bas=32, inner_max_iters=10)
        name = f"3D Laplacian {n}^3 ({A.nrows} rows), 8 smallest"
    t_build = time.perf_counter() - t0
    print(f"bench: {name}; built in {t_build:.1f} s", file=sys.stderr, flush=True)
    tune = None
    if not args.no_tune:
        # set-up, not timed: the library's autotuner picks the SpMV launch for the block width on the
        # matrix and on its FP32 copy (the correction solves of the mixed run)
        pp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        nv, dms, tms, st = (np.array([cfg["block_size"]], np.int32), np.zeros(1), np.zeros(1),
                            np.zeros(7, np.int32))
        iflag = ctypes.c_int(0)
        t0 = time.perf_counter()
        be.L.phist_Damd_sparseMat_tune(A.h, 1, pp(nv), 4, pp(dms), pp(tms), pp(st), ctypes.byref(iflag))
        if iflag.value:
            raise RuntimeError(f"phist_Damd_sparseMat_tune iflag={iflag.value}")
        tune = {"lane_bytes": int(st[2]), "blocks_per_cu": int(st[3]), "depth": int(st[4]), "col_block": int(st[5]), "stage": int(st[6]),
                "untuned_spmv_ms": round(float(dms[0]), 4), "tuned_spmv_ms": round(float(tms[0]), 4),
                "tune_s": round(time.perf_counter() - t0, 2)}
        print(f"bench: autotuner (FP64 matrix; its FP32 copy tuned too) {tune}", file=sys.stderr, flush=True)
    # untimed warm-up (three outer iterations): the solver's blocks come from the
    # allocation caches in both timed runs, not only in the second
    cfg["inner"] = args.inner
    Q0 = be.subspacejada(A, which="SR", tol=1e-8, max_iters=3, inner_fp32=False, **cfg)
    Q0[4].delete()
    Q0[5].delete()
    # several start vectors (the KISS stream from its initial state, then skipped
    # ahead by 7919*s draws): the outer iteration count moves by a few with the
    # rounding of the start-up basis, so one start vector is a noisy sample
    nstart = max(1, args.starts)
    results = {}
    runs = [r for r in (("fp64", False), ("mixed", True)) if args.precision in ("both", r[0])]
    for label, fp32 in runs:
        if fp32:  # the FP32 copy of A is made once per matrix: build it outside the timed run
            b, x = be.mvec(A.domain_map, 1), be.mvec(A.domain_map, 1)
            be.minres_shifted(A, [0.0], b, x, 1.0, 1, fp32=True)
            b.delete()
            x.delete()
        secs, iters, lows, mres, convs = [], [], [], [], []
        for st in range(nstart):
            iflag = ctypes.c_int(0)
            be.L.phist_amd_kiss_reset(ctypes.byref(iflag))
            be.L.phist_amd_kiss_skip(ctypes.c_uint64(7919 * st), ctypes.byref(iflag))
            be.synchronize()
            t0 = time.perf_counter()
            ev, res, nconv, niter, Q, R = be.subspacejada(A, which="SR", tol=1e-8, max_iters=500, inner_fp32=fp32,
                                                          **cfg)
            be.synchronize()
            secs.append(time.perf_counter() - t0)
            iters.append(niter)
            convs.append(nconv)
            lows.append(float(np.min(ev[:nconv])) if nconv else float("nan"))
            mres.append(float(np.max(res[:cfg["num_eigs"]])))
            Q.delete()
            R.delete()
        results[label] = {"seconds": round(float(np.mean(secs)), 4), "outer_iters": round(float(np.mean(iters)), 2),
                          "converged": int(min(convs)), "lowest": round(float(np.min(lows)), 10),
                          "max_res": float(np.max(mres)), "starts": nstart,
                          "seconds_each": [round(t, 4) for t in secs], "outer_iters_each": iters}
        print(f"bench: jada {label}: mean {np.mean(secs):.3f} s, iterations {iters}, converged >= {min(convs)}",
              file=sys.stderr, flush=True)
    speed = results["fp64"]["seconds"] / results["mixed"]["seconds"] if len(results) == 2 else None
    main = results.get("mixed", results.get("fp64"))
    cpu = None
    if not args.no_cpu and (prob.startswith("lap") or args.cpu_jada):
        cpu = jada_cpu_baseline(A, cfg, args.inner)
    out = {"metric": "JD time-to-1e-8 mixed vs FP64", "value": main["seconds"], "unit": "s",
           "n_gpus": 1, "higher_is_better": False, "dtype": "f64 outer / f32 inner",
           "data": "synthetic (generated matrix, KISS random start vectors)",
           "config": {"workload": name, "which": "SR", "tol": 1e-8, **cfg,
                      **({} if tune is None else {"autotuned": tune})},
           **results, "speedup_mixed_vs_fp64": None if speed is None else round(speed, 3),
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)
    A.delete()


def jada_cpu_baseline(A, cfg, inner):
    """The same JD solve on the host cores: oracle/jada_port.py (a numpy port of
    phist_subspacejada with blocked-GMRES correction solves) over the reference's
    own OpenMP SpMV (flang-built builtin dspmvm, else the oracle's C SpMV), from
    the first start vector of the GPU runs (the KISS stream from its initial
    state), FP64 throughout; one solve, wall clock."""
    import oracle_lib as ol
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import jada_port
    if inner != "GMRES":
        return None  # the port solves the correction equations by GMRES only
    ro, ci, va = A.crs()
    n = len(ro) - 1
    op = jada_port.CrsOperator(n, ro, ci, va)
    v0 = ol.mvec_random_sequence([(n, 1)])[0][:, 0]
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    print(f"bench: timing the CPU JD ({op.kind} SpMV, {threads} threads)", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    ev, res, nconv, niter, _q = jada_port.subspacejada(
        op, v0, cfg["num_eigs"], cfg["block_size"], cfg["min_bas"], cfg["max_bas"], 1e-8, cfg["inner_max_iters"],
        max_iters=500)
    secs = time.perf_counter() - t0
    return {"value": round(secs, 3), "unit": "s", "cores": threads, "kind": "port",
            "sample": f"the whole solve from start vector 0 (FP64): {niter} outer iterations, {nconv} converged, "
                      f"max residual {float(np.max(res)):.2e}, lowest {float(np.min(ev[:max(1, nconv)])):.10f}; "
                      f"SpMV: {'the reference builtin dspmvm (flang, OpenMP)' if op.kind == 'reference' else 'oracle C'}",
            **cpu_info_fields(threads)}


def tune_bench(be, args):
    """The SpMV autotuner on the matrices of the workloads: per matrix and block width the kernel time per
    product with the library's untuned choice (default SELL layout, launch table) and with the settings
    the tuner picks over SELL layouts (C 16/32/64, sigma 1/256), gather width per lane, workgroups
    per CU, depth, column blocks and LDS staging of x; one JSON line per matrix."""
    L = be.L
    pp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    cases = [
        ("banded 4194304 rows x 32 nnz (+-50k), FP64", lambda: be.banded_random(4194304, 32, 50000, 40.0, 2024),
         [1, 4, 8, 16, 32]),
        ("XXZ chain L=24 S_z=0, FP64", lambda: be.spinchain_sz(24, 12), [4, 16]),
        ("honeycomb 2000x4000 sites (8M rows), FP64", lambda: be.honeycomb(2000, 4000, 0, 0, 0, 1.0, 77), [32]),
        ("3D Laplacian 128^3, FP64", lambda: be.laplace3d(128, 128, 128), [1, 4, 8, 16]),
        ("3D Laplacian 64^3, FP64", lambda: be.laplace3d(64, 64, 64), [4]),
    ]
    iflag = ctypes.c_int(0)
    L.phist_amd_sparseMat_set_format(0, 0, ctypes.byref(iflag))
    for name, make, widths in cases:
        t0 = time.perf_counter()
        A = make()
        t_build = time.perf_counter() - t0
        n = len(widths)
        nv, dms, tms, st = np.array(widths, np.int32), np.zeros(n), np.zeros(n), np.zeros(7 * n, np.int32)
        t0 = time.perf_counter()
        L.phist_Damd_sparseMat_tune(A.h, n, pp(nv), 1 | 2, pp(dms), pp(tms), pp(st), ctypes.byref(iflag))
        if iflag.value:
            raise RuntimeError(f"tune {name}: iflag {iflag.value}")
        t_tune = time.perf_counter() - t0
        rows = []
        for i, w in enumerate(widths):
            C, sg, lb, bpc, depth, cblk, stg = (int(v) for v in st[7 * i:7 * i + 7])
            gb = spmv_bytes(A.nnz, A.nrows, w) / 1e9
            # the product with x staged in LDS (the tuned launch otherwise; best of the workgroups per CU),
            # or None where no tile window of this width fits
            staged = staged_product_ms(be, A, w, lb, depth, cblk) if not stg else None
            rows.append({"nvec": w, "untuned_ms": round(float(dms[i]), 4), "tuned_ms": round(float(tms[i]), 4),
                         "gain": round(float(dms[i] / tms[i]), 3), "C": C, "sigma": sg, "lane_bytes": lb,
                         "blocks_per_cu": bpc, "depth": depth, "col_block": cblk, "stage": stg,
                         "staged_ms": None if staged is None else round(staged, 4),
                         "tuned_GBs": round(gb / (tms[i] * 1e-3), 1)})
        out = {"metric": "SpMV autotuner: kernel ms per product, untuned vs tuned", "workload": name,
               "rows": A.nrows, "nnz": A.nnz, "build_s": round(t_build, 2), "tune_s": round(t_tune, 2),
               "widths": rows}
        print(json.dumps(out), flush=True)
        A.delete()


def staged_product_ms(be, A, nvec, lane_bytes, depth, col_block, reps=20):
    """kernel ms per product y = A x with x staged in LDS per workgroup tile (phist_amd_set_spmv_stage),
    lane bytes, depth and column block as given, the fastest over 1-8 workgroups per CU; None if the
    launch does not stage (no tile window of this width fits)"""
    L, fl = be.L, ctypes.c_int(0)
    m = be.map(A.nrows)
    x, y = be.mvec(m, nvec), be.mvec(m, nvec)
    x.random()
    best = None
    try:
        L.phist_amd_set_spmv_stage(1, ctypes.byref(fl))
        L.phist_amd_set_spmv_lane_bytes(lane_bytes, ctypes.byref(fl))
        L.phist_amd_set_spmv_depth(depth, ctypes.byref(fl))
        L.phist_amd_set_spmv_col_block(col_block, ctypes.byref(fl))
        for bpc in (1, 2, 3, 4, 6, 8):
            L.phist_amd_set_spmv_blocks_per_cu(bpc, ctypes.byref(fl))
            c0, c1 = ctypes.c_int64(0), ctypes.c_int64(0)
            L.phist_amd_spmv_staged_launches(ctypes.byref(c0), ctypes.byref(fl))
            be.spmv(1.0, A, x, 0.0, y)
            L.phist_amd_spmv_staged_launches(ctypes.byref(c1), ctypes.byref(fl))
            if c1.value == c0.value:
                return None
            be.synchronize()
            L.phist_amd_kernel_timer(1, ctypes.byref(fl))
            for _ in range(reps):
                be.spmv(1.0, A, x, 0.0, y)
            be.synchronize()
            kms, kc = ctypes.c_double(0.0), ctypes.c_int64(0)
            L.phist_amd_kernel_timer_read(ctypes.byref(kms), ctypes.byref(kc), ctypes.byref(fl))
            L.phist_amd_kernel_timer(0, ctypes.byref(fl))
            ms = kms.value / reps  # per product (a column-blocked product is several launches)
            best = ms if best is None else min(best, ms)
    finally:
        L.phist_amd_set_spmv_stage(-1, ctypes.byref(fl))
        L.phist_amd_set_spmv_blocks_per_cu(-1, ctypes.byref(fl))
        L.phist_amd_set_spmv_lane_bytes(0, ctypes.byref(fl))
        L.phist_amd_set_spmv_depth(0, ctypes.byref(fl))
        L.phist_amd_set_spmv_col_block(-1, ctypes.byref(fl))
        x.delete()
        y.delete()
        m.delete()
    return best


def unit_random_block(be, m, k):
    """k columns of KISS random numbers (phist_Dmvec_random), each scaled to unit norm on the device"""
    B = be.mvec(m, k)
    B.random()
    inv = 1.0 / be.norm2(B)
    be.call("mvec_vscale", B.h, inv.ctypes.data_as(ctypes.c_void_p))
    return B


def true_relres_dev(be, A, shifts, B, X):
    """||b_j - (A + s_j I) x_j|| / ||b_j|| computed on the GPU in FP64"""
    R = be.mvec(A.domain_map, B.nvec)
    be.call("mvec_add_mvec", be.s(1.0), B.h, be.s(0.0), R.h)
    be.spmv_shifted(-1.0, A, shifts, X, 1.0, R)
    out = be.norm2(R) / be.norm2(B)
    R.delete()
    return out


def minres_bench(be, args, W=2000, L=4000, nsys=32, tol=1e-6, max_iter=3000):
    """Batched shifted MINRES: honeycomb sheet of W x L = 8,000,000 sites (nearest neighbours -1, on-site
    U(-1/2, 1/2)), 32 shifts uniform in [-2.5, 2.5], 32 random right-hand sides of unit norm in one block,
    tol 1e-6 relative, FP64 and FP32 iterations.  The shifts lie inside the spectrum: the residual of an
    interior shifted system decays like k^(-1/2) until the Krylov space resolves the eigenvalues next to
    the shift, which takes on the order of N iterations, so on 8M rows the tolerance is out of reach of any
    unpreconditioned Krylov method (the reference's as well).  Reported for the full sheet: the time per
    iteration and the true residuals after --max-iter iterations; and time-to-tolerance of the same
    workload on a 64 x 64 sheet (4096 rows, 32 systems through 16 slots, converged columns replaced in
    place), where 1e-6 is reached in FP64 and by the FP32 iterations refined in FP64."""
    results = {"full": {}, "small": {}}
    # ---- the full sheet: throughput and residual after max_iter iterations ----
    t0 = time.perf_counter()
    A = be.honeycomb(W, L, 0, 0, 0, 1.0, 77)
    N = A.nrows
    print(f"bench: honeycomb {W}x{L} ({N} rows, {A.nnz} nnz) built in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    shifts = np.random.default_rng(11).uniform(-2.5, 2.5, nsys)
    # the library's autotuner picks the SELL layout and launch settings of the
    # 32-column product for the matrix and its FP32 copy (set-up, not timed)
    pp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    nv, dms, tms, st = np.array([nsys], np.int32), np.zeros(1), np.zeros(1), np.zeros(7, np.int32)
    iflag = ctypes.c_int(0)
    t0 = time.perf_counter()
    be.L.phist_Damd_sparseMat_tune(A.h, 1, pp(nv), 1 | 4, pp(dms), pp(tms), pp(st), ctypes.byref(iflag))
    if iflag.value:
        raise RuntimeError(f"phist_Damd_sparseMat_tune iflag={iflag.value}")
    tune = {"C": int(st[0]), "sigma": int(st[1]), "lane_bytes": int(st[2]), "blocks_per_cu": int(st[3]),
            "depth": int(st[4]), "col_block": int(st[5]), "stage": int(st[6]), "untuned_spmv_ms": round(float(dms[0]), 4), "tuned_spmv_ms": round(float(tms[0]), 4),
            "tune_s": round(time.perf_counter() - t0, 2)}
    print(f"bench: autotuner {tune}", file=sys.stderr, flush=True)
    m = A.domain_map
    B = unit_random_block(be, m, nsys)
    for label, fp32 in (("fp64", False), ("fp32", True)):
        X = be.mvec(m, nsys)
        if fp32:  # cached FP32 copy of A: built outside the timed solve
            b1, x1 = be.mvec(m, 1), be.mvec(m, 1)
            be.minres_shifted(A, [0.0], b1, x1, 1.0, 1, fp32=True)
            b1.delete()
            x1.delete()
        be.synchronize()
        t0 = time.perf_counter()
        it, rr, unconv = be.minres_shifted(A, shifts, B, X, tol, max_iter, fp32=fp32)
        be.synchronize()
        dt = time.perf_counter() - t0
        iters = int(it.max())
        tr = true_relres_dev(be, A, shifts, B, X)
        results["full"][label] = {"seconds": round(dt, 3), "iterations": iters, "unconverged": int(unconv),
                                  "true_relres_max": float(tr.max()), "true_relres_median": float(np.median(tr)),
                                  "ms_per_iteration": round(1e3 * dt / max(1, iters), 4)}
        print(f"bench: minres {label} (full sheet): {dt:.2f} s, {iters} iterations, true relres max "
              f"{tr.max():.3e} median {np.median(tr):.3e}", file=sys.stderr, flush=True)
        X.delete()
    cpu = {}
    if not args.no_cpu:
        cpu["full_sheet"] = minres_cpu_sample(A, B.get(), shifts, tol)
    B.delete()
    A.delete()
    # ---- a sheet where the tolerance is reachable: time to 1e-6 with slot refill ----
    Ws, Ls, slots = 64, 64, 16
    A = be.honeycomb(Ws, Ls, 0, 0, 0, 1.0, 77)
    ms = A.domain_map
    B = unit_random_block(be, ms, nsys)
    for label, fp32 in (("fp64", False), ("fp32", True)):
        X = be.mvec(ms, nsys)
        bv, xv = B.view(0, 1), X.view(0, 1)
        be.minres_refill(A, shifts[:2], bv, xv, tol, 5, 2, fp32=fp32)  # warm-up
        bv.delete()
        xv.delete()
        be.synchronize()
        t0 = time.perf_counter()
        it, rr, unconv = be.minres_refill(A, shifts, B, X, tol, 60000, slots, fp32=fp32)
        be.synchronize()
        dt = time.perf_counter() - t0
        tr = true_relres_dev(be, A, shifts, B, X)
        results["small"][label] = {"seconds": round(dt, 3), "unconverged_estimate": int(unconv),
                                   "true_relres_max": float(tr.max()),
                                   "true_relres_le_tol": int(np.sum(tr <= tol)),
                                   "iterations_each": [int(v) for v in it]}
        print(f"bench: minres {label} ({Ws}x{Ls}, {slots} slots): {dt:.2f} s, iterations max {it.max()} median "
              f"{int(np.median(it))}, true relres max {tr.max():.2e} ({int(np.sum(tr <= tol))}/{nsys} <= tol)",
              file=sys.stderr, flush=True)
        X.delete()
    if not args.no_cpu:
        cpu["small_sheet"] = minres_cpu_solve(A, B.get(), shifts, tol, 60000)
    B.delete()
    A.delete()
    out = {"metric": "batched shifted MINRES (32 shifts x 32 random unit RHS, tol 1e-6)",
           "value": results["full"]["fp64"]["ms_per_iteration"], "unit": "ms/iteration (FP64, 8M rows x 32)",
           "n_gpus": 1, "higher_is_better": False, "dtype": "f64 and f32",
           "data": "synthetic (honeycomb sheet with seeded on-site disorder, KISS random unit right-hand sides)",
           "config": {"workload": f"honeycomb {W}x{L} ({N} rows), nearest neighbours, on-site U(-0.5,0.5), "
                                  f"{nsys} shifts in [-2.5,2.5], random unit RHS, tol {tol}; full sheet run for "
                                  f"{max_iter} iterations; time to tol on {Ws}x{Ls} with {slots} slots"},
           "autotuned": tune, "full_sheet": results["full"], "small_sheet": results["small"],
           "cpu_baseline": cpu or None}
    print(json.dumps(out), flush=True)


def _minres_port(A):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import jada_port
    import minres_port
    ro, ci, va = A.crs()
    op = jada_port.CrsOperator(len(ro) - 1, ro, ci, va)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return minres_port, op, threads


def minres_cpu_sample(A, Bh, shifts, tol, iters=4):
    """FP64 MINRES iterations of the same block on the host cores (oracle/minres_port.py over the
    reference's flang-built OpenMP SpMV, vector passes on row blocks in a thread pool): the bounded
    sample is `iters` iterations, the first untimed; reported per iteration, the unit of the GPU line"""
    minres_port, op, threads = _minres_port(A)
    print(f"bench: timing {iters} CPU MINRES iterations ({op.kind} SpMV, {threads} threads)", file=sys.stderr,
          flush=True)
    marks = []
    minres_port.blocked_minres(op, shifts, np.ascontiguousarray(Bh), tol, iters, threads=threads,
                               on_iteration=lambda k: marks.append(time.perf_counter()))
    per = (marks[-1] - marks[0]) / max(1, len(marks) - 1)
    return {"value": round(1e3 * per, 2), "unit": "ms/iteration (FP64)", "cores": threads, "kind": "port",
            "sample": f"iterations 2..{iters} of the same {Bh.shape[1]} systems on all {Bh.shape[0]} rows (FP64); "
                      f"SpMV: {'the reference builtin dspmvm (flang, OpenMP)' if op.kind == 'reference' else 'oracle C'}",
            **cpu_info_fields(threads)}


def minres_cpu_solve(A, Bh, shifts, tol, max_iter):
    """time to tol of the same systems on the host cores, all in one block (FP64)"""
    minres_port, op, threads = _minres_port(A)
    pool = max(1, min(threads, Bh.shape[0] // 65536))  # small sheets: no row blocks
    t0 = time.perf_counter()
    _x, its, rr = minres_port.blocked_minres(op, shifts, np.ascontiguousarray(Bh), tol, max_iter, threads=pool)
    secs = time.perf_counter() - t0
    return {"value": round(secs, 3), "unit": "s", "cores": threads, "kind": "port",
            "sample": f"the whole solve, {Bh.shape[1]} systems in one block (FP64): iterations max {int(its.max())} "
                      f"median {int(np.median(its))}, {int(np.sum(rr <= tol))}/{Bh.shape[1]} at tol by the recurrence; "
                      f"vector passes on {pool} thread(s)",
            **cpu_info_fields(threads)}


def gmres_bench(be, args, n=128, nsys=16, tol=1e-8, restart=30, max_iter=600):
    """Restarted blocked GMRES on shifted systems: 3D Laplacian 128^3 (2,097,152 rows), 16 shifts in
    [0.05, 1.0], random right-hand sides, zero initial guess, GMRES(30) to tol 1e-8, FP64 and FP32;
    reports the time per Arnoldi step (the Krylov basis of a step holds nsys*(restart+1) vectors)."""
    t0 = time.perf_counter()
    A = be.laplace3d(n, n, n)
    N = A.nrows
    print(f"bench: Laplacian {n}^3 ({N} rows) built in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    shifts = np.linspace(0.05, 1.0, nsys)
    m = A.domain_map
    B = be.mvec(m, nsys)
    B.random()
    results = {}
    for label, fp32 in (("fp64", False), ("fp32", True)):
        X = be.mvec(m, nsys)
        if fp32:  # cached FP32 copy of A: built outside the timed solve
            b1, x1 = be.mvec(m, 1), be.mvec(m, 1)
            be.minres_shifted(A, [0.0], b1, x1, 1.0, 1, fp32=True)
            b1.delete()
            x1.delete()
        be.synchronize()
        t0 = time.perf_counter()
        it, rr, unconv = be.gmres_shifted(A, shifts, B, X, tol, max_iter, restart=restart, fp32=fp32)
        be.synchronize()
        dt = time.perf_counter() - t0
        iters = int(it.max())
        results[label] = {"seconds": round(dt, 3), "max_iters": iters, "median_iters": int(np.median(it)),
                          "unconverged": int(unconv), "max_relres": float(rr.max()),
                          "ms_per_step": round(1e3 * dt / max(1, iters), 4)}
        print(f"bench: gmres {label}: {dt:.2f} s, max {iters} Arnoldi steps, {unconv} unconverged",
              file=sys.stderr, flush=True)
        X.delete()
    cpu = None
    if not args.no_cpu:
        cpu = gmres_cpu_baseline(A, B.get(), shifts, restart)
    out = {"metric": "blocked shifted GMRES time (16 shifts, tol 1e-8)", "value": results["fp64"]["seconds"],
           "unit": "s", "n_gpus": 1, "higher_is_better": False, "dtype": "f64 and f32",
           "data": "synthetic (3D Laplacian, KISS random right-hand sides)",
           "config": {"workload": f"3D Laplacian {n}^3 ({N} rows), {nsys} shifts in [0.05,1.0], GMRES({restart}), "
                                  f"tol {tol}, maxIter {max_iter}, zero initial guess"},
           "fp64": results["fp64"], "fp32": results["fp32"], "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)
    B.delete()
    A.delete()


def gmres_cpu_baseline(A, Bh, shifts, restart):
    """The first GMRES(restart) cycle of the same systems on the host cores (FP64): the blocked
    GMRES of oracle/jada_port.py (one basis per system, classical Gram-Schmidt twice) over the
    reference's flang-built OpenMP SpMV and OpenBLAS; reported per Arnoldi step, the unit of the
    GPU runs' ms_per_step (the bounded sample: one cycle of `restart` steps)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import jada_port
    ro, ci, va = A.crs()
    op = jada_port.CrsOperator(len(ro) - 1, ro, ci, va)
    del ro, ci, va
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    print(f"bench: timing one CPU GMRES({restart}) cycle ({op.kind} SpMV, {threads} threads)", file=sys.stderr,
          flush=True)
    t0 = time.perf_counter()
    _x, its, rr = jada_port.blocked_gmres(op, list(shifts), np.ascontiguousarray(Bh), None,
                                          np.full(Bh.shape[1], 1e-8), restart, restart)
    secs = time.perf_counter() - t0
    steps = int(its.max())
    return {"value": round(1e3 * secs / max(1, steps), 2), "unit": "ms per Arnoldi step", "cores": threads,
            "kind": "port", "seconds": round(secs, 3),
            "sample": f"the first GMRES({restart}) cycle of all {Bh.shape[1]} systems ({steps} Arnoldi steps, FP64, "
                      f"relative residual after it {float(rr.min()):.1e}..{float(rr.max()):.1e}); SpMV: "
                      f"{'the reference builtin dspmvm (flang, OpenMP)' if op.kind == 'reference' else 'oracle C'}",
            **cpu_info_fields(threads)}


def carp_bytes(nnz, n, nvec, s=8):
    """bytes of one forward + backward Kaczmarz sweep: every row is visited twice and reads its index
    (4), row pointer (8), entries (s+4 each) and, per column of X, reads and writes x at each entry"""
    return 2 * (nnz * (s + 4) + n * 12 + 2 * nnz * nvec * s)


def carp_bench(be, args, W=2048, L=2048):
    """CARP sweeps X <- DKSWP(A - sigma I, X) as in examples/kernels/bench_carp.cpp (sigma 42.9, omega 1.5,
    rhs NULL) on a W x L graphene sheet (long-range hopping, on-site disorder 0.2, the reference driver's
    graphene<L> problem), nvec columns.  cpu_baseline: the oracle's sequential sweep on 1/64 of the rows."""
    import oracle_lib as ol
    nvec = args.nvec
    t0 = time.perf_counter()
    A = be.honeycomb(W, L, 0, 0, 1, 0.2, 7)
    N = A.nrows
    print(f"bench: graphene {W}x{L} ({N} rows, {A.nnz} nnz) built in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    m = A.domain_map
    X = be.mvec(m, nvec)
    X.random()
    sig = np.full(nvec, 42.9)
    om = np.full(nvec, 1.5)
    t0 = time.perf_counter()
    W_ = be.carp_setup(A, sig)
    cp, _ = W_.colors()
    ncol = len(cp) - 1
    print(f"bench: CARP setup (distance-2 colouring: {ncol} colours) {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    iflag = ctypes.c_int(0)
    for _ in range(args.warmup):
        W_.sweep(sig, None, X, om)
    be.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        W_.sweep(sig, None, X, om)
    be.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    be.L.phist_amd_kernel_timer(1, ctypes.byref(iflag))
    for _ in range(3):
        W_.sweep(sig, None, X, om)
    kms, kcount = ctypes.c_double(0.0), ctypes.c_int64(0)
    be.L.phist_amd_kernel_timer_read(ctypes.byref(kms), ctypes.byref(kcount), ctypes.byref(iflag))
    be.L.phist_amd_kernel_timer(0, ctypes.byref(iflag))
    dev_ms = kms.value / max(1, kcount.value)
    nbytes = carp_bytes(A.nnz, N, nvec)
    cpu = None
    if not args.no_cpu:
        print("bench: timing the CPU baseline", file=sys.stderr, flush=True)
        ro, ci, va = A.crs()
        ns = N // 64
        nnz_s = int(ro[ns] - 1)
        ro_s, ci_s, va_s = ro[:ns + 1].copy(), ci[:nnz_s].copy(), va[:nnz_s].copy()
        del ro, ci, va
        keep = ci_s <= ns  # the sample is the leading ns x ns block
        cnt = np.zeros(ns + 1, np.int64)
        rows_of = np.repeat(np.arange(ns), np.diff(ro_s))
        np.add.at(cnt, rows_of[keep] + 1, 1)
        ro_s = np.cumsum(cnt) + 1
        ci_s, va_s = ci_s[keep], va_s[keep]
        xs = np.random.default_rng(3).uniform(-1, 1, (ns, nvec))
        seq = np.concatenate([np.arange(ns), np.arange(ns)[::-1]]).astype(np.int32)
        times = []
        tb = time.perf_counter()
        while time.perf_counter() - tb < 10.0 or len(times) < 3:
            t1 = time.perf_counter()
            ol.kacz_rows(ro_s, ci_s, va_s, sig, om, xs, seq)
            times.append(time.perf_counter() - t1)
            if len(times) >= 100:
                break
        tc = float(np.median(times))
        cpu = {"value": round(carp_bytes(len(ci_s), ns, nvec) / tc / 1e9, 3), "unit": "GB/s", "cores": 1,
               "kind": "port", "ms_per_step": round(tc * 1e3, 3),
               "sample": f"leading {ns} x {ns} block ({len(ci_s)} nnz), lexicographic forward + backward "
                         f"sweep, nvec={nvec}, median of {len(times)} runs"}
    out = {"metric": f"CARP sweep GB/s (graphene {W}x{L}, nvec={nvec}, FP64)",
           "value": round(nbytes / (ms * 1e-3) / 1e9, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (graphene sheet with seeded on-site disorder, KISS random X)",
           "config": {"workload": f"carp_sweep (forward + backward, {ncol} colours), graphene {W}x{L} "
                                  f"long-range, disorder 0.2, sigma 42.9, omega 1.5, rhs NULL",
                      "rows": N, "nnz": A.nnz, "nvec": nvec, "colours": ncol, "bytes_per_sweep": nbytes,
                      "device_ms_per_sweep": round(dev_ms, 4)},
           "roofline": {"bound": "hbm", "achieved": round(nbytes / (dev_ms * 1e-3) / 1e9, 2), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(nbytes / (dev_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                        "traffic": read_traffic(os.path.join(ROOT, "profiles",
                                                             f"carp_sweep_graphene{W}_nvec{nvec}_traffic.txt"),
                                                "traffic_bytes_per_sweep")},
           "cpu_baseline": cpu}
    W_.delete()
    X.delete()
    A.delete()
    print(json.dumps(out), flush=True)


def serve_bench(be, args):
    """Served workload: Poisson arrivals (rate --rate per second for --duration seconds) of
    shifted solves (1-16 right-hand sides, tol 1e-6, FP32 iterations, shifts keeping A + sI
    positive definite) and eigenvalue requests (4-8 smallest pairs, tol 1e-8, FP32 correction
    equations), 85 % / 15 %, over 3D Laplacians 32^3..128^3 and XXZ chains L=16..22, through the
    library's continuous-batching loop (32 slots per MINRES batch, 200 GB work-space budget).
    Latency = submission to completion; all buffers are allocated before the timed run."""
    rng = np.random.default_rng(2024)
    mats = []
    t0 = time.perf_counter()
    for n in (32, 64, 96, 128):
        mats.append(("lap%d" % n, be.laplace3d(n, n, n), 0.0))  # spectrum (0, 12)
    for L in (16, 18, 20, 22):
        mats.append(("xxz%d" % L, be.spinchain_sz(L, L // 2), 1.8 * L))  # spectrum above about -1.78 L
    print(f"bench: {len(mats)} matrices built in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    srv = be.server(budget_bytes=200 * 2**30, slots=32)
    # pre-allocated buffers: per matrix a ring of right-hand-side / solution blocks
    ring = 24
    pools = []
    for name, A, lo in mats:
        Bh = rng.uniform(-1, 1, (A.nrows, 16))
        pools.append([[be.mvec(A.domain_map, 16, Bh), be.mvec(A.domain_map, 16), False] for _ in range(ring)])
        print(f"bench: buffers for {name} ({A.nrows} rows) ready", file=sys.stderr, flush=True)
    for name, A, lo in mats:  # FP32 copies of the matrices, made before the timed run
        b1, x1 = be.mvec(A.domain_map, 1), be.mvec(A.domain_map, 1)
        be.minres_shifted(A, [0.0], b1, x1, 1.0, 1, fp32=True)
        b1.delete()
        x1.delete()
    print("bench: FP32 matrix copies ready; serving", file=sys.stderr, flush=True)
    nreq = rng.poisson(args.rate * args.duration)
    arrivals = np.sort(rng.uniform(0.0, args.duration, nreq))
    kinds = rng.uniform(size=nreq) < 0.15
    which_mat = rng.integers(0, len(mats), nreq)
    be.synchronize()
    live = {}  # id -> (kind, mat index, pool slot or JD objects, t_arrival)
    lat = {"solve": [], "jada": []}
    systems = 0
    unconv = 0
    waiting = []  # arrivals whose client buffers are all busy
    T0 = time.perf_counter()
    i = 0
    last_print = T0
    while i < nreq or live or waiting:
        now = time.perf_counter() - T0
        while i < nreq and arrivals[i] <= now:
            waiting.append(i)
            i += 1
        still = []
        for r in waiting:
            mi = int(which_mat[r])
            name, A, lo = mats[mi]
            if kinds[r]:
                if A.nrows > 300000:  # eigenvalue requests on the smaller matrices
                    mi = int(rng.integers(0, 2)) * 4
                    name, A, lo = mats[mi]
                k = int(rng.integers(4, 9))
                rid, Q, R, ev, res = srv.submit_jada(A, k, "SR", 1e-8, block_size=4, min_bas=k + 12,
                                                     max_bas=k + 32, max_iters=400, inner_max_iters=10, fp32=True)
                live[rid] = ("jada", mi, (Q, R), arrivals[r])
                continue
            slot = next((j for j, p in enumerate(pools[mi]) if not p[2]), None)
            if slot is None:
                still.append(r)
                continue
            nrhs = int(rng.integers(1, 17))
            shifts = lo + rng.uniform(0.5, 3.0, nrhs)
            B, X, _ = pools[mi][slot]
            pools[mi][slot][2] = True
            rid = srv.submit_solve(A, shifts, B, X, 1e-6, 5000, fp32=True)
            live[rid] = ("solve", mi, slot, arrivals[r], nrhs)
        waiting = still
        ts = time.perf_counter()
        srv.step()
        if time.perf_counter() - ts > 2.0:
            print(f"bench: serve step took {time.perf_counter() - ts:.1f} s (live {len(live)})", file=sys.stderr,
                  flush=True)
        done = []
        for rid, info in live.items():
            st, latency, it = srv.status(rid)
            if st >= 2 or st < 0:
                done.append(rid)
                t_complete = time.perf_counter() - T0
                lat[info[0]].append(t_complete - info[3])  # from arrival (includes client-side wait)
                if info[0] == "solve":
                    systems += info[4]
                    pools[info[1]][info[2]][2] = False
                else:
                    info[2][0].delete()
                    info[2][1].delete()
                if st != 2:
                    unconv += 1
        for rid in done:
            srv.release(rid)
            del live[rid]
        if time.perf_counter() - last_print > 2.0:
            last_print = time.perf_counter()
            print(f"bench: serve t={now:.1f}s submitted {i}/{nreq} live {len(live)} waiting {len(waiting)}",
                  file=sys.stderr, flush=True)
    wall = time.perf_counter() - T0

    def pct(a, q):
        return round(float(np.percentile(a, q)) * 1e3, 2) if len(a) else None

    alll = lat["solve"] + lat["jada"]
    out = {"metric": "served p50/p99 latency and solves/s", "value": pct(alll, 99), "unit": "ms (p99)",
           "n_gpus": 1, "higher_is_better": False, "dtype": "f32 solves / f64 outer JD",
           "data": "synthetic (generated matrices, random right-hand sides, Poisson arrivals)",
           "config": {"workload": "Poisson arrivals of shifted solves (1-16 rhs, tol 1e-6) and JD requests "
                                  "(4-8 pairs, tol 1e-8) on 3D Laplacians 32^3..128^3 and XXZ L=16..22",
                      "rate_per_s": args.rate, "duration_s": args.duration, "requests": nreq,
                      "slots": 32, "budget_bytes": 200 * 2**30},
           "p50_ms": pct(alll, 50), "p99_ms": pct(alll, 99),
           "solve": {"count": len(lat["solve"]), "p50_ms": pct(lat["solve"], 50), "p99_ms": pct(lat["solve"], 99)},
           "jada": {"count": len(lat["jada"]), "p50_ms": pct(lat["jada"], 50), "p99_ms": pct(lat["jada"], 99)},
           "solves_per_s": round(systems / wall, 2), "wall_s": round(wall, 2), "unconverged": unconv}
    print(json.dumps(out), flush=True)
    for pool in pools:
        for B, X, _ in pool:
            B.delete()
            X.delete()
    srv.delete()
    for _, A, _ in mats:
        A.delete()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nvec", type=int, default=8)
    ap.add_argument("--rows", type=int, default=4194304)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--op", choices=["spmv", "gram", "cholqr", "mvsd", "jada", "minres", "serve", "carp", "gmres",
                                     "tune", "dot", "norm2", "axpby"],
                    default="spmv",
                    help="spmv (default): the fused SpMV line; gram / cholqr / mvsd (fused_mvsd_mvTmv, "
                         "W = V*C and W'W, beta=0): the tall-skinny configuration (16,777,216 x k, U(-1,1)) "
                         "reported in the same format")
    ap.add_argument("--k", type=int, default=8, help="block width for --op gram/cholqr/mvsd/dot/norm2/axpby")
    ap.add_argument("--rate", type=float, default=40.0, help="--op serve: mean request arrivals per second")
    ap.add_argument("--duration", type=float, default=20.0, help="--op serve: seconds of arrivals")
    ap.add_argument("--starts", type=int, default=5, help="--op jada: start vectors averaged per precision")
    ap.add_argument("--cpu-jada", action="store_true",
                    help="--op jada: time the CPU port also for the spin chain (default: Laplacians only)")
    ap.add_argument("--max-iter", type=int, default=None, help="--op minres: iteration limit (default 3000)")
    ap.add_argument("--precision", choices=["both", "fp64", "mixed"], default="both",
                    help="--op jada: run the FP64 solve, the mixed-precision one, or both")
    ap.add_argument("--inner", choices=["GMRES", "MINRES"], default="GMRES",
                    help="--op jada: correction-equation solver (GMRES is the reference's default)")
    ap.add_argument("--problem", default="xxz24",
                    help="--op jada: xxz<L> (S_z=0 spin chain, 16 smallest) or lap<n> (n^3 Laplacian, 8 smallest)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run the library's STREAM copy kernel (1 GiB in, 1 GiB out per launch) so a "
                         "counter profile of this command holds a known-size reference dispatch")
    ap.add_argument("--gather-calibration", type=int, default=0, metavar="BYTES",
                    help="also run the library's known-size gather kernel (every BYTES-byte segment of 1 GiB "
                         "read once, scrambled order) for a counter profile of gathers of that size")
    ap.add_argument("--halfband", type=int, default=50000,
                    help="--op spmv: columns drawn within +-HALFBAND of the diagonal (50000 is the kernel-bench "
                         "matrix; other values are a band sweep, not the headline line)")
    ap.add_argument("--lane-bytes", type=int, default=0,
                    help="--op spmv: bytes of x each SpMV lane gathers per nonzero (0 = library default)")
    ap.add_argument("--blocks-per-cu", type=int, default=-1,
                    help="--op spmv: SpMV workgroups per CU (-1 = library table)")
    ap.add_argument("--depth", type=int, default=0, choices=[0, 8, 16, 32],
                    help="--op spmv: entries per row whose loads the SpMV issues together (0: tuned / default)")
    ap.add_argument("--col-block", type=int, default=-1,
                    help="--op spmv: block columns per launch (0 = one launch, -1 = as tuned)")
    ap.add_argument("--no-tune", action="store_true",
                    help="--op spmv / jada: skip the autotuner (launch settings from the library's table)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    import phist_amd
    be = phist_amd.Backend(np.float64)
    be.init()
    L = be.L
    iflag = ctypes.c_int(0)
    if args.copy_calibration:
        mean_bw, max_bw = ctypes.c_double(), ctypes.c_double()
        L.phist_bench_stream_copy(ctypes.byref(mean_bw), ctypes.byref(max_bw), ctypes.byref(iflag))
        print(f"bench: STREAM copy {mean_bw.value:.0f} GB/s mean, {max_bw.value:.0f} GB/s max",
              file=sys.stderr, flush=True)
        for kind, label in ((0, "read"), (1, "copy")):
            L.phist_amd_bench_hbm(kind, ctypes.byref(mean_bw), ctypes.byref(max_bw), ctypes.byref(iflag))
            print(f"bench: HBM ceiling {label} (16-byte accesses, four in flight) {mean_bw.value:.0f} GB/s mean, "
                  f"{max_bw.value:.0f} GB/s max", file=sys.stderr, flush=True)
    if args.gather_calibration:
        mean_bw, max_bw = ctypes.c_double(), ctypes.c_double()
        L.phist_amd_bench_gather(args.gather_calibration, ctypes.byref(mean_bw), ctypes.byref(max_bw),
                                 ctypes.byref(iflag))
        if iflag.value:
            raise RuntimeError(f"phist_amd_bench_gather iflag={iflag.value}")
        print(f"bench: gather {args.gather_calibration}-byte segments once over 1 GiB: {mean_bw.value:.0f} GB/s "
              f"mean, {max_bw.value:.0f} GB/s max", file=sys.stderr, flush=True)
    if args.op == "jada":
        return jada_bench(be, args)
    if args.op == "minres":
        return minres_bench(be, args, max_iter=args.max_iter or 3000)
    if args.op == "serve":
        return serve_bench(be, args)
    if args.op == "carp":
        return carp_bench(be, args)
    if args.op == "gmres":
        return gmres_bench(be, args)
    if args.op == "tune":
        return tune_bench(be, args)
    if args.op in ("dot", "norm2", "axpby"):
        return blas1_bench(be, args)
    if args.op != "spmv":
        return tall_skinny(be, args)
    L.phist_amd_sparseMat_set_format(32, 256, ctypes.byref(iflag))
    L.phist_amd_set_spmv_lane_bytes(args.lane_bytes, ctypes.byref(iflag))
    L.phist_amd_set_spmv_blocks_per_cu(args.blocks_per_cu, ctypes.byref(iflag))
    if not iflag.value:
        L.phist_amd_set_spmv_depth(args.depth, ctypes.byref(iflag))
    if not iflag.value:
        L.phist_amd_set_spmv_col_block(args.col_block, ctypes.byref(iflag))
    if iflag.value:
        raise RuntimeError(f"SpMV launch settings rejected (iflag={iflag.value})")
    N, nvec = args.rows, args.nvec
    t0 = time.perf_counter()
    A = be.banded_random(N, 32, args.halfband, 40.0, 2024)
    t_build = time.perf_counter() - t0
    m = be.map(N)
    x = be.mvec(m, nvec)
    x.random()
    y = be.mvec(m, nvec)
    nbytes = spmv_bytes(A.nnz, N, nvec)
    print(f"bench: matrix {N} rows, {A.nnz} nnz built in {t_build:.1f} s; nvec={nvec}", file=sys.stderr, flush=True)
    tune = None
    if not args.no_tune and args.lane_bytes == 0 and args.blocks_per_cu < 0 and args.depth == 0 and args.col_block < 0:
        # the library's autotuner picks the launch settings for this width; the
        # layout stays the configuration's SELL-32-256
        nv, dms, tms, st = np.array([nvec], np.int32), np.zeros(1), np.zeros(1), np.zeros(7, np.int32)
        pp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        t0 = time.perf_counter()
        L.phist_Damd_sparseMat_tune(A.h, 1, pp(nv), 0, pp(dms), pp(tms), pp(st), ctypes.byref(iflag))
        if iflag.value:
            raise RuntimeError(f"phist_Damd_sparseMat_tune iflag={iflag.value}")
        tune = {"lane_bytes": int(st[2]), "blocks_per_cu": int(st[3]), "depth": int(st[4]), "col_block": int(st[5]), "stage": int(st[6]),
                "untuned_kernel_ms": round(float(dms[0]), 4),
                "tuned_kernel_ms": round(float(tms[0]), 4), "tune_s": round(time.perf_counter() - t0, 2)}
        print(f"bench: autotuner {tune}", file=sys.stderr, flush=True)

    ww = np.zeros(nvec)
    vw = np.zeros(nvec)
    pw, pv = ww.ctypes.data_as(ctypes.c_void_p), vw.ctypes.data_as(ctypes.c_void_p)
    f = be.fn("fused_spmv_mvdot")
    one, zero = ctypes.c_double(1.0), ctypes.c_double(0.0)

    def step():
        fl = ctypes.c_int(0)
        f(one, A.h, x.h, zero, y.h, pw, pv, ctypes.byref(fl))
        if fl.value:
            raise RuntimeError(f"fused_spmv_mvdot iflag={fl.value}")

    for _ in range(args.warmup):
        step()
    be.synchronize()
    L.phist_amd_kernel_timer(1, ctypes.byref(iflag))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    be.synchronize()
    t1 = time.perf_counter()
    kms = ctypes.c_double(0.0)
    kcount = ctypes.c_int64(0)
    L.phist_amd_kernel_timer_read(ctypes.byref(kms), ctypes.byref(kcount), ctypes.byref(iflag))
    L.phist_amd_kernel_timer(0, ctypes.byref(iflag))
    ms_step = (t1 - t0) * 1e3 / args.steps
    # kernel time per product: one launch, or one per column block when the
    # tuner split the block columns (then kernel_launches_timed is a multiple of steps)
    k_avg_ms = kms.value / max(1, args.steps)
    value = nbytes / (ms_step * 1e-3) / 1e9
    achieved = nbytes / (k_avg_ms * 1e-3) / 1e9
    default_run = args.halfband == 50000 and N == 4194304 and args.lane_bytes == 0 and args.blocks_per_cu < 0 \
        and args.depth == 0 and args.col_block < 0 and not args.no_tune
    traffic = read_traffic(os.path.join(ROOT, "profiles", f"sell_spmv_fused_nvec{nvec}_traffic.txt")) \
        if default_run else None

    cpu = None
    if not args.no_cpu:
        print("bench: timing the CPU baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(A, N // 4, nvec)

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (banded random matrix from a seeded generator, KISS random block vector)",
        "config": {
            "workload": f"fused_spmv_mvdot, banded random {N} rows x 32 nnz/row (+-{args.halfband} band, diag 40), "
                        f"SELL-C-sigma C=32 sigma=256, FP64, nvec={nvec}, alpha=1 beta=0, both dots",
            "rows": N, "nnz": A.nnz, "nvec": nvec, "bytes_per_step": nbytes,
            "pct_hbm_peak": round(100.0 * value / HBM_PEAK_GBS, 1),
            "kernel_avg_ms": round(k_avg_ms, 4), "kernel_launches_timed": kcount.value,
            "parallelism": "single GPU",
            **({} if args.lane_bytes == 0 else {"lane_bytes": args.lane_bytes}),
            **({} if args.blocks_per_cu < 0 else {"blocks_per_cu": args.blocks_per_cu}),
            **({} if args.depth == 0 else {"depth": args.depth}),
            **({} if args.col_block < 0 else {"col_block": args.col_block}),
            **({} if tune is None else {"autotuned": tune}),
        },
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
        },
        "cpu_baseline": cpu,
    }
    for o in (x, y, A):
        o.delete()
    m.delete()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
