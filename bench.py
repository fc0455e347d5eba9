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
absent, the oracle's single-threaded C restatement.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
METRIC = "fused block-SpMV GB/s (% HBM peak); JD time-to-1e-8 mixed vs FP64; p99"


def spmv_bytes(nnz, n, nvec, sval=8, sx=8, beta_nonzero=False):
    return nnz * (sval + 4) + 4 * n + n * nvec * sx * (2 + (1 if beta_nonzero else 0))


def read_traffic(path):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary, or None.
    The file holds lines `traffic_bytes_per_launch: <float>`."""
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


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
    return {"value": round(gbs, 2), "unit": "GB/s", "cores": threads, "kind": kind, "sample": sample,
            "ms_per_step": round(t * 1e3, 3)}


def tall_skinny(be, args):
    """Gram (self, V'V), CholQR or fused_mvsd_mvTmv on N = 16,777,216 rows x k columns"""
    L = be.L
    iflag = ctypes.c_int(0)
    N, k = 16777216, args.k
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
            be.call("mvec_add_mvec", be.s(1.0), w0.h, be.s(0.0), v.h)  # fresh input (not a timed kernel)
            fl = ctypes.c_int(0)
            f(v.h, C.h, ctypes.byref(fl))
            if fl.value != 0:
                raise RuntimeError(f"chol_QR iflag={fl.value}")
    for _ in range(args.warmup):
        step()
    be.synchronize()
    L.phist_amd_kernel_timer(1, ctypes.byref(iflag))
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
                        "unit": "GB/s", "frac": round(nbytes / (k_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}
    print(json.dumps(out), flush=True)


def jada_bench(be, args):
    """Time to 1e-8 of the symmetric block Jacobi-Davidson, FP64 throughout vs
    FP32 correction equations (outer iteration FP64), same start vector."""
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
    results = {}
    for label, fp32 in (("fp64", False), ("mixed", True)):
        iflag = ctypes.c_int(0)
        be.L.phist_amd_kiss_reset(ctypes.byref(iflag))  # identical random start vector for both runs
        if fp32:  # the FP32 copy of A is made once per matrix: build it outside the timed run
            b, x = be.mvec(A.domain_map, 1), be.mvec(A.domain_map, 1)
            be.minres_shifted(A, [0.0], b, x, 1.0, 1, fp32=True)
            b.delete()
            x.delete()
        be.synchronize()
        t0 = time.perf_counter()
        ev, res, nconv, niter, Q, R = be.subspacejada(A, which="SR", tol=1e-8, max_iters=500, inner_fp32=fp32,
                                                      **cfg)
        be.synchronize()
        dt = time.perf_counter() - t0
        results[label] = {"seconds": round(dt, 3), "outer_iters": niter, "converged": nconv,
                          "lowest": round(float(np.min(ev[:nconv])) if nconv else float("nan"), 10),
                          "max_res": float(np.max(res[:cfg["num_eigs"]]))}
        print(f"bench: jada {label}: {dt:.2f} s, {niter} iterations, {nconv} converged", file=sys.stderr, flush=True)
        Q.delete()
        R.delete()
    speed = results["fp64"]["seconds"] / results["mixed"]["seconds"]
    out = {"metric": "JD time-to-1e-8 mixed vs FP64", "value": results["mixed"]["seconds"], "unit": "s",
           "n_gpus": 1, "higher_is_better": False, "dtype": "f64 outer / f32 inner",
           "data": "synthetic (generated matrix, KISS random start vector)",
           "config": {"workload": name, "which": "SR", "tol": 1e-8, **cfg},
           "fp64": results["fp64"], "mixed": results["mixed"], "speedup_mixed_vs_fp64": round(speed, 3)}
    print(json.dumps(out), flush=True)
    A.delete()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nvec", type=int, default=8)
    ap.add_argument("--rows", type=int, default=4194304)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--op", choices=["spmv", "gram", "cholqr", "mvsd", "jada"], default="spmv",
                    help="spmv (default): the fused SpMV line; gram / cholqr / mvsd (fused_mvsd_mvTmv, "
                         "W = V*C and W'W, beta=0): the tall-skinny configuration (16,777,216 x k, U(-1,1)) "
                         "reported in the same format")
    ap.add_argument("--k", type=int, default=8, help="block width for --op gram/cholqr")
    ap.add_argument("--problem", default="xxz24",
                    help="--op jada: xxz<L> (S_z=0 spin chain, 16 smallest) or lap<n> (n^3 Laplacian, 8 smallest)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run the library's STREAM copy kernel (1 GiB in, 1 GiB out per launch) so a "
                         "counter profile of this command holds a known-size reference dispatch")
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
    if args.op == "jada":
        return jada_bench(be, args)
    if args.op != "spmv":
        return tall_skinny(be, args)
    L.phist_amd_sparseMat_set_format(32, 256, ctypes.byref(iflag))
    N, nvec = args.rows, args.nvec
    if args.copy_calibration:
        mean_bw, max_bw = ctypes.c_double(), ctypes.c_double()
        L.phist_bench_stream_copy(ctypes.byref(mean_bw), ctypes.byref(max_bw), ctypes.byref(iflag))
        print(f"bench: STREAM copy {mean_bw.value:.0f} GB/s mean, {max_bw.value:.0f} GB/s max",
              file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    A = be.banded_random(N, 32, 50000, 40.0, 2024)
    t_build = time.perf_counter() - t0
    m = be.map(N)
    x = be.mvec(m, nvec)
    x.random()
    y = be.mvec(m, nvec)
    nbytes = spmv_bytes(A.nnz, N, nvec)
    print(f"bench: matrix {N} rows, {A.nnz} nnz built in {t_build:.1f} s; nvec={nvec}", file=sys.stderr, flush=True)

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
    k_avg_ms = kms.value / max(1, kcount.value)
    value = nbytes / (ms_step * 1e-3) / 1e9
    achieved = nbytes / (k_avg_ms * 1e-3) / 1e9
    traffic = read_traffic(os.path.join(ROOT, "profiles", f"sell_spmv_fused_nvec{nvec}_traffic.txt"))

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
            "workload": "fused_spmv_mvdot, banded random 4M rows x 32 nnz/row (+-50k band, diag 40), "
                        "SELL-C-sigma C=32 sigma=256, FP64, alpha=1 beta=0, both dots",
            "rows": N, "nnz": A.nnz, "nvec": nvec, "bytes_per_step": nbytes,
            "pct_hbm_peak": round(100.0 * value / HBM_PEAK_GBS, 1),
            "kernel_avg_ms": round(k_avg_ms, 4), "kernel_launches_timed": kcount.value,
            "parallelism": "single GPU",
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
