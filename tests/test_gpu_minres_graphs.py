"""GPU: the device-resident MINRES iteration replayed from captured HIP graphs
gives bit for bit what launching its kernels one by one gives (the same
kernels in the same order with the same arguments), over many replays with
the host's own copies, uploads and synchronisations between them: blocked
MINRES, its iteration limit with several iterations queued per step, the
eigensolver's correction equations (projection fused into the Lanczos
update; FP64 and FP32 inner) and the service loop refilling slots."""
import ctypes

import numpy as np
import pytest

import phist_amd as pa

pytestmark = pytest.mark.gpu


@pytest.fixture
def graphs(be64):
    """set_minres_graphs(enable, iters_per_call, graph_len); the defaults come back afterwards"""
    yield be64.set_minres_graphs
    be64.set_minres_graphs(False, 0, 4)


def minres(be, A, shifts, Bh, tol, maxit, fp32):
    m = be.map(A.nrows)
    B, X = be.mvec(m, Bh.shape[1], Bh), be.mvec(m, Bh.shape[1])
    it, rr, unc = be.minres_shifted(A, shifts, B, X, tol, maxit, fp32=fp32)
    Xh = X.get()
    for o in (B, X, m):
        o.delete()
    return it.copy(), rr.copy(), unc, Xh


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("enable,ipc,glen", [(False, 4, 1), (True, 1, 1), (True, 4, 1), (True, 7, 4), (True, 8, 4)])
def test_minres_graphs_bitexact(be64, graphs, fp32, enable, ipc, glen):
    A = be64.laplace3d(16, 15, 14)
    rng = np.random.default_rng(11)
    shifts = np.array([0.3, -0.4, 1.1, -2.2, 0.05, 3.0, -0.9])  # several indefinite
    Bh = rng.uniform(-1, 1, (A.nrows, len(shifts)))
    tol = 1e-5 if fp32 else 1e-10
    graphs(False, 1)
    ref = minres(be64, A, shifts, Bh, tol, 800, fp32)
    graphs(enable, ipc, glen)
    c0, n0 = be64.minres_graph_stats()
    for rep in range(2):  # a fresh batch each solve: captured again, replayed hundreds of times
        got = minres(be64, A, shifts, Bh, tol, 800, fp32)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]) and got[2] == ref[2]
        assert np.array_equal(got[3], ref[3])
    c1, n1 = be64.minres_graph_stats()
    if enable:
        want = 2 * (int(ref[0].max()) // ipc) * (ipc // glen)  # full graphs among the queued runs
        assert c1 - c0 >= 2 and n1 - n0 >= want, (c1 - c0, n1 - n0, ref[0])  # a capture or two per solve
    else:
        assert (c1, n1) == (c0, n0)
    assert np.all(ref[0] > 10)  # some systems converge, the one nearest an eigenvalue runs to the limit
    A.delete()


@pytest.mark.parametrize("maxit", [1, 5, 13])
@pytest.mark.parametrize("glen", [1, 4])
def test_minres_graphs_iteration_limit(be64, graphs, maxit, glen):
    """iterations stop exactly at the limit with 4 queued per step"""
    A = be64.laplace3d(10, 11, 12)
    rng = np.random.default_rng(maxit)
    shifts = np.linspace(-1.0, 1.0, 5)
    Bh = rng.uniform(-1, 1, (A.nrows, 5))
    graphs(False, 1)
    ref = minres(be64, A, shifts, Bh, 1e-14, maxit, False)
    graphs(True, 4, glen)
    got = minres(be64, A, shifts, Bh, 1e-14, maxit, False)
    assert np.all(ref[0] == maxit) and np.array_equal(got[0], ref[0])
    assert np.array_equal(got[1], ref[1]) and np.array_equal(got[3], ref[3])
    A.delete()


@pytest.mark.parametrize("fp32", [False, True])
def test_jada_graphs_bitexact(be64, graphs, fp32):
    A = be64.laplace3d(9, 11, 13)
    runs = []
    for enable, ipc in ((False, 1), (True, 4)):
        graphs(enable, ipc, 2)
        fl = ctypes.c_int(0)
        be64.L.phist_amd_kiss_reset(ctypes.byref(fl))  # the same random start vector for both runs
        assert fl.value == 0
        c0, n0 = be64.minres_graph_stats()
        ev, res, nconv, niter, Q, R = be64.subspacejada(A, 6, "SR", 1e-8, block_size=2, min_bas=14, max_bas=28,
                                                        max_iters=400, inner_max_iters=20, inner_fp32=fp32)
        c1, n1 = be64.minres_graph_stats()
        runs.append((ev.copy(), res.copy(), nconv, niter, Q.get(), n1 - n0))
        Q.delete()
        R.delete()
    (ev0, res0, nc0, it0, Q0, l0), (ev1, res1, nc1, it1, Q1, l1) = runs
    assert nc0 >= 6 and (nc1, it1) == (nc0, it0)
    assert np.array_equal(ev0, ev1) and np.array_equal(res0, res1) and np.array_equal(Q0, Q1)
    assert l0 == 0 and l1 > 0
    A.delete()


def test_serve_graphs_bitexact(be64, graphs):
    A1 = be64.laplace3d(12, 13, 14)
    A2 = be64.spinchain_sz(12, 6)
    rng = np.random.default_rng(5)
    work = []
    for A, nrhs, fp32 in [(A1, 3, False), (A1, 5, False), (A2, 4, False), (A1, 2, True), (A2, 7, True)]:
        shifts = rng.uniform(0.5, 2.0, nrhs) + (0.0 if A is A1 else 12.0)
        work.append((A, shifts, rng.uniform(-1, 1, (A.nrows, nrhs)), fp32))
    out = []
    for enable in (False, True):
        graphs(enable, 1 if not enable else 4, 4)
        srv = be64.server(budget_bytes=1 << 30, slots=4)  # fewer slots than systems: slots are refilled
        reqs = []
        for A, shifts, Bh, fp32 in work:
            B, X = be64.mvec(A.domain_map, Bh.shape[1], Bh), be64.mvec(A.domain_map, Bh.shape[1])
            reqs.append((srv.submit_solve(A, shifts, B, X, 1e-5 if fp32 else 1e-10, 2000, fp32=fp32), B, X))
        while srv.step() > 0:
            pass
        res = []
        for rid, B, X in reqs:
            st, _, it = srv.status(rid)
            assert st == pa.Server.DONE
            res.append((it, X.get()))
            srv.release(rid)
            B.delete()
            X.delete()
        srv.delete()
        out.append(res)
    for (it0, X0), (it1, X1) in zip(*out):
        assert it0 == it1 and np.array_equal(X0, X1)
    A1.delete()
    A2.delete()
