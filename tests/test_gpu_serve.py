"""GPU: the continuous-batching service loop (phist_amd_serve.h).

Requests of different sizes and precisions share MINRES batches with fewer
slots than systems (so slots are refilled as systems converge); every
solution is checked by its true residual on the host, eigenvalue requests by
their eigenvalues; memory admission and request states are checked too."""
import itertools

import numpy as np
import pytest

import oracle_lib as ol
import phist_amd as pa

pytestmark = pytest.mark.gpu


def true_relres(A, shifts, B, X):
    ro, ci, va = A.crs()
    crs = ol.Crs(A.nrows, A.ncols, ro, ci, va.astype(float))
    AX = ol.spmv(crs, 1.0, X, 0.0, np.zeros_like(X))
    return np.linalg.norm(B - AX - shifts * X, axis=0) / np.linalg.norm(B, axis=0)


def test_serve_mixed_requests(be64):
    A1 = be64.laplace3d(12, 13, 14)
    A2 = be64.spinchain_sz(12, 6)
    srv = be64.server(budget_bytes=1 << 30, slots=6)
    rng = np.random.default_rng(3)
    reqs = []
    for i, (A, nrhs, fp32) in enumerate([(A1, 3, False), (A1, 5, False), (A2, 4, False), (A1, 2, True),
                                          (A2, 7, True), (A1, 1, False)]):
        n = A.nrows
        shifts = rng.uniform(0.5, 2.0, nrhs) + (0.0 if A is A1 else 12.0)  # SPD shifted systems
        Bh = rng.uniform(-1, 1, (n, nrhs))
        B, X = be64.mvec(A.domain_map, nrhs, Bh), be64.mvec(A.domain_map, nrhs)
        tol = 1e-5 if fp32 else 1e-10
        rid = srv.submit_solve(A, shifts, B, X, tol, 2000, fp32=fp32)
        reqs.append((rid, A, shifts, Bh, B, X, tol))
    jid, Q, R, ev, res = srv.submit_jada(A2, 4, "SR", 1e-8, block_size=2, min_bas=12, max_bas=24)
    steps = 0
    while srv.step() > 0:
        steps += 1
        assert steps < 20000
    for rid, A, shifts, Bh, B, X, tol in reqs:
        st, lat, it = srv.status(rid)
        assert st == pa.Server.DONE and lat > 0 and it > 0
        rr = true_relres(A, shifts, Bh, X.get())
        assert np.all(rr <= (20 * tol if tol > 1e-6 else 1e-8)), (rid, rr)
        srv.release(rid)
        for o in (B, X):
            o.delete()
    st, lat, it = srv.status(jid)
    assert st == pa.Server.DONE
    w = np.linalg.eigvalsh(ol.rows_to_dense(ol.spin_sz_rows(12, 6, 0), A2.nrows))
    got = np.sort(ev[:4])
    assert np.min(np.abs(got[:, None] - w[None, :]), axis=1).max() <= 1e-7
    assert abs(got[0] - w[0]) <= 1e-7
    srv.release(jid)
    for o in (Q, R, A1, A2):
        o.delete()
    srv.delete()


def test_serve_budget(be64):
    A = be64.laplace3d(10, 10, 10)
    # room for one FP64 batch of 2 slots (6*2*1000*8 bytes) but not for the JD work space
    srv = be64.server(budget_bytes=6 * 2 * 1000 * 8 + 100, slots=2)
    jid, Q, R, ev, res = srv.submit_jada(A, 2, "SR", 1e-8, block_size=1, min_bas=10, max_bas=20)
    B, X = be64.mvec(A.domain_map, 3, np.ones((1000, 3))), be64.mvec(A.domain_map, 3)
    rid = srv.submit_solve(A, [1.0, 2.0, 3.0], B, X, 1e-10, 500)
    while srv.step() > 0:
        pass
    st, _, _ = srv.status(jid)
    assert st == -44  # the eigenvalue request can never fit the budget
    st, _, it = srv.status(rid)
    assert st == pa.Server.DONE
    rr = true_relres(A, np.array([1.0, 2.0, 3.0]), np.ones((1000, 3)), X.get())
    assert np.all(rr <= 1e-8)
    with pytest.raises(pa.PhistError):
        srv.release(12345)
    for o in (B, X, Q, R, A):
        o.delete()
    srv.delete()


def test_serve_result_same_alone_and_in_the_mix(be64):
    """each request's result served alone equals its result served among others, within the solver
    tolerance: solutions to 100*tol relative (the shifted systems have condition number < 30), equal
    iteration counts to +-2, eigenvalues to 1e-9"""
    A1 = be64.laplace3d(11, 12, 13)
    A2 = be64.spinchain_sz(12, 6)
    rng = np.random.default_rng(9)
    specs = []
    for A, nrhs, fp32 in [(A1, 3, False), (A2, 5, False), (A1, 4, True), (A2, 2, False), (A1, 6, False)]:
        n = A.nrows
        shifts = rng.uniform(0.5, 2.0, nrhs) + (0.0 if A is A1 else 12.0)
        specs.append((A, shifts, rng.uniform(-1, 1, (n, nrhs)), 1e-5 if fp32 else 1e-10, fp32))

    def serve(which, with_jada):
        srv = be64.server(budget_bytes=1 << 30, slots=5)
        out, objs = {}, []
        for i in which:
            A, shifts, Bh, tol, fp32 = specs[i]
            B, X = be64.mvec(A.domain_map, len(shifts), Bh), be64.mvec(A.domain_map, len(shifts))
            objs += [B, X]
            out[i] = (srv.submit_solve(A, shifts, B, X, tol, 3000, fp32=fp32), X)
        jada = srv.submit_jada(A2, 4, "SR", 1e-8, block_size=2, min_bas=12, max_bas=24) if with_jada else None
        n = 0
        while srv.step() > 0:
            n += 1
            assert n < 40000
        res = {}
        for i, (rid, X) in out.items():
            st, _, it = srv.status(rid)
            assert st == pa.Server.DONE
            res[i] = (X.get(), it)
        ev = None
        if jada is not None:
            jid, Q, R, ev_a, _ = jada
            assert srv.status(jid)[0] == pa.Server.DONE
            ev = np.sort(ev_a[:4].copy())
            objs += [Q, R]
        for o in objs:
            o.delete()
        srv.delete()
        return res, ev

    mixed, ev_mixed = serve(range(len(specs)), True)
    for i in range(len(specs)):
        alone, _ = serve([i], False)
        xa, ita = alone[i]
        xm, itm = mixed[i]
        tol = specs[i][3]
        rel = np.linalg.norm(xm - xa, axis=0) / np.linalg.norm(xa, axis=0)
        assert np.all(rel <= 100 * tol), (i, rel)
        assert abs(ita - itm) <= 2, (i, ita, itm)
    _, ev_alone = serve([], True)
    assert np.all(np.abs(ev_mixed - ev_alone) <= 1e-9 * np.abs(ev_alone).max())
    for o in (A1, A2):
        o.delete()


def test_serve_one_batch_budget_two_matrices(be64):
    """a budget that holds the eigenvalue request's work space but not that and a batch: solves
    alternating between two matrices, and the eigenvalue request queued ahead of solves on the matrices
    of idle batches; every request must finish (idle batches kept for queued requests behind a head
    that waits for memory would block the loop for good)"""
    A1 = be64.laplace3d(8, 9, 10)
    A2 = be64.laplace3d(9, 9, 9)
    n1, n2 = A1.nrows, A2.nrows
    one_batch = 6 * 2 * max(n1, n2) * 8
    jd_bytes = 8 * n1 * (2 * 20 + 2 + 1 + 2 + 8 + 2 + 4)
    srv = be64.server(budget_bytes=max(one_batch, jd_bytes) + 64, slots=2)
    rng = np.random.default_rng(4)
    reqs = []
    for i, A in enumerate([A1, A2, A1, A2]):
        sh = rng.uniform(0.5, 2.0, 3)
        Bh = rng.uniform(-1, 1, (A.nrows, 3))
        B, X = be64.mvec(A.domain_map, 3, Bh), be64.mvec(A.domain_map, 3)
        reqs.append((srv.submit_solve(A, sh, B, X, 1e-9, 2000), A, sh, Bh, B, X))
        if i == 1:
            # queue: [JD on A1, solve on A1 ...] while the A1 batch of request 0 is idle
            jid, Q, R, ev, res = srv.submit_jada(A1, 2, "SR", 1e-8, block_size=1, min_bas=10, max_bas=20)
    steps = 0
    while srv.step() > 0:
        steps += 1
        assert steps < 50000, "service loop does not drain"
    for rid, A, sh, Bh, B, X in reqs:
        assert srv.status(rid)[0] == pa.Server.DONE
        assert np.all(true_relres(A, sh, Bh, X.get()) <= 1e-8)
        srv.release(rid)
        B.delete()
        X.delete()
    assert srv.status(jid)[0] == pa.Server.DONE
    lam = sorted(6 - 2 * np.cos(np.pi * i / 9) - 2 * np.cos(np.pi * j / 10) - 2 * np.cos(np.pi * k / 11)
                 for i in range(1, 9) for j in range(1, 10) for k in range(1, 11))[:2]
    assert np.allclose(np.sort(ev[:2]), lam, rtol=0, atol=1e-7)
    for o in (Q, R, A1, A2):
        o.delete()
    srv.delete()


def test_serve_rejects_what_can_never_run(be64):
    """a solve whose batch can never fit fails with -44 without blocking the requests behind it;
    eigenvalue options that subspacejada does not support fail at submission with -99"""
    A = be64.laplace3d(10, 10, 10)
    small = be64.laplace3d(4, 4, 4)
    srv = be64.server(budget_bytes=6 * 2 * 64 * 8 + 16, slots=2)  # a batch over `small` fits, over A not
    B, X = be64.mvec(A.domain_map, 2, np.ones((1000, 2))), be64.mvec(A.domain_map, 2)
    big = srv.submit_solve(A, [1.0, 2.0], B, X, 1e-10, 500)
    Bs, Xs = be64.mvec(small.domain_map, 2, np.ones((64, 2))), be64.mvec(small.domain_map, 2)
    ok = srv.submit_solve(small, [1.0, 2.0], Bs, Xs, 1e-10, 500)
    with pytest.raises(pa.PhistError) as e:
        srv.submit_jada(small, 2, 5, 1e-8, block_size=1)  # phist_TARGET
    assert e.value.code == -99
    with pytest.raises(pa.PhistError) as e:
        srv.submit_jada(small, 2, "SR", 1e-8, block_size=1, max_bas=5)  # maxBas < minBas
    assert e.value.code == -55
    n = 0
    while srv.step() > 0:
        n += 1
        assert n < 5000
    assert srv.status(big)[0] == -44
    assert srv.status(ok)[0] == pa.Server.DONE
    assert np.all(true_relres(small, np.array([1.0, 2.0]), np.ones((64, 2)), Xs.get()) <= 1e-8)
    for o in (B, X, Bs, Xs, A, small):
        o.delete()
    srv.delete()


def test_serve_cancel(be64):
    """cancelling a queued and a running solve and a running eigenvalue request: the other solves
    finish with correct results, the queued one taking the cancelled slot; a finished request is
    unchanged"""
    A = be64.laplace3d(12, 12, 12)
    n = A.nrows
    srv = be64.server(budget_bytes=1 << 30, slots=3)
    rng = np.random.default_rng(8)
    Bh = rng.uniform(-1, 1, (n, 5))
    sh = np.array([0.5, 1.0, 1.5, 2.0, 2.5])
    objs, ids = [], []
    for j in range(5):  # one system each; 3 slots, so requests 3 and 4 wait in the queue
        B, X = be64.mvec(A.domain_map, 1, Bh[:, j:j + 1]), be64.mvec(A.domain_map, 1)
        objs += [B, X]
        ids.append((srv.submit_solve(A, sh[j:j + 1], B, X, 1e-10, 3000), X))
    jid, Q, R, ev, res = srv.submit_jada(A, 2, "SR", 1e-8, block_size=1, min_bas=10, max_bas=20)
    srv.step()
    srv.step()
    assert srv.status(ids[1][0])[0] == pa.Server.RUNNING
    assert srv.status(ids[3][0])[0] == pa.Server.QUEUED
    assert srv.status(ids[4][0])[0] == pa.Server.QUEUED
    srv.cancel(ids[1][0])  # running
    srv.cancel(ids[3][0])  # queued
    srv.cancel(jid)  # running eigenvalue request
    for r in (ids[1][0], ids[3][0], jid):
        assert srv.status(r)[0] == pa.Server.CANCELLED
    srv.step()
    assert srv.status(ids[4][0])[0] == pa.Server.RUNNING  # took the freed slot
    k = 0
    while srv.step() > 0:
        k += 1
        assert k < 20000
    for j in (0, 2, 4):
        assert srv.status(ids[j][0])[0] == pa.Server.DONE
        assert np.all(true_relres(A, sh[j:j + 1], Bh[:, j:j + 1], ids[j][1].get()) <= 1e-8)
    srv.cancel(ids[0][0])  # finished: unchanged
    assert srv.status(ids[0][0])[0] == pa.Server.DONE
    for r in [i for i, _ in ids] + [jid]:
        srv.release(r)
    for o in objs + [Q, R, A]:
        o.delete()
    srv.delete()
