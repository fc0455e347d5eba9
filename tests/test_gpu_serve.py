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
