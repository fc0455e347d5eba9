"""GPU: batched shifted MINRES on the honeycomb tight-binding sheet (the batched-MINRES workload at test
size): nearest-neighbour hopping -1, on-site disorder U(-1/2, 1/2) from a seed, shifts uniform in
[-2.5, 2.5] (inside the spectrum, so the systems are indefinite), random right-hand sides of unit norm,
and fewer slots than systems, so converged columns are replaced in place by the next system
(phist_amd_DblockedMINRES_shifted_refill; the slot loop of phist_jadaCorrectionSolver_def.hpp:428-466).

Every solution is checked by its true residual ||b - (A + s I) x|| / ||b||, computed on the host with
the oracle's SpMV of the matrix's CRS; FP64 to the tolerance 1e-6 (the MINRES recurrence estimate and
the true residual differ by O(eps * cond^2), far below it at these sizes).  FP32 iterations alone reach
a true residual limited by FP32 rounding times the conditioning of these interior shifts; the FP32
variant refines in FP64 (true residuals between FP32 solves), so it is held to the same 1e-6 and its
reported relres must be the true residual.  The number of iterations a system needs does not depend on the slot
it runs in or on its neighbours beyond rounding: blocks of other widths reduce the dots in another
order, and over a thousand iterations on these nearly singular systems that moves the count by up to
1% (plus 2, the lag of the host's view of the states)."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def unit_rhs(n, k, seed):
    B = np.random.default_rng(seed).standard_normal((n, k))
    return B / np.linalg.norm(B, axis=0)


def true_relres(A, shifts, B, X):
    ro, ci, va = A.crs()
    crs = ol.Crs(A.nrows, A.ncols, ro, ci, va.astype(float))
    R = ol.spmv(crs, -1.0, X, 1.0, B, shifts=np.asarray(shifts, float))
    return np.linalg.norm(R, axis=0) / np.linalg.norm(B, axis=0)


@pytest.mark.parametrize("W,L,nsys,slots", [(32, 32, 12, 4), (32, 48, 16, 16), (64, 64, 24, 8)])
def test_minres_honeycomb_refill_fp64(be64, W, L, nsys, slots):
    A = be64.honeycomb(W, L, 0, 0, 0, 1.0, 5)
    n = A.nrows
    shifts = np.random.default_rng(W + L).uniform(-2.5, 2.5, nsys)
    Bh = unit_rhs(n, nsys, nsys)
    B, X = be64.mvec(A.domain_map, nsys, Bh), be64.mvec(A.domain_map, nsys)
    it, rr, unconv = be64.minres_refill(A, shifts, B, X, 1e-6, 60000, slots)
    print(f"honeycomb {W}x{L} nsys {nsys} slots {slots}: iterations {sorted(it.tolist())}", flush=True)
    assert unconv == 0
    assert np.all(rr <= 1e-6)
    tr = true_relres(A, shifts, Bh, X.get())
    assert np.all(tr <= 1.5e-6), tr
    if slots < nsys:
        # the same systems in one block (or the first 64): same iteration counts
        k = min(nsys, 64)
        B2, X2 = be64.mvec(A.domain_map, k, Bh[:, :k]), be64.mvec(A.domain_map, k)
        it2, _, unconv2 = be64.minres_shifted(A, shifts[:k], B2, X2, 1e-6, 60000)
        assert unconv2 == 0
        assert np.all(np.abs(it2 - it[:k]) <= 2 + it[:k] // 100), (it2, it[:k])
        for o in (B2, X2):
            o.delete()
    for o in (B, X, A):
        o.delete()


@pytest.mark.parametrize("refill", [True, False])
def test_minres_honeycomb_fp32_refined(be64, refill):
    A = be64.honeycomb(32, 32, 0, 0, 0, 1.0, 5)
    n = A.nrows
    nsys = 12
    shifts = np.random.default_rng(64).uniform(-2.5, 2.5, nsys)
    Bh = unit_rhs(n, nsys, 3)
    B, X = be64.mvec(A.domain_map, nsys, Bh), be64.mvec(A.domain_map, nsys)
    if refill:
        it, rr, unconv = be64.minres_refill(A, shifts, B, X, 1e-6, 60000, 4, fp32=True)
    else:
        it, rr, unconv = be64.minres_shifted(A, shifts, B, X, 1e-6, 60000, fp32=True)
    tr = true_relres(A, shifts, Bh, X.get())
    print(f"fp32 refined: iterations {it.tolist()}, true relres max {tr.max():.2e}", flush=True)
    assert unconv == 0
    assert np.all(tr <= 1.5e-6), tr
    assert np.allclose(rr, tr, rtol=1e-3, atol=1e-12), (rr, tr)  # relres is the true residual
    for o in (B, X, A):
        o.delete()


def test_minres_refill_edge_cases(be64):
    """a zero right-hand side finishes at once with x = 0, maxIter bounds a system that cannot converge
    (reported unconverged, iflag > 0), one slot runs the systems one after another"""
    A = be64.honeycomb(16, 16, 0, 0, 0, 1.0, 9)
    n = A.nrows
    Bh = unit_rhs(n, 5, 1)
    Bh[:, 2] = 0.0
    shifts = np.array([0.3, -1.1, 0.7, 2.2, -0.4])
    B, X = be64.mvec(A.domain_map, 5, Bh), be64.mvec(A.domain_map, 5, np.full((n, 5), 7.0))
    it, rr, unconv = be64.minres_refill(A, shifts, B, X, 1e-6, 20000, 1)
    assert unconv == 0
    Xh = X.get()
    assert np.all(Xh[:, 2] == 0.0)
    tr = true_relres(A, shifts[[0, 1, 3, 4]], Bh[:, [0, 1, 3, 4]], Xh[:, [0, 1, 3, 4]])
    assert np.all(tr <= 1.5e-6)
    it2, rr2, unconv2 = be64.minres_refill(A, shifts, B, X, 1e-14, 5, 2)
    assert unconv2 == 4 and np.all(it2[[0, 1, 3, 4]] == 5)
    for o in (B, X, A):
        o.delete()
